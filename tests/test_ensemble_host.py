"""Host side of the ensemble push (CPU): the batching workload's
distributions and the per-instance field set-up."""
import numpy as np

from trench01_amd import ensemble as E, model


def test_workload_ranges_and_valid_geometries():
    specs = E.batching_workload(256, seed=3)
    w = np.array([s.p.trench_w for s in specs])
    ar = np.array([s.p.trench_d / s.p.trench_w for s in specs])
    mask = np.array([s.p.mask_t for s in specs])
    tapers = np.array([(s.p.trench_taper_deg, s.p.mask_taper_deg) for s in specs])
    n = np.array([2 * s.p.n_per_batch for s in specs])
    assert w.min() >= 10 and w.max() <= 40
    assert ar.min() >= 2 - 0.05 and ar.max() <= 20 + 0.05
    assert mask.min() >= 0 and mask.max() <= 40
    assert tapers.min() >= -3 and tapers.max() <= 5
    assert 1e4 - 2 <= n.min() and n.max() <= 1e6
    # log-uniform: about half the instances below the geometric mean 1e5
    assert 0.35 < np.mean(n < 1e5) < 0.65
    kinds = {s.p.ion_distribution for s in specs}
    assert kinds == {"mono", "arcsine"}
    assert {s.p.mask_type for s in specs} == {"conductor", "dielectric"}
    for s in specs[:40]:
        geo = model.Geometry(s.p)          # raises on an invalid geometry
        b0, b1 = geo.opening_span(s.p.floor_t)
        assert b1 - b0 >= 1
    peaks = [(s.p.ion_energy_lo_eV, s.p.ion_energy_hi_eV) for s in specs if s.p.ion_distribution == "arcsine"]
    assert all(50 <= lo <= hi <= 300 for lo, hi in peaks)


def test_workload_is_reproducible():
    a = E.batching_workload(8, seed=11)
    b = E.batching_workload(8, seed=11)
    assert [s.p for s in a] == [s.p for s in b]
    assert [s.p for s in E.batching_workload(8, seed=12)] != [s.p for s in a]


def test_instance_field_is_the_poisson_solution_of_the_stated_charge():
    """The field of an instance is the stencil of phi = LU^-1(-rho/eps0) with
    charge only on the floor and wall surface cells: the discrete Laplacian of
    phi reproduces the charge there and vanishes in vacuum."""
    spec = E.batching_workload(3, seed=5, particles=(1e4, 2e4))[0]
    vac, ex, ey, phi, wf = E.instance_field(spec)
    p = spec.p
    geo = model.Geometry(p)
    f = model.Fields(p, geo)
    res = (f.A @ phi.ravel() - f.boundary_rhs.ravel()).reshape(phi.shape) * -model.EPS0
    cond = f.cond
    free = ~cond
    q = res[free]
    # charge appears only on surface cells, at the three stated densities
    nonzero = np.abs(q) > 1e-6 * max(abs(c) for c in spec.charge)
    surf = np.zeros(geo.solid.size, bool)
    surf[geo.surface] = True
    assert np.all(surf.reshape(phi.shape)[free][nonzero])
    vals = np.unique(np.round(q[nonzero] / spec.charge[0], 9))
    assert set(vals) <= {1.0, round(spec.charge[1] / spec.charge[0], 9), round(spec.charge[2] / spec.charge[0], 9)}
    ex2, ey2 = model.compute_field(phi, vac, p, f.volts["top"])
    np.testing.assert_array_equal(ex, ex2)
    np.testing.assert_array_equal(ey, ey2)
