"""Split-precision potential update on the GPU (th_charging_set_green).

The Green's matrix is split into two 16-bit halves (BF16, or FP16 of the
matrix scaled into FP16 range), products and sums run in FP32, and the
external potential is added in FP64.  At the AR 8 workload with a charged
surface the errors against the FP64 update are stated here: BF16 hi+lo within
5 mV (potential) and 1e-3 of max |E| (field), FP16 hi+lo within 1 mV and
2e-4 (the emulation below gives 1.5 mV and 0.24 mV for the potential; the
field differences neighbouring potentials and so carries their error / dx).  The device result also
matches a NumPy emulation of the same split arithmetic to FP32 summation
order, and switching back to FP64 restores the FP64 potential bit for bit."""
import numpy as np
import pytest

from trench01_amd import charging, configs

pytestmark = pytest.mark.gpu


def bf16(x):
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7fff + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def emulate(G, q, mode):
    """The device's split arithmetic (BF16: hi + lo; FP16: scaled hi + lo * 2^-11)."""
    if mode == 1:
        hi = bf16(G.astype(np.float32))
        lo = bf16((G - hi.astype(np.float64)).astype(np.float32))
        return (hi + lo) @ q.astype(np.float32), 1.0
    S = 2.0 ** (14 - int(np.ceil(np.log2(np.abs(G).max()))))
    hi = (G * S).astype(np.float32).astype(np.float16)
    lo = ((G * S - hi.astype(np.float64).astype(np.float64)) / 2.0 ** -11).astype(np.float32).astype(np.float16)
    g = hi.astype(np.float32) + lo.astype(np.float32) * np.float32(2.0 ** -11)
    return g @ q.astype(np.float32), S


@pytest.fixture(scope="module")
def charged():
    print("[gpu case] split Green's AR8", flush=True)
    p, _ = configs.product_params("AR8")
    s = charging.ChargingSetup(p)
    dev = charging.DeviceCharging(s)
    rng = np.random.default_rng(0)
    rho = rng.normal(0, 3e6, dev.m) * (1 - s.cond_s)
    rho[s.c_idx] = -4e8
    dev.set_rho(rho)
    q = np.zeros(s.green.shape[1])
    has = s.col_of >= 0
    np.add.at(q, s.col_of[has], rho[has] * p.dx ** 2)
    _, phi64, _ = dev.read(counts=False)
    e64 = dev.tracer.read_field()
    yield s, dev, q, phi64, e64
    dev.close()


@pytest.mark.parametrize("mode, tol_v, tol_e", [(1, 5e-3, 1e-3), (2, 1e-3, 2e-4)])
def test_split_green_potential_error(charged, mode, tol_v, tol_e):
    s, dev, q, phi64, e64 = charged
    dev.set_green_precision(mode)
    try:
        _, phi, _ = dev.read(counts=False)
        rows = s.rows
        ext = s.fields.external.ravel()[rows]
        err = np.abs(phi[rows] - phi64[rows])
        scale = np.abs(phi64[rows]).max()
        print(f"mode {mode}: max |dphi| {err.max():.3g} V of {scale:.1f} V", flush=True)
        assert err.max() < tol_v
        ref, S = emulate(s.green, q, mode)
        np.testing.assert_allclose(phi[rows], ref.astype(np.float64) / S + ext, rtol=0, atol=2e-6 * scale)
        ex, ey = dev.tracer.read_field()
        emax = max(np.abs(e64[0]).max(), np.abs(e64[1]).max())
        ferr = max(np.abs(ex - e64[0]).max(), np.abs(ey - e64[1]).max()) / emax
        print(f"mode {mode}: max field error {ferr:.3g} of max |E|", flush=True)
        assert ferr < tol_e
    finally:
        dev.set_green_precision(0)
    _, back, _ = dev.read(counts=False)
    np.testing.assert_array_equal(back, phi64)


def test_batch_runs_in_split_precision(charged):
    """A whole device batch with the BF16 split potential: deposits, leak,
    conductor and the split potential run, and the result stays within the
    split error of the FP64 update of the same tallies."""
    s, dev, q, phi64, e64 = charged
    p = s.p
    n = 2 * 20000
    dev.reserve(n)
    rho0, _, _ = dev.read(phi=False, counts=False)
    out = {}
    for mode in (0, 1):
        dev.set_rho(rho0)
        dev.set_green_precision(mode)
        dev.inject_device(0, 20000, p.seed, 7)
        d = dev.step(0, n)
        assert d.tally.off_surface == 0
        out[mode] = dev.read(counts=True)
    dev.set_green_precision(0)
    (r0, f0, c0), (r1, f1, c1) = out[0], out[1]
    # same field before the push (up to the split error) -> tallies differ by a handful of particles at most
    assert np.abs(c0.astype(np.int64) - c1).sum() <= 0.01 * n
    assert np.abs(f1[s.rows] - f0[s.rows]).max() < 0.05


@pytest.mark.parametrize("mode, scalar, tol_v, tol_e", [(3, 1, 5e-3, 1e-3), (4, 2, 1e-3, 2e-4)])
def test_split_green_on_matrix_cores(charged, mode, scalar, tol_v, tol_e):
    """The split update on MFMA (modes 3 / 4: the charges split into two
    halves of the matrix's 16-bit type too) within the same bounds against
    FP64 as the scalar split, and within 2e-5 of the potential scale of it."""
    s, dev, _, _, _ = charged
    rows = s.rows
    dev.set_green_precision(0)          # the FP64 reference of the fixture's current charge
    _, phi64, _ = dev.read(counts=False)
    e64 = dev.tracer.read_field()
    scale = np.abs(phi64[rows]).max()
    try:
        dev.set_green_precision(scalar)
        _, phis, _ = dev.read(counts=False)
        dev.set_green_precision(mode)
        _, phi, _ = dev.read(counts=False)
        err = np.abs(phi[rows] - phi64[rows])
        print(f"mode {mode}: max |dphi| {err.max():.3g} V of {scale:.1f} V; "
              f"vs scalar split {np.abs(phi[rows] - phis[rows]).max():.3g} V", flush=True)
        assert err.max() < tol_v
        assert np.abs(phi[rows] - phis[rows]).max() < 2e-5 * scale
        ex, ey = dev.tracer.read_field()
        emax = max(np.abs(e64[0]).max(), np.abs(e64[1]).max())
        assert max(np.abs(ex - e64[0]).max(), np.abs(ey - e64[1]).max()) / emax < tol_e
    finally:
        dev.set_green_precision(0)
    _, back, _ = dev.read(counts=False)
    np.testing.assert_array_equal(back, phi64)


@pytest.mark.parametrize("mode, tol", [(3, 5e-3), (4, 1e-3)])
def test_potentials_of_many_charge_vectors(charged, mode, tol):
    """th_charging_potentials (DeviceCharging.potentials): G Q for 16 charge
    vectors in one pass over G, against the FP64 product on the host; one
    vector at a time gives each column within rounding; 17 vectors or a
    wrong column count are refused."""
    s, dev, q, phi64, e64 = charged
    rng = np.random.default_rng(5)
    R = 16
    Q = q[None, :] * rng.uniform(0.5, 1.5, (R, q.size))
    ref = Q @ s.green.T                                       # [R][n_rows]
    dev.set_green_precision(mode)
    try:
        out, ms = dev.potentials(Q)
        err = np.abs(out - ref).max()
        print(f"mode {mode}: 16 right-hand sides, max error {err:.3g} V, kernel {ms * 1e3:.1f} us", flush=True)
        assert err < tol
        one, ms = dev.potentials(Q[3])
        print(f"mode {mode}: 1 right-hand side, kernel {ms * 1e3:.1f} us", flush=True)
        np.testing.assert_allclose(one[0], out[3], rtol=0, atol=1e-6 * np.abs(ref).max())
        for bad in (np.vstack([Q, Q[:1]]), Q[:, :-1]):
            with pytest.raises(ValueError):
                dev.potentials(bad)
    finally:
        dev.set_green_precision(0)
