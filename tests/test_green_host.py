"""The host side of the device Green's solve (charging.lu_factors,
charging.device_solve's permutations) and th_lu_solve's argument checks, which
run before any device work."""
import numpy as np
import pytest
import scipy.sparse as sps
from scipy.sparse.linalg import spsolve_triangular

from tests import cases
from trench01_amd import _lib, charging, model


def _solver(p):
    names = set(model.Params.__dataclass_fields__)
    p = model.Params(**{k: getattr(p, k) for k in names if hasattr(p, k)})
    return model.Fields(p, model.Geometry(p)).solver


SOLVERS = {
    "floating conductor": lambda: _solver(cases.ar5_params()),
    "dielectric mask": lambda: _solver(cases.taper_params(mask_type="dielectric")),
    "driven conductor": lambda: _solver(cases.taper_params(field_mode="fixed", mask_voltage=-20.0)),
}


@pytest.mark.parametrize("name", list(SOLVERS))
def test_factors_rebuild_the_lu_and_its_solve(name):
    """The CSR arrays handed to th_lu_solve are SuperLU's L and U exactly; the
    permutations applied around the triangular solves reproduce lu.solve
    (here with scipy's triangular solves in place of the device's)."""
    s = SOLVERS[name]()
    lu = s.lu
    n = lu.shape[0]
    f = charging.lu_factors(lu)
    L = sps.csr_matrix((f["l_val"], f["l_idx"], f["l_ptr"]), shape=(n, n)) + sps.identity(n)
    U = sps.csr_matrix((f["u_val"], f["u_idx"], f["u_ptr"]), shape=(n, n)) + sps.diags(f["u_diag"])
    assert (L != lu.L).nnz == 0 and (U != lu.U).nnz == 0
    rows = np.repeat(np.arange(n), np.diff(f["l_ptr"]))
    assert np.all(f["l_idx"] < rows)
    rows = np.repeat(np.arange(n), np.diff(f["u_ptr"]))
    assert np.all(f["u_idx"] > rows)
    rng = np.random.default_rng(1)
    B = rng.normal(size=(s.A.shape[0], 3))
    r = np.asarray(s.reduce(B))
    z = np.empty_like(r)
    z[lu.perm_r] = r
    w = spsolve_triangular(U, spsolve_triangular(L, z, lower=True, unit_diagonal=True), lower=False)
    x = s.expand(w[lu.perm_c])
    ref = s.solve(B)
    np.testing.assert_allclose(x, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def _call(n, lp, li, lv, up, ui, uv, ud, k=1):
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    arrs = [i32(lp), i32(li), f64(lv), i32(up), i32(ui), f64(uv), f64(ud)]
    b = np.ones((n, k))
    x = np.empty((n, k))
    types = [_lib.i32p, _lib.i32p, _lib.dp, _lib.i32p, _lib.i32p, _lib.dp, _lib.dp]
    rc = _lib.load().th_lu_solve(0, n, *(_lib.ptr(a, t) for a, t in zip(arrs, types)), k, _lib.ptr(b, _lib.dp),
                                 _lib.ptr(x, _lib.dp), None)
    return rc, _lib.load().th_last_error().decode()


# a valid 3x3 system: L = [[1],[2,1],[0,3,1]], U = [[4,1,0],[5,2],[6]]
GOOD = dict(n=3, lp=[0, 0, 1, 2], li=[0, 1], lv=[2.0, 3.0], up=[0, 1, 2, 2], ui=[1, 2], uv=[1.0, 2.0],
            ud=[4.0, 5.0, 6.0])


@pytest.mark.parametrize("change,message", [
    (dict(li=[1, 1]), "strictly lower"),          # row 1 may only reach column 0
    (dict(li=[0, 5]), "strictly lower"),          # out of range
    (dict(lp=[1, 0, 1, 2]), "strictly lower"),    # does not start at 0
    (dict(lp=[0, 1, 0, 2]), "strictly lower"),    # decreasing
    (dict(ui=[0, 2]), "strictly upper"),          # row 0 may not reach column 0
    (dict(ui=[1, 3]), "strictly upper"),          # out of range
    (dict(uv=[1.0, np.nan]), "strictly upper"),   # non-finite entry
    (dict(ud=[4.0, 0.0, 6.0]), "zero pivot"),
    (dict(ud=[4.0, np.inf, 6.0]), "zero pivot"),
])
def test_lu_solve_rejects_malformed_factors(change, message):
    rc, msg = _call(**{**GOOD, **change})
    assert rc != 0 and message in msg


def test_lu_solve_accepts_no_right_hand_sides():
    rc, _ = _call(**GOOD, k=0)
    assert rc == 0


@pytest.mark.parametrize("name", list(SOLVERS))
def test_unit_source_solve_gathers_the_solver_columns(name, monkeypatch):
    """unit_source_solve's host side (the right-hand side built permuted and
    padded, potentials read back through the cell-to-unknown map): with
    scipy's triangular solves standing in for th_lu_solve, every cell's
    potential equals solver.solve of the unit-source block, including a
    floating conductor's shared unknown and a driven conductor's fixed cells."""
    def host_lu_solve(lu, z, device):
        n = lu.shape[0]
        f = charging.lu_factors(lu)
        L = sps.csr_matrix((f["l_val"], f["l_idx"], f["l_ptr"]), shape=(n, n)) + sps.identity(n)
        U = sps.csr_matrix((f["u_val"], f["u_idx"], f["u_ptr"]), shape=(n, n)) + sps.diags(f["u_diag"])
        y = spsolve_triangular(L.tocsr(), z, lower=True, unit_diagonal=True)
        return spsolve_triangular(U.tocsr(), y, lower=False), 0.0

    monkeypatch.setattr(charging, "_lu_solve", host_lu_solve)
    s = SOLVERS[name]()
    n_cells = s.A.shape[0]
    cells = np.flatnonzero(s.unknown_of() >= 0)[::301][:6]
    potential, _ = charging.unit_source_solve(s, cells, 2.5)
    B = np.zeros((n_cells, cells.size))
    B[cells, np.arange(cells.size)] = 2.5
    ref = s.solve(B)
    every = np.arange(n_cells)
    np.testing.assert_allclose(potential(every), ref, rtol=0, atol=1e-13 * np.abs(ref).max())
    np.testing.assert_array_equal(potential(every, 1), potential(every)[:, 1])
    if s.free is not None:
        with pytest.raises(ValueError, match="driven"):
            charging.unit_source_solve(s, np.flatnonzero(~s.free)[:1], 1.0)
