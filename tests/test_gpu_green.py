"""The Green's columns solved on the device (th_lu_solve through
charging.device_solve) against the host LU solve of the same factors
(scipy's SuperLU, the reference's solver: trench_charging_2d.py:297-309), for
the three conductor treatments and at the production size."""
import numpy as np
import pytest

from tests import cases
from tests.test_green_host import SOLVERS, _solver
from trench01_amd import charging, configs, model

pytestmark = pytest.mark.gpu

ALL = {**SOLVERS, "AR8 production": lambda: _solver(configs.product_params("AR8")[0])}


def _case(name):
    print(f"[gpu case] {name}", flush=True)


@pytest.mark.parametrize("name", list(ALL))
@pytest.mark.parametrize("k", [1, 65, 200])
def test_device_solve_matches_the_host_lu(name, k):
    """Ragged right-hand-side counts (1, one past a wave, several waves): the
    device's triangular solves agree with SuperLU's to 1e-12 of the largest
    entry."""
    _case(f"lu solve {name} k={k}")
    s = ALL[name]()
    rng = np.random.default_rng(k)
    B = rng.normal(size=(s.A.shape[0], k))
    x, ms = charging.device_solve(s, B)
    ref = s.solve(B)
    assert ms > 0
    np.testing.assert_allclose(x, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def test_setup_green_matrix_equals_the_host_solve():
    """The set-up's Green's matrix, surface response and conductor column from
    the device solve equal the host LU's to 1e-12 of their largest entry."""
    _case("setup green")
    p = cases.ar5_params()
    names = set(model.Params.__dataclass_fields__)
    p = model.Params(**{k: getattr(p, k) for k in names if hasattr(p, k)})
    s = charging.ChargingSetup(p)
    m = s.geo.surface.size
    reps = [np.flatnonzero(s.col_of == c)[0] for c in range(s.green.shape[1])]
    B = np.zeros((s.geo.solid.size, len(reps)))
    B[s.geo.surface[reps], np.arange(len(reps))] = -1.0 / (p.dx ** 2 * model.EPS0)
    G = s.fields.solve(B)
    np.testing.assert_allclose(s.green, G[s.rows], rtol=0, atol=1e-12 * np.abs(G).max())
    P = np.zeros((m, m))
    P[:, s.col_of >= 0] = G[s.geo.surface][:, s.col_of[s.col_of >= 0]]
    np.testing.assert_allclose(s.surface_response, P, rtol=0, atol=1e-12 * np.abs(P).max())
    np.testing.assert_allclose(s.g_cond.ravel(), G[:, s.cond_col], rtol=0, atol=1e-12 * np.abs(G).max())


def test_tiny_system_matches_dense_solve():
    """A 3x3 system written out by hand (L's first row and U's last row hold
    no off-diagonal entry, so each triangle has three levels): the device
    solve equals the dense solve of (L U) x = b for two right-hand sides."""
    _case("lu solve 3x3")
    from tests.test_green_host import GOOD
    from trench01_amd import _lib
    L = np.array([[1.0, 0, 0], [2.0, 1, 0], [0, 3.0, 1]])
    U = np.array([[4.0, 1.0, 0], [0, 5.0, 2.0], [0, 0, 6.0]])
    b = np.array([[1.0, -2.0], [0.5, 3.0], [2.0, 1.0]])
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    arrs = [(i32(GOOD["lp"]), _lib.i32p), (i32(GOOD["li"]), _lib.i32p), (f64(GOOD["lv"]), _lib.dp),
            (i32(GOOD["up"]), _lib.i32p), (i32(GOOD["ui"]), _lib.i32p), (f64(GOOD["uv"]), _lib.dp),
            (f64(GOOD["ud"]), _lib.dp)]
    x = np.empty_like(b)
    _lib.check(_lib.load().th_lu_solve(0, 3, *(_lib.ptr(a, t) for a, t in arrs), 2, _lib.ptr(b, _lib.dp),
                                       _lib.ptr(x, _lib.dp), None))
    np.testing.assert_allclose(x, np.linalg.solve(L @ U, b), rtol=1e-14, atol=1e-14)
