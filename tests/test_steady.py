"""The steady-state solver's host parts on the CPU: the Anderson step is
pinned by what Anderson acceleration must do on linear maps (Walker & Ni,
SIAM J. Numer. Anal. 49(4), 2011: type II with full memory is GMRES on
(I - A) x = b, so it reaches the fixed point of an n-dimensional affine map
in at most n + 1 evaluations), and the potential coordinates by round trips
through the Green's response of a real geometry."""
import numpy as np
import pytest

from tests import cases
from trench01_amd import charging, model, steady


def affine(n, rho, seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    A = Q @ np.diag(rng.uniform(-rho, rho, n)) @ Q.T
    b = rng.normal(size=n)
    return A, b, np.linalg.solve(np.eye(n) - A, b)


def test_full_memory_anderson_solves_affine_map_in_n_plus_one_steps():
    """GMRES terminates in n steps on an n-dimensional system, so the (n+1)-th
    evaluation of full-memory Anderson sits at the fixed point up to the
    rounding of the normal equations."""
    n = 12
    A, b, xs = affine(n, 0.99, 1)
    res = []
    steady.anderson_iterate(lambda x: (None, A @ x + b), np.zeros(n), [], tol=0.0, memory=n + 1,
                            max_evals=n + 1, log=lambda k, r: res.append(r))
    assert len(res) == n + 1
    assert res[n] < 1e-5 * res[0]
    assert res[n] < 1e-2 * res[n - 1]          # the drop comes at the n+1-th evaluation


def test_anderson_beats_plain_iteration_on_slow_map():
    n = 40
    A, b, xs = affine(n, 0.995, 2)                # plain iteration: ~ 5000 steps for 1e-10
    _, gx, evals = steady.anderson_iterate(lambda x: (None, A @ x + b), np.zeros(n), [], tol=1e-9, memory=10,
                                           max_evals=400)
    np.testing.assert_allclose(gx, xs, atol=1e-6)
    assert evals < 200


def test_anderson_update_without_pairs_is_the_plain_step():
    gx = np.arange(5.0)
    out = steady.anderson_update([], gx - 1.0, gx)
    np.testing.assert_array_equal(out, gx)
    assert out is not gx


def test_secant_pairs_carry_over_to_a_perturbed_map():
    """A second map with the same Jacobian and a shifted fixed point (a new
    particle sample) converges in fewer evaluations with the first map's
    secant pairs than from an empty memory."""
    n = 30
    A, b, xs = affine(n, 0.99, 3)
    b2 = b + 1e-3 * np.random.default_rng(4).normal(size=n)
    pairs = []
    _, x1, _ = steady.anderson_iterate(lambda x: (None, A @ x + b), np.zeros(n), pairs, tol=1e-9, memory=n)
    warm = steady.anderson_iterate(lambda x: (None, A @ x + b2), x1, pairs, tol=1e-9, memory=n)[2]
    cold = steady.anderson_iterate(lambda x: (None, A @ x + b2), x1, [], tol=1e-9, memory=n)[2]
    assert warm < cold


@pytest.mark.gpu                  # the set-up solves its Green's columns on the device
@pytest.mark.parametrize("name", ["taper", "ar5"])
def test_potential_coordinates_round_trip(name):
    rp = {"taper": cases.taper_params, "ar5": cases.ar5_params}[name]()
    names = set(model.Params.__dataclass_fields__)
    p = model.Params(**{k: getattr(rp, k) for k in names if hasattr(rp, k)})
    s = charging.ChargingSetup(p)
    to_u, to_rho = steady.potential_map(s)
    rng = np.random.default_rng(0)
    rho = rng.normal(0, 1e6, s.geo.surface.size) * (1 - s.cond_s)
    if s.fields.floating:
        rho[s.c_idx] = -4e8
        rho[np.flatnonzero(s.cond_s)[-1]] += 1e8     # a conductor's charge may sit on several cells
    u = to_u(rho)
    # u is the surface potential of the charge: the Green's response times the column totals
    q = np.zeros(s.green.shape[1])
    np.add.at(q, s.col_of[s.col_of >= 0], rho[s.col_of >= 0] * p.dx ** 2)
    P = s.surface_response
    cols = [np.flatnonzero(s.col_of == c)[0] for c in range(q.size)]
    np.testing.assert_allclose(u, P[cols] @ rho * p.dx ** 2, rtol=1e-10, atol=1e-10 * np.abs(u).max())
    back = to_rho(u, rho)
    np.testing.assert_allclose(back, rho, rtol=1e-9, atol=1e-9 * np.abs(rho).max())
    u2 = u + rng.normal(0, 1.0, u.size)
    np.testing.assert_allclose(to_u(to_rho(u2, rho)), u2, atol=1e-9 * np.abs(u2).max())
    # the within-conductor distribution of the template is kept apart from the representative cell
    if s.fields.floating:
        moved = to_rho(u2, rho)
        keep = s.cond_s.astype(bool)
        keep[cols] = False
        np.testing.assert_array_equal(moved[keep], rho[keep])


def test_patience_stops_a_noisy_map():
    """A map whose residual cannot fall below its noise stops after `patience`
    evaluations without progress instead of running to max_evals."""
    rng = np.random.default_rng(5)
    n = 8
    A, b, xs = affine(n, 0.5, 6)
    calls = []

    def g(x):
        calls.append(1)
        return None, A @ x + b + rng.normal(0, 1e-3, n)

    steady.anderson_iterate(g, np.zeros(n), [], tol=1e-12, memory=5, max_evals=200, patience=6)
    assert len(calls) < 60


def test_gmres_solves_a_small_system_in_n_steps():
    rng = np.random.default_rng(8)
    n = 9
    A = rng.normal(size=(n, n)) + 5 * np.eye(n)
    b = rng.normal(size=n)
    x, k = steady.gmres(lambda v: A @ v, b, krylov=n)
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)
    assert k == n


def test_gmres_forcing_term_stops_at_the_requested_residual():
    """With rtol, GMRES stops at the first iteration whose least-squares
    residual |b - A x| is at most rtol |b| -- the true residual of the x it
    returns -- and the previous iteration's was above it."""
    rng = np.random.default_rng(9)
    n = 40
    A = np.eye(n) + 0.3 * rng.normal(size=(n, n)) / np.sqrt(n)
    b = rng.normal(size=n)
    for rtol in (0.3, 0.1, 1e-3):
        x, k = steady.gmres(lambda v: A @ v, b, krylov=30, rtol=rtol)
        assert np.linalg.norm(A @ x - b) <= rtol * np.linalg.norm(b) * (1 + 1e-9)
        xp, _ = steady.gmres(lambda v: A @ v, b, krylov=k - 1)
        assert k == 1 or np.linalg.norm(A @ xp - b) > rtol * np.linalg.norm(b)
        assert k < 30


def test_newton_krylov_solves_an_affine_map_in_one_newton_step():
    """With exact differences (affine map) and a Krylov space of full
    dimension, one Newton step lands on the fixed point."""
    n = 15
    A, b, xs = affine(n, 0.99, 6)
    res = []
    _, gx, evals = steady.newton_krylov_iterate(lambda x: (None, A @ x + b), np.zeros(n), tol=1e-8, h=1.0,
                                                krylov=n, max_newton=3, log=lambda k, r, e: res.append(r))
    np.testing.assert_allclose(gx, xs, atol=1e-8)
    assert len(res) == 2 and res[1] < 1e-8          # converged after one Newton step
    assert evals == 1 + n + 1                        # no backtracking on an affine map
