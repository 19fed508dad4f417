"""Closed-form checks that pin the oracle where golden vectors cannot: the RK
schemes (the reference's rk_* needs mpi.h and is not built here) and
invariants of the flux and residual operator that a dropped term, a wrong sign
or a transposed operand would break."""
import numpy as np
import pytest

import amrhip
from conftest import parity_err


def cons(rho, u, v, p, gamma=1.4):
    return np.array([rho, rho * u, rho * v, p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)])


def phys_flux(q, n):
    """F(q).n from the Euler equations (textbook form, gamma = 1.4)"""
    rho, u, v = q[0], q[1] / q[0], q[2] / q[0]
    p = 0.4 * (q[3] - 0.5 * rho * (u * u + v * v))
    fx = np.array([rho * u, rho * u * u + p, rho * u * v, (q[3] + p) * u])
    fy = np.array([rho * v, rho * u * v, rho * v * v + p, (q[3] + p) * v])
    return fx * n[0] + fy * n[1]


@pytest.mark.parametrize("seed", range(5))
def test_flux_consistency(oracle, seed):
    rng = np.random.default_rng(seed)
    for _ in range(50):
        q = cons(rng.uniform(0.1, 2), *rng.uniform(-1, 1, 2), rng.uniform(0.1, 2))
        nrm = rng.uniform(-1, 1, 2)
        f = oracle.flux(q, q, nrm)
        ref = phys_flux(q, nrm)
        scale = np.abs(phys_flux(q, [1, 0])).max() * np.hypot(*nrm)
        assert np.max(np.abs(f - ref)) <= 1e-14 * scale


def test_flux_supersonic_upwind(oracle):
    rng = np.random.default_rng(7)
    for _ in range(100):
        rho, p = rng.uniform(0.5, 2), rng.uniform(0.1, 1)
        c = np.sqrt(1.4 * p / rho)
        ql = cons(rho, 3 * c + rng.uniform(0, 1), rng.uniform(-0.1, 0.1), p)
        qr = cons(rho * 1.01, 3 * c + rng.uniform(0, 1), rng.uniform(-0.1, 0.1), p * 1.01)
        h = 0.125
        f = oracle.flux(ql, qr, [h, 0.0])
        assert np.allclose(f, phys_flux(ql, [h, 0.0]), rtol=4e-16, atol=0)
        # mirror case: everything moving left takes the right state
        f = oracle.flux(ql * [1, -1, 1, 1], qr * [1, -1, 1, 1], [h, 0.0])
        assert np.allclose(f, phys_flux(qr * [1, -1, 1, 1], [h, 0.0]), rtol=4e-16, atol=0)


def test_flux_mirror_symmetry(oracle):
    rng = np.random.default_rng(3)
    sw = [0, 2, 1, 3]
    for _ in range(200):
        ql = cons(rng.uniform(0.1, 2), *rng.uniform(-2, 2, 2), rng.uniform(0.1, 2))
        qr = cons(rng.uniform(0.1, 2), *rng.uniform(-2, 2, 2), rng.uniform(0.1, 2))
        h = 0.25
        fx = oracle.flux(ql, qr, [h, 0.0])
        fy = oracle.flux(ql[sw], qr[sw], [0.0, h])
        assert np.allclose(fy[sw], fx, rtol=1e-15, atol=1e-15 * np.abs(fx).max())


def test_slip_wall_pure_pressure(oracle):
    rng = np.random.default_rng(5)
    for _ in range(50):
        rho, u, v, p = rng.uniform(0.2, 2), *rng.uniform(-1, 1, 2), rng.uniform(0.2, 2)
        q = cons(rho, u, v, p)
        for side, nrm in enumerate([(-0.5, 0.0), (0.5, 0.0), (0.0, -0.5), (0.0, 0.5)]):
            qb = oracle.bc(1, nrm, q)
            ut = v if side < 2 else u
            pb = 0.4 * (q[3] - 0.5 * rho * ut * ut)
            f = oracle.flux_bc(qb, nrm)
            exp = np.array([0.0, pb * nrm[0], pb * nrm[1], 0.0])
            assert np.allclose(f, exp, rtol=1e-14, atol=1e-14 * abs(pb))


def forest(n, periodic, seed, lmin=2, lmax=4, frac=0.3):
    rng = np.random.default_rng(seed)
    f = amrhip.Forest(1, 1, periodic, 0.0, 0.0, 1.0, n, lmin, lmax)
    for _ in range(2):
        f.refine((rng.uniform(size=f.nquad) < frac).astype(np.int32))
        f.balance()
    return f, rng


@pytest.mark.parametrize("n", [2, 4, 8])
def test_free_stream_preservation(oracle, n):
    f, _ = forest(n, (1, 1), 11)
    q = cons(1.3, 0.7, -0.4, 0.9)
    f.soln()[:] = q
    ci, fi, geom, soln = f.external()
    assert np.any(fi.reshape(-1, 11)[:, 0] == 3)
    p = oracle.params(n, bc=(0, 0, 0, 0))
    rhs = oracle.rhs(p, ci, fi, geom, soln)
    scale = np.abs(phys_flux(q, [1, 0])).max() / (2.0 ** -4 / n)
    assert np.max(np.abs(rhs)) <= 1e-12 * scale


@pytest.mark.parametrize("n", [2, 4, 8])
def test_discrete_conservation(oracle, n):
    f, rng = forest(n, (1, 1), 12)
    s = f.soln()
    s[..., 0] = rng.uniform(0.5, 1.5, s.shape[:-1])
    s[..., 1] = rng.uniform(-0.5, 0.5, s.shape[:-1])
    s[..., 2] = rng.uniform(-0.5, 0.5, s.shape[:-1])
    s[..., 3] = rng.uniform(2.0, 3.0, s.shape[:-1])
    ci, fi, geom, soln = f.external()
    p = oracle.params(n, bc=(0, 0, 0, 0))
    rhs = oracle.rhs(p, ci, fi, geom, soln).reshape(f.nquad, n * n, 4)
    vol = ((1.0 / 2.0 ** f.levels()) / n) ** 2
    tot = np.einsum("q,qcf->f", vol, rhs)
    mag = np.einsum("q,qcf->f", vol, np.abs(rhs))
    assert np.all(np.abs(tot) <= 1e-13 * mag)


def test_time_step_closed_form(oracle):
    n = 4
    f = amrhip.Forest(2, 1, (0, 0), 0.0, 0.0, 2.0, n, 1, 1)
    rho, u, v, p = 0.8, 0.3, -0.6, 1.1
    f.soln()[:] = cons(rho, u, v, p)
    ci, fi, geom, soln = f.external()
    pr = oracle.params(n, cfl=0.3)
    h = 2.0 / 2 / n
    exp = 0.3 * h / (np.hypot(u, v) + np.sqrt(1.4 * p / rho))
    assert oracle.time_step(pr, ci, geom, soln, 1.0) == pytest.approx(exp, rel=1e-14)
    assert oracle.time_step(pr, ci, geom, soln, 1e-6) == 1e-6


RK_POLY = {
    1: [1, 1],
    2: [1, 1, 1 / 2],
    3: [1, 1, 1 / 2, 1 / 6],
    4: [1, 1, 1 / 2, 1 / 6, 1 / 24],
}


@pytest.mark.parametrize("scheme", [1, 2, 3, 4])
def test_rk_stability_polynomial(oracle, scheme):
    lam, dt = -1.7, 0.3
    q0 = np.array([1.0, -2.0, 0.5])
    q1 = oracle.rk_generic(scheme, dt, q0, lambda q: lam * q)
    z = lam * dt
    R = sum(c * z ** k for k, c in enumerate(RK_POLY[scheme]))
    assert np.allclose(q1, R * q0, rtol=1e-14)


def test_rk_systems_rotation(oracle):
    # L(Q) = A Q with A a rotation generator: order conditions on a system
    A = np.array([[0.0, -1.0], [1.0, 0.0]])
    q0 = np.array([1.0, 0.0])
    for scheme, order in [(1, 1), (2, 2), (3, 3), (4, 4), (5, 4)]:
        errs = []
        for dt in (0.02, 0.01):
            q = q0.copy()
            for _ in range(int(round(1.0 / dt))):
                q = oracle.rk_generic(scheme, dt, q, lambda x: A @ x)
            errs.append(np.linalg.norm(q - np.array([np.cos(1.0), np.sin(1.0)])))
        rate = np.log2(errs[0] / errs[1])
        assert rate == pytest.approx(order, abs=0.25), (scheme, rate)


def test_lserk_local_error_order(oracle):
    lam = -1.0
    q0 = np.array([1.0])
    e = []
    for dt in (0.02, 0.01):
        q1 = oracle.rk_generic(5, dt, q0, lambda q: lam * q)
        e.append(abs(q1[0] - np.exp(lam * dt)))
    assert np.log2(e[0] / e[1]) == pytest.approx(5.0, abs=0.3)


def test_parity_metric_catches_planted_mistake(oracle):
    """a sign error in one field of one cell is caught by the 1e-12 bar"""
    f, rng = forest(4, (1, 1), 13)
    s = f.soln()
    s[..., 0] = rng.uniform(0.5, 1.5, s.shape[:-1])
    s[..., 1] = rng.uniform(-0.2, 0.2, s.shape[:-1])
    s[..., 3] = 2.5
    ci, fi, geom, soln = f.external()
    rhs = oracle.rhs(oracle.params(4), ci, fi, geom, soln)
    bad = rhs.copy()
    k = 4 * int(np.argmax(np.abs(rhs[1::4]))) + 1     # largest x-momentum residual
    bad[k] *= -1.0
    assert parity_err(rhs, rhs) == 0.0
    assert parity_err(bad, rhs) > 1e-6


@pytest.mark.parametrize("n", [2, 4, 8, 16])
def test_norm_closed_forms(oracle, n):
    """the residual sum of squares counts exactly the real patches at their
    cell_info offsets: all-ones gives 4 n^2 ncell exactly, ghost patches and
    patches no cell names do not count, and the value equals the correctly
    rounded sum (math.fsum) to within the rounding of a plain running sum"""
    import math
    rng = np.random.default_rng(n)
    ncell, nghost = 11, 4
    m = 4 * n * n
    ci = np.zeros((ncell, 3), np.int32)
    perm = rng.permutation(ncell + nghost)
    ci[:, 2] = perm[:ncell] * m
    ones = np.ones((ncell + nghost) * m)
    assert oracle.norm(n, ci.reshape(-1), ones) == 4.0 * n * n * ncell
    res = rng.standard_normal((ncell + nghost) * m)
    real = np.concatenate([res[o:o + m] for o in ci[:, 2]])
    exact = math.fsum(real * real)
    got = oracle.norm(n, ci.reshape(-1), res)
    assert abs(got - exact) <= 1e-13 * exact
    poisoned = res.copy()
    for g in perm[ncell:]:
        poisoned[g * m:(g + 1) * m] = np.nan
    assert oracle.norm(n, ci.reshape(-1), poisoned) == got


@pytest.mark.parametrize("scheme", [1, 2, 3, 4])
def test_rk_norm_is_last_stage(oracle, scheme):
    """the step's residual norm is the 2-norm of the LAST stage's residual
    (rk_time_integrators.c keeps only the last physics_norm_L2_residual):
    on L(Q) = lambda Q the last stage evaluates L at a known polynomial of
    z = lambda dt times Q0 (RK1: Q0; SSP-RK2: 1 + z; SSP-RK3: 1 + z/2 + z^2/4;
    classical RK4: 1 + z + z^2/2 + z^3/4)"""
    lam, dt = -0.7, 0.3
    z = lam * dt
    last = {1: 1.0, 2: 1.0 + z, 3: 1.0 + z / 2 + z * z / 4, 4: 1.0 + z + z * z / 2 + z ** 3 / 4}[scheme]
    q0 = np.array([1.0, -2.0, 0.5, 3.0])
    _, rn = oracle.rk_generic(scheme, dt, q0, lambda q: lam * q, with_norm=True)
    assert rn == pytest.approx(abs(lam) * abs(last) * np.linalg.norm(q0), rel=1e-14)
