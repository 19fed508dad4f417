"""Project wall types 3-5 (double-Mach reflection: prescribed inflow,
post-shock / slip bottom wall, moving shock along the top) in the CPU
oracle.  They have no reference counterpart; closed forms pin them: a
uniform state equal to the prescribed one leaves every residual zero, and
over a gas at rest exactly the boundary cells whose face centre lies behind
the shock line (or before x0 on the bottom) change."""
import numpy as np
import pytest

import amrhip

POST, PRE, W = amrhip.dmr_states()
SHOCK = (1.0 / 6.0, 60.0, W)


def brick(n=4, lmin=2, lmax=3, seed=0):
    """[0,4] x [0,1] as four unit trees, some quadrants one level finer"""
    rng = np.random.default_rng(seed)
    f = amrhip.Forest(4, 1, (0, 0), 0.0, 0.0, 1.0, n, lmin, lmax)
    f.refine((rng.uniform(size=f.nquad) < 0.3).astype(np.int32))
    f.balance()
    return f


def fill(f, state):
    s = f.soln()
    s[:] = np.asarray(state, float)
    return f


def top_bottom_cells(f, geom, side):
    """(quad, i) of the top (side 3) or bottom (side 2) edge cells on the
    domain boundary, with their face-centre coordinates"""
    n = f.n
    out = []
    for k, (tree, lv, x, y) in enumerate(f.quads()):
        g = geom[12 * k:12 * k + 12].reshape(4, 3)
        h = (g[1, 0] - g[0, 0]) / n
        on = (side == 3 and abs(g[3, 1] - 1.0) < 1e-14) or (side == 2 and abs(g[0, 1]) < 1e-14)
        if on:
            for i in range(n):
                out.append((k, i, g[0, 0] + (i + 0.5) * h, g[3, 1] if side == 3 else g[0, 1]))
    return out


def test_uniform_prescribed_state_is_steady(oracle):
    f = fill(brick(), POST)
    ci, fi, geom, soln = f.external()
    p = oracle.params(4, bc=(3, 3, 3, 3), bstate=(POST, PRE), shock=SHOCK, time=0.0)
    r = oracle.rhs(p, ci, fi, geom, soln)
    assert np.abs(r).max() <= 1e-11 * np.abs(soln).max()


@pytest.mark.parametrize("t", [0.0, 0.05, 0.15])
def test_moving_shock_top(oracle, t):
    f = fill(brick(seed=1), PRE)
    ci, fi, geom, soln = f.external()
    p = oracle.params(4, bc=(0, 0, 1, 5), bstate=(POST, PRE), shock=SHOCK, time=t)
    r = oracle.rhs(p, ci, fi, geom, soln).reshape(f.nquad, f.n, f.n, 4)
    xs = lambda y: SHOCK[0] + y / np.sqrt(3.0) + W * 2.0 / np.sqrt(3.0) * t
    behind = ahead = 0
    for k, i, xf, yf in top_bottom_cells(f, geom, 3):
        changed = np.abs(r[k, f.n - 1, i]).max() > 1e-9
        assert changed == (xf < xs(yf)), (k, i, xf, xs(yf))
        behind += changed
        ahead += not changed
    assert behind > 0 and ahead > 0
    # nothing else moves: the gas is at rest and the other walls are copy/slip
    r2 = r.copy()
    for k, i, _, _ in top_bottom_cells(f, geom, 3):
        r2[k, f.n - 1, i] = 0.0
    assert np.abs(r2).max() <= 1e-11


def test_post_shock_then_reflecting_bottom(oracle):
    f = fill(brick(seed=2), PRE)
    ci, fi, geom, soln = f.external()
    p = oracle.params(4, bc=(0, 0, 4, 1), bstate=(POST, PRE), shock=SHOCK, time=0.3)
    r = oracle.rhs(p, ci, fi, geom, soln).reshape(f.nquad, f.n, f.n, 4)
    before = 0
    for k, i, xf, _ in top_bottom_cells(f, geom, 2):
        changed = np.abs(r[k, 0, i]).max() > 1e-9
        assert changed == (xf < SHOCK[0])
        before += changed
    assert before > 0
