"""Host forest (p4est stand-in) invariants: p4est is absent, so topology,
balance and the boundary arrays are pinned by construction checks."""
import ctypes as C

import numpy as np
import pytest

import amrhip

ROOT_LEN = 1 << 30


def corners(geom, k):
    g = geom[12 * k:12 * k + 12].reshape(4, 3)
    return g[0, 0], g[3, 0], g[0, 1], g[3, 1]   # xlo, xhi, ylo, yhi


def random_forest(seed, n=4, ntx=2, nty=1, periodic=(1, 0), lmin=1, lmax=5, rounds=3):
    rng = np.random.default_rng(seed)
    f = amrhip.Forest(ntx, nty, periodic, -1.0, 0.5, 0.75, n, lmin, lmax)
    for _ in range(rounds):
        f.refine((rng.uniform(size=f.nquad) < 0.3).astype(np.int32))
        f.balance()
    return f


def test_uniform_counts():
    f = amrhip.Forest(3, 2, (1, 1), 0.0, 0.0, 1.0, 8, 3, 3)
    assert f.nquad == 6 * 64
    ci, fi, geom, soln = f.external()
    fi = fi.reshape(-1, 11)
    assert len(fi) == 2 * f.nquad and np.all(fi[:, 0] == 2)
    assert np.array_equal(ci.reshape(-1, 3)[:, 2], 4 * 64 * np.arange(f.nquad))


def test_geometry_closed_form():
    f = amrhip.Forest(2, 3, (0, 0), -1.0, 0.5, 0.75, 2, 2, 2)
    _, _, geom, _ = f.external()
    for k, (tree, level, x, y) in enumerate(f.quads()):
        tx, ty = tree % 2, tree // 2
        size = 0.75 / 2 ** level
        x0 = -1.0 + 0.75 * tx + 0.75 * x / ROOT_LEN
        y0 = 0.5 + 0.75 * ty + 0.75 * y / ROOT_LEN
        g = geom[12 * k:12 * k + 12].reshape(4, 3)
        exp = np.array([[x0, y0], [x0 + size, y0], [x0, y0 + size], [x0 + size, y0 + size]])
        assert np.allclose(g[:, :2], exp, rtol=0, atol=1e-15)


@pytest.mark.parametrize("seed", range(6))
def test_balance_and_face_table(seed):
    f = random_forest(seed)
    assert f.is_balanced()
    ci, fi, geom, _ = f.external()
    ci = ci.reshape(-1, 3)
    fi = fi.reshape(-1, 11)
    lv = ci[:, 1]
    seen = np.zeros((f.nquad, 4), int)
    Lx, Ly = 2 * 0.75, 1 * 0.75
    for face in fi:
        t, s, c = face[0], face[1], face[2]
        seen[c, s] += 1
        if t == 1:
            xlo, xhi, ylo, yhi = corners(geom, c)
            # boundary faces only on non-periodic y sides
            assert s in (2, 3)
            assert np.isclose(ylo if s == 2 else yhi, 0.5 if s == 2 else 0.5 + Ly)
            continue
        nbs = [face[3]] if t == 2 else [face[3], face[4]]
        assert np.all(face[7:7 + len(nbs)] == s ^ 1)
        assert np.all(face[3 + len(nbs):7] == -1)
        for j, d in enumerate(nbs):
            seen[d, s ^ 1] += 1
            assert lv[d] == lv[c] + (t == 3)
            a = corners(geom, c)
            b = corners(geom, d)
            if s in (0, 1):
                xa = a[1] if s == 1 else a[0]
                xb = b[0] if s == 1 else b[1]
                assert np.isclose((xa - xb) % Lx, 0) or np.isclose((xa - xb) % Lx, Lx)
                assert a[2] - 1e-12 <= b[2] and b[3] <= a[3] + 1e-12
            else:
                ya = a[3] if s == 3 else a[2]
                yb = b[2] if s == 3 else b[3]
                assert np.isclose(ya, yb)
                assert a[0] - 1e-12 <= b[0] and b[1] <= a[1] + 1e-12
        if t == 3:
            # hanging order: increasing coordinate along the face
            b0, b1 = corners(geom, nbs[0]), corners(geom, nbs[1])
            if s in (0, 1):
                assert b0[2] < b1[2]
            else:
                assert b0[0] < b1[0]
    assert np.all(seen == 1), "every quadrant side is described exactly once"


@pytest.mark.parametrize("seed", range(6))
def test_direct_mesh_table_equals_face_table(seed):
    """amr_forest_build_nbr (the device table straight from the forest) is the
    table the library derives from the reference-format face_info, and its
    levels and origins are cell_info's levels and geom's lower-left corners"""
    f = random_forest(seed)
    ci, fi, geom, _ = f.external()
    nbr, level, org = f.mesh_table()
    assert np.array_equal(nbr, amrhip.nbr_from_faces(f.nquad, fi))
    assert np.array_equal(level, ci.reshape(-1, 3)[:, 1])
    assert np.array_equal(org.reshape(-1, 2), geom.reshape(-1, 4, 3)[:, 0, :2])
    # a planted inconsistency in face_info is rejected, not turned into a table
    bad = fi.copy().reshape(-1, 11)
    k = int(np.nonzero(bad[:, 0] == 2)[0][0])
    bad[k, 3] = bad[k, 2]
    bad[k, 1] ^= 1
    with pytest.raises(amrhip.GpuError):
        amrhip.nbr_from_faces(f.nquad, bad.reshape(-1))


def side_cells(n, s):
    """(i, j) of the n sub-cells along side s (-x, +x, -y, +y) of a patch"""
    t = np.arange(n)
    return {0: (0 * t, t), 1: (0 * t + n - 1, t), 2: (t, 0 * t), 3: (t, 0 * t + n - 1)}[s]


def raw_mesh_table(f):
    """amr_forest_build_nbr's table and status, also on an unbalanced forest
    (sides facing a neighbour two or more levels away are left at -1)"""
    nbr = np.empty(8 * f.nquad, np.int32)
    level = np.empty(f.nquad, np.int32)
    org = np.empty(2 * f.nquad, np.float64)
    pi = C.POINTER(C.c_int)
    rc = amrhip.lib().amr_forest_build_nbr(f.h, nbr.ctypes.data_as(pi), level.ctypes.data_as(pi),
                                           org.ctypes.data_as(C.POINTER(C.c_double)))
    return rc, nbr.reshape(-1, 4, 2)


@pytest.mark.parametrize("seed", range(5))
def test_host_indicator_per_quadrant_sides(seed):
    """amr_indicator_gradient walks the reference-format face list; the same
    indicator from every quadrant's own sides through the direct mesh table
    (index-aligned pairs across conforming and hanging faces, the jump
    |a - b| / ((a + b) / 2) being symmetric) must agree bit for bit.  Seed 4
    refines one quadrant twice without balancing, as the passes of the
    initial refinement do: sides facing a neighbour two levels away have no
    pairs, in both formulations"""
    n = 4
    if seed < 4:
        f = random_forest(seed, n=n, periodic=(1, 1) if seed % 2 else (1, 0))
    else:
        # uniform level 1, the lower-left quadrant refined, then its
        # upper-right child again: level 3 next to level 1
        f = amrhip.Forest(1, 1, (1, 0), 0.0, 0.0, 1.0, n, 1, 5)
        f.refine(np.array([1, 0, 0, 0], np.int32))
        f.refine(np.eye(f.nquad, dtype=np.int32)[3])
        assert not f.is_balanced()
    rng = np.random.default_rng(300 + seed)
    s = f.soln()
    s[:] = rng.uniform(0.5, 1.5, s.shape)
    got = f.indicator()
    rho = f.soln().reshape(f.nquad, n, n, 4)[..., 0]      # [quad, j, i]
    rc, nbr = raw_mesh_table(f)
    assert rc == (-1 if seed == 4 else 0)
    lv = f.levels()

    def jump(a, b):
        return np.abs(a - b) / (0.5 * (a + b))
    exp = np.zeros(f.nquad)
    for k in range(f.nquad):
        r = rho[k]
        m = max(jump(r[:, :-1], r[:, 1:]).max(), jump(r[:-1, :], r[1:, :]).max())
        for sd in range(4):
            i, j = side_cells(n, sd)
            ib, jb = side_cells(n, sd ^ 1)
            for d in nbr[k, sd]:
                if d >= 0 and abs(int(lv[d]) - int(lv[k])) <= 1:
                    m = max(m, jump(r[j, i], rho[d][jb, ib]).max())
        exp[k] = m
    assert np.array_equal(got, exp)


def test_refine_copies_and_coarsen_averages(oracle):
    import pyoracle
    f = amrhip.Forest(1, 1, (1, 1), 0.0, 0.0, 1.0, 4, 1, 3)
    rng = np.random.default_rng(4)
    s = f.soln()
    s[:] = rng.uniform(0.5, 1.5, s.shape)
    parent = s[2].copy()
    flags = np.zeros(f.nquad, np.int32)
    flags[2] = 1
    f.refine(flags)
    s = f.soln()
    for c in range(4):
        assert np.array_equal(s[2 + c], parent)
    # perturb children, then coarsen the family (tags are 0)
    s[2:6] = rng.uniform(0.5, 1.5, s[2:6].shape)
    ch = [np.ascontiguousarray(s[2 + c].reshape(-1)) for c in range(4)]
    exp = np.zeros(ch[0].size)
    oracle.L.orc_coarsen_2d(4, *[pyoracle.dpt(c) for c in ch], pyoracle.dpt(exp))
    tags = f.tags()
    tags[:] = 1
    tags[2:6] = 0
    assert f.coarsen() == 1
    assert f.nquad == 4
    assert np.array_equal(f.soln()[2].reshape(-1), exp)


def test_find_leaf_and_levels():
    f = random_forest(9)
    L = amrhip.lib()
    for k, (tree, level, x, y) in enumerate(f.quads()):
        assert L.amr_forest_find_leaf(f.h, tree, x, y) == k
        size = ROOT_LEN >> level
        assert L.amr_forest_find_leaf(f.h, tree, x + size - 1, y + size - 1) == k


def test_regrid_initial_through_unbalanced_passes():
    """the initial refinement balances only after its last pass, so its
    indicator sees forests with two-level jumps; it must still end balanced
    (periodic vortex, levels 2-4: the mesh of the device-loop lag test)"""
    ic = amrhip.vortex_ic(5.0, 5.0, 5.0)
    f = amrhip.Forest(1, 1, (1, 1), 0.0, 0.0, 10.0, 8, 2, 4)
    f.apply_ic(ic)
    nq = f.regrid_initial(ic, 0.02)
    assert nq == f.nquad > 16
    assert f.is_balanced() and f.levels().max() == 4
    rc, _ = raw_mesh_table(f)
    assert rc == 0


def test_regrid_refines_discontinuity():
    ic = amrhip.riemann4_ic(0)
    f = amrhip.Forest(1, 1, (0, 0), 0.0, 0.0, 1.0, 8, 2, 5)
    f.apply_ic(ic)
    f.regrid_initial(ic, 0.1)
    assert f.is_balanced()
    lv = f.levels()
    assert lv.max() == 5 and lv.min() >= 2
    ci, fi, geom, _ = f.external()
    # the finest quadrants touch the discontinuity lines x = 0.5 or y = 0.5
    for k in np.nonzero(lv == 5)[0]:
        xlo, xhi, ylo, yhi = corners(geom, k)
        near = min(abs(xlo - 0.5), abs(xhi - 0.5), abs(ylo - 0.5), abs(yhi - 0.5))
        assert near < 2 * (xhi - xlo) + 1e-12


@pytest.mark.parametrize("seed", range(4))
def test_regrid_tracking_reproduces_host_data(seed):
    """the recorded source map (copy / family mean) applied to the old data
    equals the data the host operators moved, bit for bit"""
    rng = np.random.default_rng(100 + seed)
    n = 4
    f = amrhip.Forest(2, 1, (1, 0), 0.0, 0.0, 1.0, n, 1, 5)
    for _ in range(3):
        f.refine((rng.uniform(size=f.nquad) < 0.3).astype(np.int32))
        f.balance()
    s = f.soln()
    s[:] = rng.uniform(0.5, 1.5, s.shape)
    old = s.copy()
    tags = (rng.uniform(size=f.nquad) < 0.2).astype(np.int32)
    f.track_begin(keep_data=True)
    f.regrid_tags(tags, allow_coarsening=True)
    ok, src = f.track_end()
    assert ok
    assert np.any(src < 0) and f.is_balanced()
    new = f.soln()
    for k, sk in enumerate(src):
        if sk >= 0:
            exp = old[sk]
        else:
            b = -sk - 1
            exp = 0.25 * (((old[b] + old[b + 1]) + old[b + 2]) + old[b + 3])
        assert np.array_equal(new[k], exp)


def test_tracking_without_data_gives_same_topology():
    rng = np.random.default_rng(7)
    fa = amrhip.Forest(1, 1, (0, 0), 0.0, 0.0, 1.0, 4, 2, 5)
    fb = amrhip.Forest(1, 1, (0, 0), 0.0, 0.0, 1.0, 4, 2, 5)
    for _ in range(2):
        fl = (rng.uniform(size=fa.nquad) < 0.3).astype(np.int32)
        fa.refine(fl); fa.balance(); fb.refine(fl); fb.balance()
    tags = (rng.uniform(size=fa.nquad) < 0.2).astype(np.int32)
    fa.track_begin(keep_data=True); fa.regrid_tags(tags); oka, sa = fa.track_end()
    fb.track_begin(keep_data=False); fb.regrid_tags(tags); okb, sb = fb.track_end()
    assert oka and okb and np.array_equal(sa, sb) and fa.quads() == fb.quads()


@pytest.mark.parametrize("seed", range(12))
def test_worklist_balance_equals_full_balance(seed):
    """amr_regrid_tags balances only around what it refined / coarsened; the
    result must be the same closure as full passes over every quadrant"""
    rng = np.random.default_rng(500 + seed)
    per = (bool(seed % 2), bool(seed % 3 == 0))
    f = amrhip.Forest(3, 2, per, 0.0, 0.0, 1.0, 4, 1, 7)
    for _ in range(4):
        f.refine((rng.uniform(size=f.nquad) < 0.25).astype(np.int32))
        f.balance()
    L = amrhip.lib()
    for step in range(4):
        tags = rng.choice([0, 1, 2], size=f.nquad, p=[0.6, 0.15, 0.25]).astype(np.int32)
        g = amrhip.Forest.from_quads(3, 2, per, 0.0, 0.0, 1.0, 4, 1, 7, f.quads())
        # reference: the same refine / coarsen, then full-pass balance
        gt = g.tags()
        gt[:] = np.where(tags == 1, 1, np.where(tags != 0, 2, 0))
        g.refine((tags == 1).astype(np.int32))
        L.amr_forest_coarsen(g.h)
        g.balance()
        f.regrid_tags(tags, allow_coarsening=True)
        assert f.is_balanced()
        assert f.quads() == g.quads(), (seed, step)
