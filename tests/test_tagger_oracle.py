"""Pins of the learned tagger's CPU oracle (oracle/tagger_oracle.c).

The reference has no learned tagger, so nothing upstream fixes these
functions; they are pinned by what the mathematics fixes instead:
  * BF16 rounding against an integer restatement of round-to-nearest-even;
  * the backward pass (exact mode) against central finite differences of the
    loss and against an independent float64 autograd implementation (torch);
  * a planted transposition of one gradient block, which the finite
    differences must catch;
  * features against closed forms (constant and linear patches) and against
    neighbours found geometrically in the forest, not through face_info.
"""
import numpy as np
import pytest

import amrhip

NP = amrhip.MLP_NPARAM
OB1 = 64 * 384
OW2 = OB1 + 64
OB2 = OW2 + 64 * 64
OW3 = OB2 + 64
OB3 = OW3 + 3 * 64
BLOCKS = {"W1": (0, OB1), "b1": (OB1, OW2), "W2": (OW2, OB2), "b2": (OB2, OW3), "W3": (OW3, OB3),
          "b3": (OB3, NP)}


def bf16_bits(x):
    """float32 -> bf16 bits, round to nearest even, in integer arithmetic"""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    lower = u & 0xFFFF
    upper = u >> 16
    up = (lower > 0x8000) | ((lower == 0x8000) & ((upper & 1) == 1))
    return (upper + up).astype(np.uint16)


def bf16_val(bits):
    return (np.asarray(bits, np.uint32) << 16).view(np.float32)


def random_problem(seed, n=48, wscale=1.0):
    rng = np.random.default_rng(seed)
    X = bf16_bits(rng.normal(0, 0.6, size=(n, 384)).astype(np.float32))
    P = np.zeros(NP, np.float32)
    P[:OB1] = rng.uniform(-1, 1, OB1) * np.sqrt(6 / 384) * wscale
    P[OB1:OW2] = rng.normal(0, 0.1, 64)
    P[OW2:OB2] = rng.uniform(-1, 1, 64 * 64) * np.sqrt(6 / 64) * wscale
    P[OB2:OW3] = rng.normal(0, 0.1, 64)
    P[OW3:OB3] = rng.uniform(-1, 1, 3 * 64) * np.sqrt(6 / 64) * wscale
    P[OB3:] = rng.normal(0, 0.1, 3)
    y = rng.integers(0, 3, size=n).astype(np.int32)
    cw = np.array([0.7, 1.0, 2.5], np.float32)
    return P, X, y, cw


def test_bf16_rounding(oracle):
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 10, 2000).astype(np.float32),
                        # exact ties: lower half 0x8000 with even and odd upper halves
                        np.array([0x3F808000, 0x3F818000, 0xBF808000, 0x00018000], np.uint32).view(np.float32),
                        np.array([0.0, -0.0, 1.0, 65504.0, 1e-38], np.float32)])
    got = np.array([oracle.L.orc_bf16(float(v)) for v in x], np.uint16)
    assert np.array_equal(got, bf16_bits(x))
    # ties go to even: 0x3F808000 -> 0x3F80, 0x3F818000 -> 0x3F82
    assert got[2000] == 0x3F80 and got[2001] == 0x3F82


def torch_exact(P, X, y, cw):
    torch = pytest.importorskip("torch")
    p = torch.tensor(P.astype(np.float64), requires_grad=True)
    x = torch.tensor(bf16_val(X).astype(np.float64))
    W1 = p[:OB1].view(64, 384); b1 = p[OB1:OW2]
    W2 = p[OW2:OB2].view(64, 64); b2 = p[OB2:OW3]
    W3 = p[OW3:OB3].view(3, 64); b3 = p[OB3:]
    h1 = torch.relu(x @ W1.T + b1)
    h2 = torch.relu(h1 @ W2.T + b2)
    z = h2 @ W3.T + b3
    w = torch.tensor(cw.astype(np.float64))[torch.tensor(y, dtype=torch.long)]
    ce = torch.nn.functional.cross_entropy(z, torch.tensor(y, dtype=torch.long), reduction="none")
    loss = (w * ce).sum() / w.sum()
    loss.backward()
    return float(loss.detach()), p.grad.numpy(), z.detach().numpy()


def test_exact_grad_matches_independent_autograd(oracle):
    P, X, y, cw = random_problem(1)
    loss, g = oracle.mlp_grad(P, X, y, cw, exact=True)
    tl, tg, tz = torch_exact(P, X, y, cw)
    assert abs(loss - tl) <= 1e-12 * abs(tl)
    assert np.allclose(oracle.mlp_forward(P, X, exact=True), tz, rtol=1e-12, atol=1e-12)
    for name, (a, b) in BLOCKS.items():
        scale = np.abs(tg[a:b]).max()
        assert scale > 0, name
        assert np.abs(g[a:b] - tg[a:b]).max() <= 1e-10 * scale, name


def fd_check(oracle, P, X, y, cw, g, picks, rel=2e-3):
    """central differences of the exact loss at the picked parameters"""
    bad = []
    for i in picks:
        h = np.float32(1e-4 * max(abs(float(P[i])), 0.02))
        Pp, Pm = P.copy(), P.copy()
        Pp[i] = np.float32(P[i] + h)
        Pm[i] = np.float32(P[i] - h)
        lp, _ = oracle.mlp_grad(Pp, X, y, cw, exact=True)
        lm, _ = oracle.mlp_grad(Pm, X, y, cw, exact=True)
        fd = (lp - lm) / (float(Pp[i]) - float(Pm[i]))
        if abs(fd - g[i]) > rel * max(abs(fd), abs(g[i])) + 1e-9:
            bad.append((i, fd, g[i]))
    return bad


def fd_picks(seed, g):
    rng = np.random.default_rng(seed)
    picks = []
    for a, b in BLOCKS.values():
        blk = np.arange(a, b)
        big = blk[np.argsort(-np.abs(g[a:b]))[:4]]       # entries that matter
        picks += list(big) + list(rng.choice(blk, size=min(3, b - a), replace=False))
    return picks


def test_exact_grad_matches_finite_differences(oracle):
    P, X, y, cw = random_problem(2)
    _, g = oracle.mlp_grad(P, X, y, cw, exact=True)
    bad = fd_check(oracle, P, X, y, cw, g, fd_picks(3, g))
    assert not bad, bad[:5]


def test_finite_differences_catch_transposed_block(oracle):
    """planted mistake: dW2 transposed (the classic operand-order slip)"""
    P, X, y, cw = random_problem(4)
    _, g = oracle.mlp_grad(P, X, y, cw, exact=True)
    wrong = g.copy()
    wrong[OW2:OB2] = g[OW2:OB2].reshape(64, 64).T.reshape(-1)
    a, b = OW2, OB2
    picks = [a + int(k) for k in np.argsort(-np.abs(g[a:b]))[:8]]
    assert fd_check(oracle, P, X, y, cw, wrong, picks), "transposed dW2 went unnoticed"


def test_rounded_grad_close_to_exact(oracle):
    """the BF16 pass stays close to the smooth gradient: rounding the weights
    to BF16 moves some pre-activations across zero, so a few percent differ"""
    P, X, y, cw = random_problem(5, n=128)
    l0, g0 = oracle.mlp_grad(P, X, y, cw, exact=True)
    l1, g1 = oracle.mlp_grad(P, X, y, cw, exact=False)
    assert abs(l1 - l0) <= 2e-2 * abs(l0)
    for name, (a, b) in BLOCKS.items():
        rel = np.linalg.norm(g1[a:b] - g0[a:b]) / np.linalg.norm(g0[a:b])
        assert rel < 0.15, (name, rel)


# ---------------------------------------------------------------------------
# features
# ---------------------------------------------------------------------------
def test_features_constant_and_linear(oracle):
    f = amrhip.Forest(1, 1, (0, 0), 0.0, 0.0, 1.0, 8, 1, 1)   # 2x2 quads, walls all round
    ci, fi, geom, soln = f.external()
    q = soln.reshape(4, 8, 8, 4)          # [quad][j][i][field]
    ii = np.arange(8)[None, :]
    q[:] = 0.0
    q[..., 0] = 2.0                        # constant density
    q[..., 1] = 0.5 * ii                   # linear x momentum (max 3.5)
    q[..., 2] = -1.0
    q[..., 3] = 4.0
    X = oracle.tagger_features(8, ci, fi, soln).reshape(4, 6, 8, 8).astype(np.uint32)
    one = bf16_bits(np.float32(1.0))
    assert np.all(X[:, 0] == one) and np.all(X[:, 3] == one)
    assert np.all(X[:, 2] == bf16_bits(np.float32(-1.0)))
    exp1 = bf16_bits((0.5 * np.arange(8) / 3.5).astype(np.float32))
    assert np.all(X[:, 1] == exp1[None, None, :])
    # constant density: both differences vanish, inside and across every edge
    assert np.all(X[:, 4] == 0) and np.all(X[:, 5] == 0)


def geometric_neighbour(f, k, side, idx):
    """the quadrant holding the neighbour's facing edge cell at sub-face
    index idx, found by point location in the forest (independent of
    face_info); None outside the domain"""
    tree, level, x, y = f.quads()[k]
    size = (1 << 30) >> level
    off = (idx * size) // 8
    px, py = {0: (x - 1, y + off), 1: (x + size, y + off), 2: (x + off, y - 1), 3: (x + off, y + size)}[side]
    if not (0 <= px < (1 << 30) and 0 <= py < (1 << 30)):
        return None
    return amrhip.lib().amr_forest_find_leaf(f.h, tree, px, py)


def test_features_edge_differences_by_geometry(oracle):
    """uniform mesh with walls: the edge differences use the geometric
    neighbour's facing cell (on a uniform mesh the index-aligned cell is it)
    and the cell itself at a wall"""
    rng = np.random.default_rng(6)
    f = amrhip.Forest(1, 1, (0, 0), 0.0, 0.0, 1.0, 8, 2, 2)
    ci, fi, geom, soln = f.external()
    soln[:] = rng.uniform(0.5, 2.0, soln.shape)
    X = oracle.tagger_features(8, ci, fi, soln).reshape(-1, 6, 8, 8)
    q = soln.reshape(-1, 8, 8, 4)
    for k in range(f.nquad):
        s0 = np.abs(q[k, :, :, 0]).max()
        for j in range(8):
            e, w = geometric_neighbour(f, k, 1, j), geometric_neighbour(f, k, 0, j)
            re = q[k, j, 7, 0] if e is None else q[e, j, 0, 0]
            rw = q[k, j, 0, 0] if w is None else q[w, j, 7, 0]
            assert X[k, 4, j, 7] == bf16_bits(np.float32((re - q[k, j, 6, 0]) / s0))
            assert X[k, 4, j, 0] == bf16_bits(np.float32((q[k, j, 1, 0] - rw) / s0))
            assert X[k, 4, j, 3] == bf16_bits(np.float32((q[k, j, 4, 0] - q[k, j, 2, 0]) / s0))
        for i in range(8):
            n_, s_ = geometric_neighbour(f, k, 3, i), geometric_neighbour(f, k, 2, i)
            rn = q[k, 7, i, 0] if n_ is None else q[n_, 0, i, 0]
            rs = q[k, 0, i, 0] if s_ is None else q[s_, 7, i, 0]
            assert X[k, 5, 7, i] == bf16_bits(np.float32((rn - q[k, 6, i, 0]) / s0))
            assert X[k, 5, 0, i] == bf16_bits(np.float32((q[k, 1, i, 0] - rs) / s0))


def test_features_hanging_side_uses_both_fine_neighbours(oracle):
    f = amrhip.Forest(2, 1, (0, 0), 0.0, 0.0, 1.0, 8, 0, 1)
    flags = np.array([0, 1], np.int32)     # refine the right tree: left quad has 2 fine +x neighbours
    f.refine(flags)
    ci, fi, geom, soln = f.external()
    rng = np.random.default_rng(7)
    soln[:] = rng.uniform(0.5, 2.0, soln.shape)
    X = oracle.tagger_features(8, ci, fi, soln).reshape(-1, 6, 8, 8)
    q = soln.reshape(-1, 8, 8, 4)
    fine = [k for k, (t, lv, x, y) in enumerate(f.quads()) if t == 1 and x == 0]
    fine.sort(key=lambda k: f.quads()[k][3])          # increasing y
    assert len(fine) == 2
    s0 = np.abs(q[0, :, :, 0]).max()
    for j in range(8):
        re = 0.5 * (q[fine[0], j, 0, 0] + q[fine[1], j, 0, 0])
        assert X[0, 4, j, 7] == bf16_bits(np.float32((re - q[0, j, 6, 0]) / s0))
    # each fine quadrant sees the coarse one across its -x side (index-aligned)
    for k in fine:
        sk = np.abs(q[k, :, :, 0]).max()
        for j in range(8):
            assert X[k, 4, j, 0] == bf16_bits(np.float32((q[k, j, 1, 0] - q[0, j, 7, 0]) / sk))


def pool16(soln16):
    """16x16 patches (reference layout) coarsened to 8x8: ((a + b) + c) + d
    times 0.25 over each 2x2 block in z-order"""
    q = np.asarray(soln16).reshape(-1, 16, 16, 4)
    a, b = q[:, 0::2, 0::2], q[:, 0::2, 1::2]
    c, d = q[:, 1::2, 0::2], q[:, 1::2, 1::2]
    return (0.25 * (((a + b) + c) + d)).reshape(-1)


def test_features_16_equal_features_of_coarsened_patches(oracle):
    """a 16x16 forest's features are those of the same forest with 8x8
    patches holding the 2x2 means, neighbours included (hanging faces and
    walls)"""
    rng = np.random.default_rng(16)
    out = []
    for n in (16, 8):
        f = amrhip.Forest(2, 1, (1, 0), 0.0, 0.0, 1.0, n, 1, 3)
        r = np.random.default_rng(3)
        for _ in range(2):
            f.refine((r.uniform(size=f.nquad) < 0.4).astype(np.int32))
            f.balance()
        out.append((f, f.external()))
    (f16, (ci16, fi16, _, s16)), (f8, (ci8, fi8, _, _)) = out
    assert f16.quads() == f8.quads() and np.array_equal(fi16, fi8)
    s16[:] = rng.uniform(0.2, 2.0, s16.shape)
    X16 = oracle.tagger_features(16, ci16, fi16, s16)
    X8 = oracle.tagger_features(8, ci8, fi8, pool16(s16))
    assert np.array_equal(X16, X8)


def test_features_16_constant_and_linear(oracle):
    f = amrhip.Forest(1, 1, (0, 0), 0.0, 0.0, 1.0, 16, 1, 1)
    ci, fi, geom, soln = f.external()
    q = soln.reshape(4, 16, 16, 4)
    ii = np.arange(16)[None, :]
    q[:] = 0.0
    q[..., 0] = 2.0
    q[..., 1] = 0.5 * ii                   # pooled: I + 0.25, max 7.25
    q[..., 2] = -1.0
    q[..., 3] = 4.0
    X = oracle.tagger_features(16, ci, fi, soln).reshape(4, 6, 8, 8).astype(np.uint32)
    one = bf16_bits(np.float32(1.0))
    assert np.all(X[:, 0] == one) and np.all(X[:, 3] == one)
    assert np.all(X[:, 2] == bf16_bits(np.float32(-1.0)))
    exp1 = bf16_bits(((np.arange(8) + 0.25) / 7.25).astype(np.float32))
    assert np.all(X[:, 1] == exp1[None, None, :])
    assert np.all(X[:, 4] == 0) and np.all(X[:, 5] == 0)
    # linear density along y: interior pooled differences are 4 fine steps
    q[..., 0] = 1.0 + 0.125 * np.arange(16)[:, None]     # exact in binary; pooled 1 + 0.125 (2J + 0.5)
    X = oracle.tagger_features(16, ci, fi, soln).reshape(4, 6, 8, 8)
    s0 = 1.0 + 0.125 * 14.5                              # the largest pooled density (J = 7)
    assert np.all(X[:, 5, 1:7, :] == bf16_bits(np.float32(0.5 / s0)))
    assert np.all(X[:, 4] == 0)
