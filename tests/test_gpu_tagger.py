"""Learned tagger on the GPU, through the C ABI: patch features bitwise equal
to the oracle, the BF16-MFMA forward and backward passes against the oracle
and against a plain PyTorch fp32 reference, determinism, training, snapshot
collection and deployment in the device evolve loop."""
import numpy as np
import pytest

import amrhip
from test_tagger_oracle import BLOCKS, bf16_bits, bf16_val, random_problem

pytestmark = pytest.mark.gpu


def adaptive_forest(seed=3, lmin=2, lmax=4, periodic=(1, 0)):
    """8x8 patches, hanging faces, a periodic pair and two walls"""
    rng = np.random.default_rng(seed)
    f = amrhip.Forest(2, 1, periodic, 0.0, 0.0, 0.5, 8, lmin, lmax)
    for _ in range(3):
        f.refine((rng.uniform(size=f.nquad) < 0.3).astype(np.int32))
        f.balance()
    s = f.soln()
    s[..., 0] = rng.uniform(0.2, 2.0, s.shape[:-1])
    s[..., 1] = rng.uniform(-1.0, 1.0, s.shape[:-1])
    s[..., 2] = rng.uniform(-1.0, 1.0, s.shape[:-1])
    s[..., 3] = rng.uniform(1.0, 4.0, s.shape[:-1])
    return f


def gpu_of(f, scheme=3):
    ci, fi, geom, soln = f.external()
    g = amrhip.Gpu()
    g.setup(8, scheme, [amrhip.make_inst(0.5, (0, 0, 0, 0), cfl=0.3, dt_cap=1.0)], ci, fi, geom, soln)
    return g, (ci, fi, geom, soln)


@pytest.mark.parametrize("seed", [3, 11])
def test_features_bitwise(oracle, seed):
    f = adaptive_forest(seed)
    g, (ci, fi, geom, soln) = gpu_of(f)
    got = g.tagger_features()
    exp = oracle.tagger_features(8, ci, fi, soln)
    print(f"features seed={seed} nquad={f.nquad}", flush=True)
    assert got.shape == exp.shape
    assert np.array_equal(got, exp)
    g.close()


def test_features_bitwise_16(oracle):
    """16x16 patches: the features of the patches coarsened to 8x8, bitwise
    equal to the oracle (hanging faces, a periodic pair, walls)"""
    rng = np.random.default_rng(5)
    f = amrhip.Forest(2, 1, (1, 0), 0.0, 0.0, 0.5, 16, 1, 3)
    for _ in range(2):
        f.refine((rng.uniform(size=f.nquad) < 0.35).astype(np.int32))
        f.balance()
    s = f.soln()
    s[..., 0] = rng.uniform(0.2, 2.0, s.shape[:-1])
    s[..., 1:3] = rng.uniform(-1.0, 1.0, s.shape[:-1] + (2,))
    s[..., 3] = rng.uniform(1.0, 4.0, s.shape[:-1])
    ci, fi, geom, soln = f.external()
    g = amrhip.Gpu()
    g.setup(16, 3, [amrhip.make_inst(0.5, (0, 0, 0, 0), cfl=0.3, dt_cap=1.0)], ci, fi, geom, soln)
    got = g.tagger_features()
    exp = oracle.tagger_features(16, ci, fi, soln)
    print(f"16x16 features nquad={f.nquad}", flush=True)
    assert np.array_equal(got, exp)
    g.close()


def torch_fp32_forward(P, X):
    torch = pytest.importorskip("torch")
    x = torch.tensor(bf16_val(X))
    W1 = torch.tensor(P[:64 * 384].reshape(64, 384))
    o = 64 * 384
    b1 = torch.tensor(P[o:o + 64]); o += 64
    W2 = torch.tensor(P[o:o + 4096].reshape(64, 64)); o += 4096
    b2 = torch.tensor(P[o:o + 64]); o += 64
    W3 = torch.tensor(P[o:o + 192].reshape(3, 64)); o += 192
    b3 = torch.tensor(P[o:o + 3])
    h1 = torch.relu(x @ W1.T + b1)
    h2 = torch.relu(h1 @ W2.T + b2)
    return (h2 @ W3.T + b3).numpy()


@pytest.mark.parametrize("n", [1, 64, 100, 1000])
def test_forward_matches_oracle_and_torch(oracle, n):
    P, X, _, _ = random_problem(20 + n, n=n)
    m = amrhip.Mlp(0)
    m.set_params(P)
    print(f"forward n={n}", flush=True)
    z = m.forward(X).astype(np.float64)
    zo = oracle.mlp_forward(P, X, exact=False)
    scale = np.abs(zo).max() + 1.0
    # FP32 accumulation order differs from the oracle's FP64 sums; a BF16
    # activation can then round the other way, moving a logit by ~1e-3
    assert np.abs(z - zo).max() <= 2e-3 * scale
    assert np.median(np.abs(z - zo)) <= 1e-5 * scale
    zt = torch_fp32_forward(P, X)       # unrounded weights / activations
    assert np.abs(z - zt).max() <= 6e-2 * (np.abs(zt).max() + 1.0)
    m.close()


@pytest.mark.parametrize("n", [64, 100, 257, 2048])
def test_grad_matches_oracle(oracle, n):
    P, X, y, cw = random_problem(40 + n, n=n)
    m = amrhip.Mlp(0)
    m.set_params(P)
    print(f"grad n={n}", flush=True)
    loss, g = m.grad(X, y, cw)
    lo, go = oracle.mlp_grad(P, X, y, cw, exact=False)
    assert abs(loss - lo) <= 1e-4 * abs(lo)
    for name, (a, b) in BLOCKS.items():
        rel = np.linalg.norm(g[a:b] - go[a:b]) / np.linalg.norm(go[a:b])
        assert rel <= 1e-2, (name, rel)
    # bitwise reproducible: fixed reduction order
    loss2, g2 = m.grad(X, y, cw)
    assert loss2 == loss and np.array_equal(g, g2)
    m.close()


def test_grad_matches_torch_fp32_autograd():
    torch = pytest.importorskip("torch")
    P, X, y, cw = random_problem(77, n=512)
    m = amrhip.Mlp(0)
    m.set_params(P)
    loss, g = m.grad(X, y, cw)
    p = torch.tensor(P, requires_grad=True)
    x = torch.tensor(bf16_val(X))
    o = 64 * 384
    h1 = torch.relu(x @ p[:o].view(64, 384).T + p[o:o + 64])
    h2 = torch.relu(h1 @ p[o + 64:o + 4160].view(64, 64).T + p[o + 4160:o + 4224])
    z = h2 @ p[o + 4224:o + 4416].view(3, 64).T + p[o + 4416:]
    yt = torch.tensor(y, dtype=torch.long)
    w = torch.tensor(cw)[yt]
    L = (w * torch.nn.functional.cross_entropy(z, yt, reduction="none")).sum() / w.sum()
    L.backward()
    tg = p.grad.numpy()
    assert abs(loss - float(L.detach())) <= 2e-2 * float(L.detach())
    for name, (a, b) in BLOCKS.items():
        cos = np.dot(g[a:b], tg[a:b]) / (np.linalg.norm(g[a:b]) * np.linalg.norm(tg[a:b]))
        assert cos > 0.98, (name, cos)
    m.close()


def test_training_learns_separable_classes():
    """three classes told apart by the sign pattern of two features"""
    rng = np.random.default_rng(5)
    n = 8192
    Xf = rng.normal(0, 0.5, size=(n, 384)).astype(np.float32)
    y = np.where(Xf[:, 3] > 0.25, 2, np.where(Xf[:, 200] < -0.1, 0, 1)).astype(np.int32)
    X = bf16_bits(Xf)
    ts = amrhip.Tagset()
    ts.append(X, y, np.zeros(n))
    m = amrhip.Mlp(1)
    losses = m.train(ts, epochs=30, batch=1024, lr=3e-3, seed=2)
    print("losses", losses[[0, 9, 19, 29]], flush=True)
    assert losses[-1] < 0.3 * losses[0]
    pred = m.predict(ts)
    assert np.mean(pred == y) > 0.95
    # predict_set agrees with argmax of the forward logits
    z = m.forward(X[:500])
    assert np.array_equal(pred[:500], np.argmax(z, axis=1))
    ts.close()
    m.close()


def test_tagset_round_trip():
    rng = np.random.default_rng(8)
    X = rng.integers(0, 65535, size=(300, 384), dtype=np.uint16)
    y = rng.integers(0, 3, size=300).astype(np.int32)
    ind = rng.uniform(size=300)
    a, b = amrhip.Tagset(), amrhip.Tagset()
    a.append(X[:200], y[:200], ind[:200])
    a.append(X[200:], y[200:], ind[200:])
    b.append_from(a, 50, 100)
    Xa, ya, ia = a.get()
    Xb, yb, ib = b.get()
    assert np.array_equal(Xa, X) and np.array_equal(ya, y) and np.array_equal(ia, ind)
    assert np.array_equal(Xb, X[50:150]) and np.array_equal(yb, y[50:150])


def riemann_forest(seed, lmin=2, lmax=5, t_end=0.05):
    amrhip.set_params(nelem_subgrid=8, time_scheme=3, time_steps=10 ** 9, dt=1.0, cfl=0.3,
                      coarse_grid_h=1.0, regrid_interval=10, checkpoint_interval=2 ** 31 - 1,
                      visualization_interval=2 ** 31 - 1, bc=(0, 0, 0, 0, 1, 1), t_end=t_end, verbose=0)
    f = amrhip.Forest(1, 1, (0, 0), 0.0, 0.0, 1.0, 8, lmin, lmax)
    ic = amrhip.riemann4_ic(seed)
    f.apply_ic(ic)
    f.regrid_initial(ic, 0.1)
    return f


def test_collect_snapshots(oracle):
    f = riemann_forest(0)
    ci, fi, geom, soln = f.external()
    X0 = oracle.tagger_features(8, ci, fi, soln)
    ind0 = f.indicator()
    nq0 = f.nquad
    ts = amrhip.Tagset()
    rows = f.collect(ts, tol=0.1, window=20, label_stride=5)
    print(f"collect rows={rows} first={nq0}", flush=True)
    assert rows == ts.size and rows > nq0
    X, y, ind = ts.get()
    # the first window's rows are the initial forest's features and indicator
    assert np.array_equal(X[:nq0], X0)
    assert np.allclose(ind[:nq0], ind0, rtol=1e-14, atol=0)
    assert set(np.unique(y)) <= {0, 1, 2} and np.any(y == 2)
    # quadrants the indicator tags now mostly stay tagged 5-20 steps later
    now = ind > 0.1
    assert np.mean(y[now] == 2) > 0.5
    ts.close()


def test_predict_ctx_matches_features_path():
    f = adaptive_forest(9)
    g, _ = gpu_of(f)
    m = amrhip.Mlp(4)
    cls = m.predict_ctx(g)
    z = m.forward(g.tagger_features())
    assert np.array_equal(cls, np.argmax(z, axis=1))
    g.close()
    m.close()


def test_deployed_tagger_keeps_plain_fallback():
    f = riemann_forest(1, t_end=0.03)
    m = amrhip.Mlp(2)                   # untrained: the fallback must still tag the fronts
    tg = amrhip.MlpTagger(m, tol=0.1, union_plain=True)
    st = f.evolve_device(0.1, 10, tagger=tg)
    s = tg.stats()
    print("deploy", st.steps, st.regrids, s, flush=True)
    assert st.regrids > 0 and st.t == pytest.approx(0.03, abs=1e-12)
    assert sum(s["predicted"]) > 0 and s["plain_refine"] > 0
    assert 0 <= s["added_by_plain"] <= s["plain_refine"]
    assert f.is_balanced()
    m.close()


def test_trained_tagger_beats_plain_indicator_at_the_same_interval():
    """the gain the network brings: trained on solver snapshots of three
    seeds, deployed every 40 steps on a held-out seed, its run ends closer
    to the uniform fine solution (L1 density) than the plain indicator
    regridding at the same interval, and it predicts held-out future refine
    flags at least as well as the indicator evaluated now at an equal refine
    fraction"""
    from amrhip import learned as L
    ts = L.collect([1, 2, 3], lmin=3, lmax=6, t_end=0.1, window=40)
    held = L.collect([100], lmin=3, lmax=6, t_end=0.1, window=40)
    m, losses = L.train(ts, epochs=20)
    print("tagger losses", losses[0], losses[-1], "rows", ts.size, flush=True)
    assert losses[-1] < 0.6 * losses[0]
    eq = L.recall_at_equal_fraction(m, held)
    print("recall at equal refine fraction", eq, flush=True)
    assert eq["recall_mlp"] >= eq["recall_indicator_now"] - 0.02
    ref, _ = L.uniform_reference(0, 6, 0.1)
    plain, _ = L.deploy(0, 3, 6, 0.1, 40, None, 0.1, ref, 6)
    learned, _ = L.deploy(0, 3, 6, 0.1, 40, amrhip.MlpTagger(m, 0.1, True), 0.1, ref, 6)
    print("L1 plain@40", plain["l1_density"], "learned@40", learned["l1_density"], flush=True)
    assert learned["l1_density"] < plain["l1_density"]
    m.close()
    ts.close()
    held.close()
