"""The reference's own tag source on the device (amrgpu_feature_tags,
amr_feature_tagger): solver_tag_feature's density thresholds
(src/solver/src/solver_tag_feature.F90:9-52, tagging_methods_module.F90:14-71)
checked bit for bit against the golden tags the reference's build wrote
(tests/golden/ops.npz) and against the oracle's orc_tag_feature, and run
inside the device evolve loop against the same loop tagged by the oracle."""
import ctypes as C
import os

import numpy as np
import pytest

import amrhip
from amrhip import dpt, ipt
from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def oracle_tags(oracle, methods, tols, n, soln):
    m = np.ascontiguousarray(methods, np.int32)
    t = np.ascontiguousarray(tols, np.float64)
    q = np.ascontiguousarray(soln, np.float64).reshape(-1, 4 * n * n)
    return np.array([oracle.L.orc_tag_feature(m.size, ipt(m), dpt(t), n, dpt(np.ascontiguousarray(row)))
                     for row in q], np.int32)


def uniform_gpu(n, level, soln):
    f = amrhip.Forest(1, 1, (1, 1), 0.0, 0.0, 1.0, n, level, level)
    ci, fi, geom, _ = f.external()
    g = amrhip.Gpu()
    g.setup(n, 3, [amrhip.make_inst(1.0)], ci, fi, geom, np.ascontiguousarray(soln, np.float64))
    return f, g


def test_golden_tags():
    """the 40 golden cases (4x4 patches, one or two methods each, method 0
    as padding tags nothing, as the reference's `case default`)"""
    d = np.load(os.path.join(GOLDEN, "ops.npz"))
    q = d["tag_q"]
    nq = 64
    soln = np.concatenate([q, q[:nq - len(q)]]).reshape(-1)
    f, g = uniform_gpu(4, 3, soln)
    assert f.nquad == nq
    for k in range(len(d["tags"])):
        got = g.feature_tags(d["tag_methods"][k], d["tag_tols"][k])
        assert got[k] == int(d["tags"][k]), f"case {k}"
    g.close()


@pytest.mark.parametrize("n", [2, 4, 8, 16])
def test_tags_equal_oracle(oracle, n):
    """thresholds straddle the patch densities so the tags are mixed;
    densities equal to the threshold (<= and >= are inclusive) and NaN
    densities (which satisfy neither comparison) included"""
    rng = np.random.default_rng(n)
    level = {2: 4, 4: 3, 8: 3, 16: 2}[n]
    nq = 4 ** level
    s = np.empty((nq, n * n, 4))
    s[..., 0] = rng.uniform(0.5, 1.5, (nq, n * n))
    s[..., 1:3] = rng.uniform(-1, 1, (nq, n * n, 2))
    s[..., 3] = rng.uniform(2, 3, (nq, n * n))
    s[3, :, 0] = np.nan                       # an all-NaN patch
    s[5, 1, 0] = np.nan                       # one NaN among numbers
    f, g = uniform_gpu(n, level, s.reshape(-1))
    thr = 1.0
    s[7, 2, 0] = thr                          # exactly at the threshold
    s[9, :, 0] = 1.2
    s[9, 0, 0] = thr
    g.upload(s.reshape(-1))
    rho = s[..., 0]
    finite = ~np.isnan(rho).all(axis=1)
    lo = float(np.median(np.nanmin(rho[finite], axis=1)))     # about half the patches below
    hi = float(np.median(np.nanmax(rho[finite], axis=1)))     # about half above
    cases = [([1], [lo]), ([2], [hi]), ([1], [thr]), ([2], [thr]), ([1, 2], [lo, hi]),
             ([2, 1], [hi, lo]), ([3, 4], [0.0, 0.0]), ([], [])]
    for methods, tols in cases:
        got = g.feature_tags(methods, tols)
        exp = oracle_tags(oracle, methods, tols, n, s.reshape(-1))
        print(f"n={n} methods={methods} tols={tols}: {int(got.sum())} of {nq} tagged", flush=True)
        assert np.array_equal(got, exp)
        if len(methods) == 1 and tols[0] in (lo, hi):
            assert 0 < got.sum() < nq
    assert g.feature_tags([1], [thr])[7] == 1 and g.feature_tags([2], [thr])[9] == 1
    assert g.feature_tags([1], [2.0])[3] == 0 and g.feature_tags([2], [0.0])[3] == 0
    g.close()


def test_rejects_too_many_methods():
    f, g = uniform_gpu(8, 1, np.ones(4 * 64 * 4))
    with pytest.raises(amrhip.GpuError):
        g.feature_tags([1] * 9, [1.0] * 9)
    g.close()


def riemann(seed, lmax, n=8):
    ic = amrhip.riemann4_ic(seed)
    f = amrhip.Forest(1, 1, (0, 0), 0.0, 0.0, 1.0, n, 2, lmax)
    f.apply_ic(ic)
    f.regrid_initial(ic, 0.1)
    return f


def params(t_end):
    amrhip.set_params(nelem_subgrid=8, time_scheme=3, time_steps=10 ** 9, dt=1.0, cfl=0.3,
                      coarse_grid_h=1.0, regrid_interval=10, checkpoint_interval=2 ** 31 - 1,
                      visualization_interval=2 ** 31 - 1, bc=(0, 0, 0, 0, 1, 1), t_end=t_end, verbose=0)


def test_feature_tagger_in_device_loop(oracle):
    """the device loop tagged by amr_feature_tagger equals the same loop
    tagged by the oracle from the downloaded state, bit for bit: same
    regrids, mesh and state"""
    L = amrhip.lib()
    methods, tols = [2, 1], [1.05, 0.35]
    params(0.05)
    fa = riemann(4, 5)
    sa = fa.evolve_device(0.1, 10, tagger=amrhip.FeatureTagger(methods, tols))

    calls = []

    def cb(user, ctx, forest, tags):
        ncell = L.amrgpu_num_cells(ctx)
        s = np.empty(ncell * 4)
        rc = L.amrgpu_download_soln(ctx, dpt(s))
        if rc:
            return rc
        t = oracle_tags(oracle, methods, tols, 8, s)
        C.memmove(tags, t.ctypes.data, t.nbytes)
        calls.append(int(t.sum()))
        return 0

    fn = amrhip.TAGGER_FN(cb)
    params(0.05)
    fb = riemann(4, 5)
    sb = fb.evolve_device(0.1, 10, tagger=fn)
    print(f"feature-tagged loop: {sa.steps} steps, {sa.regrids} regrids, {fa.nquad} quads; "
          f"oracle-tagged {sb.steps} / {sb.regrids} / {fb.nquad}; tags per regrid {calls[:6]}", flush=True)
    assert sa.regrids == sb.regrids == len(calls) > 0
    assert any(0 < c for c in calls)
    assert sa.steps == sb.steps
    assert fa.quads() == fb.quads()
    assert np.array_equal(fa.soln(), fb.soln())
