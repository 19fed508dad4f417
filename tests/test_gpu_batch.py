"""Many independent adaptive problems in one device context
(amr_evolve_batch): each instance must evolve exactly as it does alone under
amr_evolve_device with the same settings -- same steps, regrids, final mesh
and state bit for bit -- although the instances have different sizes, time
steps, regrid outcomes and finish at different step counts."""
import numpy as np
import pytest

import amrhip

pytestmark = pytest.mark.gpu

T_END = 0.04


def params(t_end=T_END):
    amrhip.set_params(nelem_subgrid=8, time_scheme=3, time_steps=10 ** 9, dt=1.0, cfl=0.3,
                      coarse_grid_h=1.0, regrid_interval=10, checkpoint_interval=2 ** 31 - 1,
                      visualization_interval=2 ** 31 - 1, bc=(0, 0, 0, 0, 1, 1), t_end=t_end, verbose=0)


def forest(seed, lmax, lmin=2):
    ic = amrhip.riemann4_ic(seed)
    f = amrhip.Forest(1, 1, (0, 0), 0.0, 0.0, 1.0, 8, lmin, lmax)
    f.apply_ic(ic)
    f.regrid_initial(ic, 0.1)
    return f


CASES = [(0, 4), (1, 5), (2, 5), (3, 6)]


@pytest.mark.parametrize("allow_coarsening,lag", [(True, 0), (False, 0), (True, 4)])
def test_batch_equals_each_instance_alone(allow_coarsening, lag):
    """lag > 0: the host regrids while the GPU advances `lag` more steps
    (alone and batched alike)"""
    params()
    batch = [forest(s, lm) for s, lm in CASES]
    insts = [amrhip.make_inst(1.0, (0, 0, 0, 0), 0.3, 1.0, T_END) for _ in CASES]
    print(f"batch sizes {[f.nquad for f in batch]} lag {lag}", flush=True)
    per, tot = amrhip.evolve_batch(batch, insts, 0.1, 10, allow_coarsening=allow_coarsening, regrid_lag=lag)
    assert tot.regrids > 0
    steps = []
    for (s, lm), fb, st in zip(CASES, batch, per):
        fa = forest(s, lm)
        sa = fa.evolve_device(0.1, 10, allow_coarsening=allow_coarsening, regrid_lag=lag)
        print(f"seed {s}: alone {sa.steps} steps {sa.regrids} regrids; batch {st.steps} / {st.regrids}",
              flush=True)
        assert st.t == sa.t == pytest.approx(T_END, abs=0)
        assert st.regrids == sa.regrids
        assert fb.quads() == fa.quads()
        assert np.array_equal(fb.soln(), fa.soln())
        steps.append(sa.steps)
    # the instances finish at different step counts: the batch really froze some
    assert len(set(steps)) > 1
    # the batch regrid ran its first worker on this thread: the thread's own
    # host-parallel setting is as it was
    assert amrhip.lib().amr_forest_host_parallel_enabled() == 1
    assert tot.steps == max(steps) or tot.steps - max(steps) < 16


@pytest.mark.parametrize("groups", [2, 3])
def test_grouped_batch_equals_one_context(groups):
    """groups > 1: the instances split over contexts advanced on their own
    host threads give every instance the steps, regrids, mesh and state of
    the single-context batch, bit for bit"""
    params()
    one = [forest(s, lm) for s, lm in CASES]
    grouped = [forest(s, lm) for s, lm in CASES]
    insts = [amrhip.make_inst(1.0, (0, 0, 0, 0), 0.3, 1.0, T_END) for _ in CASES]
    print(f"groups {groups}, sizes {[f.nquad for f in one]}", flush=True)
    per1, tot1 = amrhip.evolve_batch(one, insts, 0.1, 10)
    perg, totg = amrhip.evolve_batch(grouped, insts, 0.1, 10, groups=groups)
    for fa, fb, sa, sb in zip(one, grouped, per1, perg):
        assert (sb.steps, sb.regrids, sb.t) == (sa.steps, sa.regrids, sa.t)
        assert sb.cell_stages == sa.cell_stages
        assert fb.quads() == fa.quads()
        assert np.array_equal(fb.soln(), fa.soln())
    assert totg.steps == tot1.steps or tot1.steps - totg.steps < 16
    assert totg.regrids >= tot1.regrids


def test_batch_with_learned_tagger_runs():
    params()
    batch = [forest(s, lm) for s, lm in CASES[:3]]
    insts = [amrhip.make_inst(1.0, (0, 0, 0, 0), 0.3, 1.0, T_END) for _ in batch]
    m = amrhip.Mlp(3)
    tg = amrhip.MlpTagger(m, 0.1, union_plain=True)
    per, tot = amrhip.evolve_batch(batch, insts, 0.1, 20, tagger=tg)
    s = tg.stats()
    # the same tagger shared by two groups (one call at a time)
    batch2 = [forest(s_, lm) for s_, lm in CASES[:3]]
    per2, tot2 = amrhip.evolve_batch(batch2, insts, 0.1, 20, tagger=tg, groups=2)
    assert tot2.regrids > 0
    for f, st in zip(batch2, per2):
        assert st.t == pytest.approx(T_END, abs=0)
        assert f.is_balanced()
    assert tot.regrids > 0 and sum(s["predicted"]) > 0
    for f, st in zip(batch, per):
        assert st.t == pytest.approx(T_END, abs=0)
        assert f.is_balanced()
    m.close()


def test_lagged_regrid_conserves_and_differs_from_immediate():
    """the lagged loop is a different (legitimate) schedule: it conserves
    mass on a periodic box to round-off and ends on a balanced mesh"""
    amrhip.set_params(nelem_subgrid=8, time_scheme=3, time_steps=60, dt=1.0, cfl=0.3, coarse_grid_h=10.0,
                      regrid_interval=10, checkpoint_interval=2 ** 31 - 1, visualization_interval=2 ** 31 - 1,
                      bc=(0, 0, 0, 0, 1, 1), t_end=0.0, verbose=0)
    ic = amrhip.vortex_ic(5.0, 5.0, 5.0)
    out = []
    for lag in (0, 5):
        f = amrhip.Forest(1, 1, (1, 1), 0.0, 0.0, 10.0, 8, 2, 4)
        f.apply_ic(ic)
        f.regrid_initial(ic, 0.02)
        s0 = f.soln().reshape(f.nquad, -1, 4)
        lv0 = f.levels()
        m0 = float(np.sum(s0[:, :, 0].sum(axis=1) * (10.0 / 2.0 ** lv0 / 8) ** 2))
        st = f.evolve_device(0.02, 10, regrid_lag=lag)
        assert st.steps == 60 and f.is_balanced()
        s1 = f.soln().reshape(f.nquad, -1, 4)
        m1 = float(np.sum(s1[:, :, 0].sum(axis=1) * (10.0 / 2.0 ** f.levels() / 8) ** 2))
        assert abs(m1 - m0) <= 1e-11 * abs(m0)
        out.append((f.quads(), s1.copy()))
    assert out[0][0] != out[1][0] or not np.array_equal(out[0][1], out[1][1])
