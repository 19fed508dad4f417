"""The batched workload at its full shape: 64 independent four-quadrant
Riemann problems (seeds 1000..1063, levels 3..lmax with lmax drawn from
{5, 6, 7, 8}, 8x8 FP64 patches, density-gradient regrid every 10 steps,
SSP-RK3, input CFL 0.4) advanced together to t = 0.25 in one device context,
each instance equal to its own run alone bit for bit: steps, regrids, final
mesh and state; and the same batch split into two groups equal to it."""
import time

import numpy as np
import pytest

import amrhip

pytestmark = pytest.mark.gpu

K, LMIN, TOL, T_END = 64, 3, 0.1, 0.25


def instances():
    rng = np.random.default_rng(2024)
    lmaxs = [int(x) for x in rng.choice([5, 6, 7, 8], size=K)]
    return [1000 + k for k in range(K)], lmaxs


def forest(seed, lmax):
    ic = amrhip.riemann4_ic(seed)
    f = amrhip.Forest(1, 1, (0, 0), 0.0, 0.0, 1.0, 8, LMIN, lmax)
    f.apply_ic(ic)
    f.regrid_initial(ic, TOL)
    return f


def params():
    amrhip.set_params(nelem_subgrid=8, time_scheme=3, time_steps=10 ** 9, dt=1.0, cfl=0.25 * 0.4 * 3,
                      coarse_grid_h=1.0, regrid_interval=10, checkpoint_interval=2 ** 31 - 1,
                      visualization_interval=2 ** 31 - 1, bc=(0, 0, 0, 0, 1, 1), t_end=T_END, verbose=0)


def test_64_instances_to_t025_each_equal_to_alone():
    params()
    p = amrhip.get_params()
    seeds, lmaxs = instances()
    assert sorted(set(lmaxs)) == [5, 6, 7, 8]
    batch = [forest(s, lm) for s, lm in zip(seeds, lmaxs)]
    insts = [amrhip.make_inst(1.0, (0, 0, 0, 0), p.cfl, 1.0, T_END) for _ in seeds]
    t0 = time.perf_counter()
    per, tot = amrhip.evolve_batch(batch, insts, TOL, 10)
    print(f"batch of {K}: {tot.steps} steps, {tot.regrids} regrids, peak {tot.peak_quads * 64} cells, "
          f"{time.perf_counter() - t0:.1f} s", flush=True)
    sizes = []
    for k, (s, lm) in enumerate(zip(seeds, lmaxs)):
        fa = forest(s, lm)
        sa = fa.evolve_device(TOL, 10)
        st = per[k]
        sizes.append(st.peak_quads * 64)
        assert st.t == sa.t == T_END, f"instance {k}"
        assert (st.steps, st.regrids) == (sa.steps, sa.regrids), f"instance {k}"
        assert batch[k].quads() == fa.quads(), f"instance {k}"
        assert np.array_equal(batch[k].soln(), fa.soln()), f"instance {k}"
    print(f"peak cells per instance {min(sizes)}..{max(sizes)}", flush=True)
    # uneven sizes: peaks from about 18k to 0.5M cells
    assert min(sizes) < 30_000 and max(sizes) > 400_000
    # the same 64 instances split into two groups (own contexts and host
    # threads, regrids overlapping steps): each instance as in one context
    grouped = [forest(s, lm) for s, lm in zip(seeds, lmaxs)]
    t0 = time.perf_counter()
    perg, totg = amrhip.evolve_batch(grouped, insts, TOL, 10, groups=2)
    print(f"two groups: {time.perf_counter() - t0:.1f} s", flush=True)
    for k in range(K):
        assert (perg[k].steps, perg[k].regrids, perg[k].t) == (per[k].steps, per[k].regrids, per[k].t), k
        assert grouped[k].quads() == batch[k].quads(), f"instance {k}"
        assert np.array_equal(grouped[k].soln(), batch[k].soln()), f"instance {k}"
