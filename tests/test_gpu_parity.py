"""GPU parity: the HIP path through the C ABI against the reference's golden
outputs and against the CPU oracle (which reproduces the reference bit for
bit).  Bars: residual per cell and field within 1e-12 relative to
max(|r_ref|, ||r_ref_field||_inf); CFL time step bit-exact; RK trajectories
within 1e-11 of the oracle's k-form (the GPU uses the algebraically identical
Shu-Osher / low-storage forms); conservation to round-off."""
import math
import os

import numpy as np
import pytest

import amrhip
from conftest import golden_meshes, parity_err

pytestmark = pytest.mark.gpu

RHS_TOL = 1e-12
TRAJ_TOL = 1e-11


def log(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def gpu():
    g = amrhip.Gpu()
    yield g
    g.close()


def setup(g, n, scheme, h_base, bc, cfl, dt_cap, ci, fi, geom, soln, t_end=0.0):
    g.setup(n, scheme, [amrhip.make_inst(h_base, bc, cfl, dt_cap, t_end)], ci, fi, geom, soln)


@pytest.mark.parametrize("path", golden_meshes(), ids=lambda p: os.path.basename(p)[5:-4])
def test_rhs_against_reference_golden(gpu, path):
    m = np.load(path)
    n = int(m["n"])
    log(f"golden {os.path.basename(path)} n={n}")
    setup(gpu, n, 3, float(m["h_base"]), tuple(int(b) for b in m["bc"]), float(m["cfl"]), 1.0,
          m["cell_info"], m["face_info"], m["geom"], m["soln"])
    err = parity_err(gpu.rhs(), m["rhs"])
    assert err <= RHS_TOL, err
    assert gpu.cfl()[0] == float(m["dt"])     # bit-exact CFL minimum


def riemann_forest(n, seed, lmin=2, lmax=5, tol=0.1):
    ic = amrhip.riemann4_ic(seed)
    f = amrhip.Forest(1, 1, (0, 0), 0.0, 0.0, 1.0, n, lmin, lmax)
    f.apply_ic(ic)
    f.regrid_initial(ic, tol)
    return f


def vortex_forest(n, level, extra=True):
    f = amrhip.Forest(1, 1, (1, 1), 0.0, 0.0, 10.0, n, level, level + 1)
    if extra:
        flags = np.zeros(f.nquad, np.int32)
        flags[f.nquad // 3] = 1
        flags[f.nquad // 2] = 1
        f.refine(flags)
        f.balance()
    f.apply_ic(amrhip.vortex_ic(5.0, 5.0, 5.0))
    return f


@pytest.mark.parametrize("n", [2, 4, 8, 16])
@pytest.mark.parametrize("bc", [(0, 0, 0, 0), (1, 2, 1, 0)])
def test_rhs_adaptive_forest(gpu, oracle, n, bc):
    log(f"adaptive riemann n={n} bc={bc}")
    f = riemann_forest(n, seed=n, lmax=4 if n == 16 else 5)
    ci, fi, geom, soln = f.external()
    assert np.any(fi.reshape(-1, 11)[:, 0] == 3)
    setup(gpu, n, 3, 1.0, bc, 0.3, 1.0, ci, fi, geom, soln)
    ref = oracle.rhs(oracle.params(n, bc=bc, h_base=1.0), ci, fi, geom, soln)
    assert parity_err(gpu.rhs(), ref) <= RHS_TOL


@pytest.mark.parametrize("n", [2, 4, 8, 16])
def test_every_tile_variant(gpu, oracle, n):
    """every stage-kernel variant the autotuner may pick meets the bar, on a
    mesh with hanging faces and all three wall types"""
    f = riemann_forest(n, seed=5 + n, lmax=4 if n == 16 else 5)
    ci, fi, geom, soln = f.external()
    bc = (2, 1, 0, 1)
    setup(gpu, n, 3, 1.0, bc, 0.3, 1.0, ci, fi, geom, soln)
    ref = oracle.rhs(oracle.params(n, bc=bc, h_base=1.0), ci, fi, geom, soln)
    p = oracle.params(n, bc=bc, cfl=0.3, dt=1.0, h_base=1.0, scheme=3)
    q_ref, _, _ = oracle.run(p, ci, fi, geom, soln, 3)
    variants = amrhip.tile_variants(n)
    assert variants
    for ppb, minw, design in variants:
        log(f"variant n={n} ppb={ppb} minw={minw} design={design}")
        gpu.setup(n, 3, [amrhip.make_inst(1.0, bc, 0.3, 1.0)], ci, fi, geom, soln)
        gpu.set_tile(ppb, minw, design)
        assert parity_err(gpu.rhs(), ref) <= RHS_TOL
        gpu.run(3)
        assert parity_err(gpu.soln(), q_ref) <= TRAJ_TOL


@pytest.mark.parametrize("n", [2, 4, 8, 16])
def test_tile_variants_agree_bit_for_bit(gpu, n):
    """every compiled tile of a patch size computes the same bits: the
    kernels fuse only the multiply-adds their source writes as fma, so the
    tile an autotuner picks does not change a run's results (a batch equals
    its instances run one by one whatever tiles each run picked); residual
    and 5 SSP-RK3 steps on an adaptive forest with hanging faces, walls of
    every reference type and a copy side"""
    f = riemann_forest(n, seed=7 + n, lmax=4 if n == 16 else 5)
    ci, fi, geom, soln = f.external()
    bc = (2, 1, 0, 1)
    out = {}
    for v in amrhip.tile_variants(n):
        log(f"tile n={n} {tuple(v)}")
        gpu.setup(n, 3, [amrhip.make_inst(1.0, bc, 0.3, 1.0)], ci, fi, geom, soln)
        gpu.set_tile(*v)
        r = gpu.rhs()
        gpu.run(5)
        out[tuple(v)] = (r, gpu.soln(), gpu.time()[0][0])
    first, (r0, q0, t0) = next(iter(out.items()))
    for tile, (r, q, t) in out.items():
        assert np.array_equal(r, r0), (first, tile)
        assert np.array_equal(q, q0), (first, tile)
        assert t == t0, (first, tile)


def test_autotune_restores_state_and_keeps_parity(gpu, oracle):
    """the autotuner's trial steps (every tile, then both launch modes) leave
    state and time as they were; the tile and mode it keeps meet the bar"""
    f = riemann_forest(8, seed=11, lmax=5)
    ci, fi, geom, soln = f.external()
    bc = (0, 0, 1, 1)
    setup(gpu, 8, 3, 1.0, bc, 0.3, 1.0, ci, fi, geom, soln)
    log(f"autotune on {f.nquad} quads")
    res = gpu.autotune(3)
    variants = amrhip.tile_variants(8)
    assert len(res) == len(variants) and all(ms > 0 for *_, ms in res)
    assert gpu.tile() in [tuple(v) for v in variants]
    assert gpu.graphs() in (True, False)
    t, _ = gpu.time()
    assert t[0] == 0.0
    assert np.array_equal(gpu.soln(), soln)
    gpu.run(3)
    p = oracle.params(8, bc=bc, cfl=0.3, dt=1.0, h_base=1.0, scheme=3)
    q_ref, _, _ = oracle.run(p, ci, fi, geom, soln, 3)
    assert parity_err(gpu.soln(), q_ref) <= TRAJ_TOL


@pytest.mark.parametrize("n", [2, 4, 8, 16])
@pytest.mark.parametrize("periodic", [True, False])
def test_single_quadrant(gpu, oracle, n, periodic):
    """the smallest mesh: one quadrant, whose neighbour across every periodic
    side is itself (or four walls), in a workgroup whose other patch slots
    are empty; residual, bit-exact CFL and 5 SSP-RK3 steps against the
    oracle.  Random states (density and pressure in [0.5, 1.5], velocities
    in [-0.3, 0.3]): a symmetric initial condition would leave whole
    residual fields at round-off, where a per-field relative bar is void"""
    per = (1, 1) if periodic else (0, 0)
    bc = (0, 0, 0, 0) if periodic else (1, 2, 0, 1)
    f = amrhip.Forest(1, 1, per, 0.0, 0.0, 10.0, n, 0, 0)
    rng = np.random.default_rng(40 + n)
    q = f.soln().reshape(-1, 4)
    rho, u, v, pr = (rng.uniform(0.5, 1.5, len(q)), rng.uniform(-0.3, 0.3, len(q)),
                     rng.uniform(-0.3, 0.3, len(q)), rng.uniform(0.5, 1.5, len(q)))
    q[:] = np.stack([rho, rho * u, rho * v, pr / 0.4 + 0.5 * rho * (u * u + v * v)], axis=1)
    ci, fi, geom, soln = f.external()
    assert f.nquad == 1
    setup(gpu, n, 3, 10.0, bc, 0.3, 1.0, ci, fi, geom, soln)
    ref = oracle.rhs(oracle.params(n, bc=bc, h_base=10.0), ci, fi, geom, soln)
    assert parity_err(gpu.rhs(), ref) <= RHS_TOL
    assert gpu.cfl()[0] == oracle.time_step(oracle.params(n, bc=bc, cfl=0.3, h_base=10.0), ci, geom, soln, 1.0)
    gpu.run(5)
    p = oracle.params(n, bc=bc, cfl=0.3, dt=1.0, h_base=10.0, scheme=3)
    q_ref, _, _ = oracle.run(p, ci, fi, geom, soln, 5)
    assert parity_err(gpu.soln(), q_ref) <= TRAJ_TOL


@pytest.mark.parametrize("n", [4, 8, 16])
def test_nan_residual_is_reported(gpu, n):
    """a non-physical state (negative density in one cell: sqrt(rho) of the
    Roe averages is NaN here as in the reference) turns the residual into
    NaN: the run reports AMRGPU_ERR_NAN (3), where the reference aborts
    (physics_runge_kutta_interface.c:352-355); the context stays usable for
    a valid state.  (A negative energy alone need not: the Roe-averaged
    enthalpy, and with it the sound speed, can stay real.)"""
    f = vortex_forest(n, 2)
    ci, fi, geom, soln = f.external()
    bad = np.array(soln, copy=True).reshape(-1, 4)
    bad[5, 0] = -1.0
    setup(gpu, n, 3, 10.0, (0, 0, 0, 0), 0.3, 1.0, ci, fi, geom, bad.reshape(-1))
    with pytest.raises(FloatingPointError):
        gpu.run(2)
    assert b"nan" in amrhip.lib().amrgpu_last_error()
    gpu.upload(soln)
    gpu.set_time([0.0])
    assert math.isfinite(gpu.run(2))


def test_rhs_config1_vortex(gpu, oracle):
    """the isentropic vortex on the uniform level-4 forest with 8x8 patches"""
    log("vortex level 4 n=8")
    f = vortex_forest(8, 4, extra=False)
    ci, fi, geom, soln = f.external()
    assert f.nquad == 256
    setup(gpu, 8, 3, 10.0, (0, 0, 0, 0), 0.3, 1.0, ci, fi, geom, soln)
    ref = oracle.rhs(oracle.params(8, h_base=10.0), ci, fi, geom, soln)
    assert parity_err(gpu.rhs(), ref) <= RHS_TOL
    p = oracle.params(8, cfl=0.3, h_base=10.0)
    assert gpu.cfl()[0] == oracle.time_step(p, ci, geom, soln, np.inf)


@pytest.mark.parametrize("n", [4, 8, 16])
@pytest.mark.parametrize("scheme", [1, 2, 3, 4, 5])
def test_rk_trajectory(gpu, oracle, scheme, n):
    """every RK scheme's stage plan (Shu-Osher two-register stages, the RK4
    accumulator, LSERK45's 2N register) on a forest with hanging faces, for
    the patch sizes whose kernels differ (several patches per wave, one, or
    several waves per patch)"""
    log(f"trajectory scheme {scheme} n {n}")
    f = vortex_forest(n, 2)
    ci, fi, geom, soln = f.external()
    cfl = 0.25 * 0.4 * scheme
    dt_cap = 1.0 if scheme != 2 else 0.01
    setup(gpu, n, scheme, 10.0, (0, 0, 0, 0), cfl, dt_cap, ci, fi, geom, soln)
    nsteps = 10
    rn = gpu.run(nsteps)
    q_gpu = gpu.soln()
    p = oracle.params(n, cfl=cfl, dt=dt_cap, h_base=10.0, scheme=scheme)
    q_ref, rn_ref, dt_ref = oracle.run(p, ci, fi, geom, soln, nsteps)
    err = parity_err(q_gpu, q_ref)
    assert err <= TRAJ_TOL, err
    t, dt = gpu.time()
    assert dt[0] == pytest.approx(dt_ref, rel=1e-12)
    # the residual norm of the run's last stage (physics_norm_L2_residual
    # after the SUM allreduce, then sqrt): the fused block sums of the GPU
    # against the oracle's running sum
    assert rn == pytest.approx(rn_ref, rel=1e-10), (rn, rn_ref)


@pytest.mark.parametrize("n", [4, 8, 16])
@pytest.mark.parametrize("scheme", [3, 4, 5])
def test_graph_and_direct_launch_identical(gpu, n, scheme):
    """replays of the captured step graphs equal direct launches bit for bit,
    checked after every replay: each run of 2 steps replays one captured
    step, then launches the last (norm) step directly, with the run's own
    eager work (CFL priming, norm download, synchronisation) between
    replays"""
    f = riemann_forest(n, seed=3, lmax=4 if n < 16 else 3)
    ci, fi, geom, soln = f.external()
    setup(gpu, n, scheme, 1.0, (0, 1, 2, 0), 0.3, 1.0, ci, fi, geom, soln)
    assert not gpu.graphs()          # direct launches unless asked otherwise
    direct = []
    for _ in range(5):
        gpu.run(2, graph=True)       # graph requested, replay disabled: direct
        direct.append(gpu.soln())
    gpu.upload(soln)
    gpu.set_time([0.0])
    gpu.set_graphs(True)
    for k in range(5):
        gpu.run(2, graph=True)       # one replay of a captured step per run
        assert np.array_equal(gpu.soln(), direct[k]), k
    gpu.set_graphs(False)


@pytest.mark.parametrize("n", [2, 4, 8, 16])
def test_conservation_periodic(gpu, n):
    """mass, momentum and energy over 50 SSP-RK3 steps on the periodic box
    with hanging faces, for every patch size"""
    f = vortex_forest(n, 3)
    ci, fi, geom, soln = f.external()
    setup(gpu, n, 3, 10.0, (0, 0, 0, 0), 0.3, 1.0, ci, fi, geom, soln)
    vol = ((10.0 / 2.0 ** f.levels()) / n) ** 2
    def totals(s):
        return np.einsum("q,qcf->f", vol, s.reshape(f.nquad, n * n, 4))
    m0 = totals(soln)
    gpu.run(50)
    m1 = totals(gpu.soln())
    assert np.all(np.abs(m1 - m0) <= 1e-12 * np.abs(m0))


def test_conservation_level9_full_size(gpu):
    """the benchmark's HBM-scale mesh (uniform level 9, 16.8M cells, two
    quadrants refined so that hanging faces are present): mass, momentum and
    energy conserved to round-off over 20 SSP-RK3 steps on the periodic box"""
    n = 8
    f = vortex_forest(n, 9)
    ci, fi, geom, soln = f.external()
    log(f"level 9: {f.nquad} quads, {f.nquad * n * n} cells")
    setup(gpu, n, 3, 10.0, (0, 0, 0, 0), 0.3, 1.0, ci, fi, geom, soln)
    vol = ((10.0 / 2.0 ** f.levels()) / n) ** 2
    def totals(s):
        # correctly rounded sums: a plain sum over 16.8M terms carries a
        # rounding error of order 1e-11 relative, far above the drift tested
        q = s.reshape(f.nquad, n * n, 4) * vol[:, None, None]
        return np.array([math.fsum(q[:, :, k].ravel()) for k in range(4)])
    m0 = totals(soln)
    gpu.run(20)
    m1 = totals(gpu.soln())
    log(f"drift {m1 - m0}")
    assert np.all(np.abs(m1 - m0) <= 1e-12 * np.maximum(np.abs(m0), np.abs(m0).max()))


@pytest.mark.parametrize("n", [4, 8, 16])
def test_batched_instances_match_single(gpu, n):
    """two independent forests in one context, with their own walls, cell
    size and CFL number, give each forest's own result bit for bit (with
    4x4 patches a workgroup holds patches of both instances)"""
    ga = []
    parts = []
    for seed, bc, hb, cfl in ((1, (0, 1, 2, 0), 1.0, 0.3), (2, (1, 1, 1, 1), 2.0, 0.25)):
        f = riemann_forest(n, seed=seed, lmax=4 if n < 16 else 3)
        parts.append((tuple(np.array(a, copy=True) for a in f.external()), bc, hb, cfl))
    for (ci, fi, geom, soln), bc, hb, cfl in parts:
        g = amrhip.Gpu()
        setup(g, n, 3, hb, bc, cfl, 1.0, ci, fi, geom, soln, t_end=0.01)
        g.run(5)
        ga.append((g.soln(), g.time()[0][0]))
        g.close()
    # concatenate
    cis, fis, gs, ss, inst = [], [], [], [], []
    off_c, off_d = 0, 0
    for k, ((ci, fi, geom, soln), *_) in enumerate(parts):
        c = ci.reshape(-1, 3).copy()
        c[:, 2] += off_d
        fr = fi.reshape(-1, 11).copy()
        for col in range(2, 7):
            fr[:, col] = np.where(fr[:, col] >= 0, fr[:, col] + off_c, -1)
        cis.append(c.reshape(-1)); fis.append(fr.reshape(-1)); gs.append(geom); ss.append(soln)
        inst.append(np.full(len(c), k, np.int32))
        off_c += len(c)
        off_d += soln.size
    g = amrhip.Gpu()
    g.setup(n, 3, [amrhip.make_inst(hb, bc, cfl, 1.0, 0.01) for (_, bc, hb, cfl) in parts],
            np.concatenate(cis), np.concatenate(fis), np.concatenate(gs), np.concatenate(ss),
            cell_inst=np.concatenate(inst))
    g.run(5)
    out = g.soln()
    t, _ = g.time()
    g.close()
    a = ga[0][0].size
    assert np.array_equal(out[:a], ga[0][0])
    assert np.array_equal(out[a:], ga[1][0])
    assert t[0] == ga[0][1] and t[1] == ga[1][1]


@pytest.mark.parametrize("n", [4, 8, 16])
@pytest.mark.parametrize("scheme", [1, 2, 3, 4, 5])
def test_physics_go_gpu_interval(oracle, scheme, n):
    """the drop-in entry point: interval semantics and state vs the oracle,
    for every time scheme and patch size (RK2 takes no CFL step: the input
    dt, kept small, is its step)"""
    import ctypes as C
    f = vortex_forest(n, 2)
    ci, fi, geom, soln = f.external()
    dt_in = 0.01 if scheme == 2 else 1.0
    amrhip.set_params(nelem_subgrid=n, time_scheme=scheme, time_steps=12, dt=dt_in, cfl=0.3,
                      coarse_grid_h=10.0, regrid_interval=5, checkpoint_interval=2 ** 31 - 1,
                      visualization_interval=2 ** 31 - 1, bc=(0, 0, 0, 0, 1, 1), t_end=0.0)
    L = amrhip.lib()
    q = soln.copy()
    ci32 = np.ascontiguousarray(ci, np.int32)
    fi32 = np.ascontiguousarray(fi, np.int32)
    ints = {k: C.c_int(v) for k, v in dict(dim=2, uns=0, nf=4, n=n, ncell=f.nquad, ngh=0,
                                           nface=fi.size // 11, dof=q.size, ns=0, nr=0).items()}
    istep, rg, vis, chk, ev = (C.c_int(0) for _ in range(5))
    ev.value = 1
    tt, ct, rt = C.c_double(0), C.c_double(0), C.c_double(0)
    dummy = np.zeros(4, np.int32)
    rc = L.physics_go_gpu(*(C.byref(ints[k]) for k in ("dim", "uns", "nf", "n", "ncell", "ngh", "nface",
                                                       "dof")),
                          amrhip.ipt(ci32), amrhip.ipt(fi32), C.byref(ints["ns"]), C.byref(ints["nr"]),
                          amrhip.ipt(dummy), amrhip.ipt(dummy), amrhip.ipt(dummy),
                          amrhip.dpt(geom), amrhip.dpt(q), C.byref(istep), C.byref(rg), C.byref(vis),
                          C.byref(chk), C.byref(ev), None, C.byref(tt), C.byref(ct), C.byref(rt))
    assert rc == 0
    assert istep.value == 5 and rg.value == 1 and ev.value == 1
    p = oracle.params(n, cfl=0.3, dt=dt_in, h_base=10.0, scheme=scheme)
    q_ref, _, _ = oracle.run(p, ci, fi, geom, soln, 5)
    assert parity_err(q, q_ref) <= TRAJ_TOL


def test_device_indicator_matches_host(gpu):
    for n in (4, 8, 16):
        f = riemann_forest(n, seed=11, lmax=4 if n == 16 else 5)
        ci, fi, geom, soln = f.external()
        setup(gpu, n, 3, 1.0, (0, 0, 0, 0), 0.3, 1.0, ci, fi, geom, soln)
        assert np.array_equal(gpu.indicator(), f.indicator())


def test_device_remesh_matches_host_operators(gpu):
    rng = np.random.default_rng(3)
    f = riemann_forest(8, seed=12, lmax=5)
    ci, fi, geom, soln = f.external()
    setup(gpu, 8, 3, 1.0, (0, 0, 0, 0), 0.3, 1.0, ci, fi, geom, soln)
    tags = (rng.uniform(size=f.nquad) < 0.3).astype(np.int32)
    f.track_begin(keep_data=True)
    f.regrid_tags(tags)
    ok, src = f.track_end()
    assert ok and np.any(src < 0)
    ci2, fi2, geom2, soln2 = f.external()
    gpu.remesh(ci2, fi2, src)
    assert np.array_equal(gpu.soln(), soln2)


def test_direct_table_remesh_equals_face_route(gpu):
    """amrgpu_remesh_nbr with the forest's own mesh table (the device loops'
    route) gives the state and residual of amrgpu_remesh with the reference
    face list; tables with an id outside the new mesh, a level or an
    instance out of range are refused and leave the context as it was
    (position walls included: their coordinates stay valid)"""
    rng = np.random.default_rng(5)
    f = dmr_forest(5)
    ci, fi, geom, soln = (np.array(a, copy=True) for a in f.external())
    insts = [amrhip.dmr_inst(1.0, 0.3, 1.0)]
    gpu.setup(8, 3, insts, ci, fi, geom, soln)
    tags = (rng.uniform(size=f.nquad) < 0.3).astype(np.int32)
    f.track_begin(keep_data=True)
    f.regrid_tags(tags)
    ok, src = f.track_end()
    assert ok and np.any(src < 0)
    ci2, fi2, geom2, soln2 = (np.array(a, copy=True) for a in f.external())
    nbr, level, org = f.mesh_table()
    r_before = gpu.rhs()
    bad_nbr = nbr.copy()
    bad_nbr[8 * 3 + 2] = f.nquad                     # an id one past the new mesh
    bad_level = level.copy()
    bad_level[0] = 30
    for args, kw in (((level, bad_nbr, src), {}), ((bad_level, nbr, src), {}),
                     ((level, nbr, src), {"cell_inst": np.full(f.nquad, 1, np.int32)})):
        with pytest.raises(amrhip.GpuError):
            gpu.remesh_nbr(*args, **kw)
    assert np.array_equal(gpu.soln(), soln) and np.array_equal(gpu.rhs(), r_before)
    gpu.remesh_nbr(level, nbr, src)
    with pytest.raises(amrhip.GpuError):             # position walls need the new origins
        gpu.rhs()
    gpu.set_origins(org)
    assert np.array_equal(gpu.soln(), soln2)
    r_tab = gpu.rhs()
    g2 = amrhip.Gpu()
    try:
        g2.setup(8, 3, insts, ci, fi, geom, soln)
        g2.remesh(ci2, fi2, src)
        amrhip.check(amrhip.lib().amrgpu_set_geometry(g2.h, amrhip.dpt(geom2)), "geom")
        assert np.array_equal(g2.rhs(), r_tab)
    finally:
        g2.close()


@pytest.mark.parametrize("allow_coarsening", [True, False])
def test_device_evolve_equals_host_regrid_loop(allow_coarsening):
    """state resident in HBM across regrids gives bit for bit the result of
    the interval-by-interval host path (download, host regrid, upload)"""
    def make():
        return riemann_forest(8, seed=21, lmin=2, lmax=5)
    amrhip.set_params(nelem_subgrid=8, time_scheme=3, time_steps=60, dt=1.0, cfl=0.3,
                      coarse_grid_h=1.0, regrid_interval=10, checkpoint_interval=2 ** 31 - 1,
                      visualization_interval=2 ** 31 - 1, bc=(0, 0, 0, 0, 1, 1), t_end=0.0)
    fa = make()
    sa = fa.evolve(0.1, allow_coarsening)
    fb = make()
    sb = fb.evolve_device(0.1, 10, allow_coarsening)
    assert sa.steps == sb.steps == 60 and sa.regrids == sb.regrids
    assert fa.quads() == fb.quads()
    assert np.array_equal(fa.soln(), fb.soln())


def test_headline_workload_full_size(gpu, oracle):
    """the benchmark's workload at full size: the four-quadrant Riemann
    problem, levels 3-7, 8x8 patches, regrid every 10 steps, to t = 0.25
    (~2200 steps, ~220 regrids, ~207k cells at the end).  The state-resident
    loop equals the interval-by-interval host loop bit for bit over the whole
    run; on the mesh and state reached, the residual meets the 1e-12 bar
    against the oracle and 5 SSP-RK3 steps the trajectory bar"""
    def make():
        return riemann_forest(8, seed=0, lmin=3, lmax=7)
    amrhip.set_params(nelem_subgrid=8, time_scheme=3, time_steps=10 ** 9, dt=1.0, cfl=0.25 * 0.4 * 3,
                      coarse_grid_h=1.0, regrid_interval=10, checkpoint_interval=2 ** 31 - 1,
                      visualization_interval=2 ** 31 - 1, bc=(0, 0, 0, 0, 1, 1), t_end=0.25)
    fb = make()
    sb = fb.evolve_device(0.1, 10)
    log(f"device loop: {sb.steps} steps, {sb.regrids} regrids, {fb.nquad} quads at t={sb.t}")
    fa = make()
    sa = fa.evolve(0.1)
    log(f"host loop: {sa.steps} steps, {sa.regrids} regrids")
    assert sb.t == pytest.approx(0.25, abs=0) and sb.regrids > 200 and fb.nquad * 64 > 150000
    assert sa.steps == sb.steps and sa.regrids == sb.regrids
    assert fa.quads() == fb.quads()
    assert np.array_equal(fa.soln(), fb.soln())
    ci, fi, geom, soln = fb.external()
    p = oracle.params(8, bc=(0, 0, 0, 0), cfl=0.3, dt=1.0, h_base=1.0, scheme=3)
    setup(gpu, 8, 3, 1.0, (0, 0, 0, 0), 0.3, 1.0, ci, fi, geom, soln)
    assert parity_err(gpu.rhs(), oracle.rhs(p, ci, fi, geom, soln)) <= RHS_TOL
    gpu.run(5)
    q_ref, _, dt_ref = oracle.run(p, ci, fi, geom, soln, 5)
    assert parity_err(gpu.soln(), q_ref) <= TRAJ_TOL
    assert gpu.time()[1][0] == pytest.approx(dt_ref, rel=1e-12)


def vortex_l1(f, soln, t):
    """L1 density error per unit area against the exactly advected vortex
    (amr_ic_vortex_density at the sub-cell centres, nearest periodic image)"""
    ic = amrhip.vortex_ic(5.0, 5.0, 5.0)
    L = amrhip.lib()
    n = f.n
    _, _, geom, _ = f.external()
    q = np.asarray(soln).reshape(f.nquad, n, n, 4)
    err, area = 0.0, 0.0
    for k in range(f.nquad):
        g = geom[12 * k:12 * k + 12].reshape(4, 3)
        x0, y0, h = g[0, 0], g[0, 1], (g[1, 0] - g[0, 0]) / n
        for j in range(n):
            for i in range(n):
                ex = L.amr_ic_vortex_density(C_IC(ic), x0 + (i + 0.5) * h, y0 + (j + 0.5) * h, t,
                                             0.0, 10.0, 0.0, 10.0)
                err += abs(q[k, j, i, 0] - ex) * h * h
                area += h * h
    return err / area


def C_IC(ic):
    import ctypes
    return ctypes.byref(ic)


def test_config1_vortex_200_steps_l1(gpu, oracle):
    """config 1: 200 SSP-RK3 steps of the strength-5 vortex on the uniform
    level-4 forest (128^2 cells, CFL 0.4 input = 0.3 effective): the GPU
    trajectory matches the oracle's, and both have the same L1 density error
    against the exactly advected vortex; halving h (level 5, same physical
    time) lowers it as a first-order scheme must"""
    log("config1 200 steps")
    f = amrhip.Forest(1, 1, (1, 1), 0.0, 0.0, 10.0, 8, 4, 4)
    f.apply_ic(amrhip.vortex_ic(5.0, 5.0, 5.0))
    ci, fi, geom, soln = f.external()
    setup(gpu, 8, 3, 10.0, (0, 0, 0, 0), 0.3, 1.0, ci, fi, geom, soln)
    gpu.run(200)
    t4 = float(gpu.time()[0][0])
    q_gpu = gpu.soln()
    p = oracle.params(8, cfl=0.3, h_base=10.0)
    q_orc, _, _ = oracle.run(p, ci, fi, geom, soln, 200, threads=8)
    assert parity_err(q_gpu, q_orc) <= TRAJ_TOL
    e_gpu = vortex_l1(f, q_gpu, t4)
    e_orc = vortex_l1(f, q_orc, t4)
    log(f"L1 level 4: gpu {e_gpu:.6e} oracle {e_orc:.6e} at t={t4:.4f}")
    assert abs(e_gpu - e_orc) <= 1e-9 * e_orc
    assert 1e-5 < e_gpu < 5e-2
    # level 5 to the same physical time
    f5 = amrhip.Forest(1, 1, (1, 1), 0.0, 0.0, 10.0, 8, 5, 5)
    f5.apply_ic(amrhip.vortex_ic(5.0, 5.0, 5.0))
    ci5, fi5, g5, s5 = f5.external()
    setup(gpu, 8, 3, 10.0, (0, 0, 0, 0), 0.3, 1.0, ci5, fi5, g5, s5, t_end=t4)
    while gpu.time()[0][0] < t4:
        gpu.run(50)
    assert gpu.time()[0][0] == pytest.approx(t4, abs=1e-14)
    e5 = vortex_l1(f5, gpu.soln(), t4)
    log(f"L1 level 5: {e5:.6e} (ratio {e_gpu / e5:.2f})")
    assert e5 < 0.75 * e_gpu


def dmr_forest(seed=4, n=8):
    rng = np.random.default_rng(seed)
    f = amrhip.Forest(4, 1, (0, 0), 0.0, 0.0, 1.0, n, 1, 3)
    f.refine((rng.uniform(size=f.nquad) < 0.5).astype(np.int32))
    f.balance()
    f.refine((rng.uniform(size=f.nquad) < 0.3).astype(np.int32))
    f.balance()
    f.apply_ic(amrhip.dmr_ic())
    return f


def dmr_oracle_params(oracle, t, n=8):
    post, pre, w = amrhip.dmr_states()
    return oracle.params(n, bc=(3, 0, 4, 5), cfl=0.3, bstate=(post, pre), shock=(1.0 / 6.0, 60.0, w), time=t)


@pytest.mark.parametrize("n", [4, 8, 16])
@pytest.mark.parametrize("t", [0.0, 0.07])
def test_rhs_dmr_walls_every_tile(gpu, oracle, t, n):
    """wall types 3-5 (position- and time-dependent states, evaluated at the
    sub-cell face centres) on a hanging-face mesh of [0,4] x [0,1], every
    tile variant of each patch size, against the oracle"""
    f = dmr_forest(n=n)
    ci, fi, geom, soln = f.external()
    gpu.setup(n, 3, [amrhip.dmr_inst(1.0, 0.3, 1.0)], ci, fi, geom, soln)
    gpu.set_time([t])
    ref = oracle.rhs(dmr_oracle_params(oracle, t, n), ci, fi, geom, soln)
    for ppb, minw, design in amrhip.tile_variants(n):
        log(f"dmr walls n={n} t={t} tile {(ppb, minw, design)}")
        gpu.set_tile(ppb, minw, design)
        assert parity_err(gpu.rhs(), ref) <= RHS_TOL


def test_dmr_trajectory_against_oracle(gpu, oracle):
    """10 SSP-RK3 steps with the moving-shock walls; the boundary time of
    every stage is the step's start time on both sides"""
    import ctypes
    import pyoracle
    f = dmr_forest(5)
    ci, fi, geom, soln = f.external()
    gpu.setup(8, 3, [amrhip.dmr_inst(1.0, 0.3, 1.0)], ci, fi, geom, soln)
    gpu.run(10)
    q_gpu = gpu.soln()
    q = np.array(soln, dtype=float, copy=True)
    work = np.zeros(5 * q.size)
    t = 0.0
    cii = np.ascontiguousarray(ci, np.int32)
    fii = np.ascontiguousarray(fi, np.int32)
    g = np.ascontiguousarray(geom, float)
    for _ in range(10):
        p = dmr_oracle_params(oracle, t)
        dt = oracle.time_step(p, ci, geom, q, 1.0)
        oracle.L.orc_rk_step(ctypes.byref(p), cii.size // 3, pyoracle.ipt(cii), pyoracle.dpt(g),
                             fii.size // 11, pyoracle.ipt(fii), q.size, dt, pyoracle.dpt(q),
                             pyoracle.dpt(work))
        t += dt
    assert gpu.time()[0][0] == pytest.approx(t, rel=1e-13)
    assert parity_err(q_gpu, q) <= TRAJ_TOL


def test_position_walls_need_geometry_after_remesh(gpu):
    f = dmr_forest(6)
    ci, fi, geom, soln = f.external()
    gpu.setup(8, 3, [amrhip.dmr_inst(1.0, 0.3, 1.0)], ci, fi, geom, soln)
    src = np.arange(f.nquad, dtype=np.int32)
    gpu.remesh(ci, fi, src)
    with pytest.raises(amrhip.GpuError):
        gpu.run(1)
    amrhip.check(amrhip.lib().amrgpu_set_geometry(gpu.h, amrhip.dpt(np.ascontiguousarray(geom))), "geom")
    gpu.run(1)


def test_rhs_after_regrow_remesh_keeps_step_time(oracle):
    """a remesh that outgrows the context's buffers (every quadrant of a
    uniform level-1 forest refined: 16 -> 64 quadrants, capacity 36) keeps
    the step start time that the moving-shock wall (type 5) reads, so the
    residual right after it matches the oracle's at that time"""
    f = amrhip.Forest(4, 1, (0, 0), 0.0, 0.0, 1.0, 8, 1, 3)
    f.apply_ic(amrhip.dmr_ic())
    ci, fi, geom, soln = f.external()
    nq0 = f.nquad
    g = amrhip.Gpu()
    try:
        g.setup(8, 3, [amrhip.dmr_inst(1.0, 0.3, 1.0)], ci, fi, geom, soln)
        t = 0.07
        g.set_time([t])
        f.track_begin(keep_data=True)
        f.regrid_tags(np.ones(nq0, np.int32))
        ok, src = f.track_end()
        assert ok and f.nquad == 4 * nq0 > nq0 + nq0 // 4 + 16
        ci2, fi2, geom2, soln2 = f.external()
        g.remesh(ci2, fi2, src)
        amrhip.check(amrhip.lib().amrgpu_set_geometry(g.h, amrhip.dpt(np.ascontiguousarray(geom2))), "geom")
        assert np.array_equal(g.soln(), soln2)
        ref = oracle.rhs(dmr_oracle_params(oracle, t), ci2, fi2, geom2, soln2)
        assert parity_err(g.rhs(), ref) <= RHS_TOL
        assert g.time()[0][0] == t
    finally:
        g.close()


@pytest.mark.parametrize("scheme", [3, 4, 5])
def test_persistent_design_equals_paired_faces(oracle, scheme):
    """the persistent, prefetching 16x16 kernel (several patches per
    workgroup, the next patch's state and halo loaded into LDS while the
    current one is evaluated) against the one-patch-per-workgroup paired-face
    kernel it restates, on a batch of three instances (adaptive forests with
    hanging faces and walls, and a periodic one) large enough that every
    workgroup walks several patches: residual and five steps bit for bit
    (both kernels fuse only the multiply-adds their source writes as fma),
    and the residual of each instance against the oracle"""
    n = 16
    forests = [riemann_forest(n, seed=21, lmin=3, lmax=7), riemann_forest(n, seed=22, lmin=3, lmax=7),
               vortex_forest(n, 5)]
    insts = [amrhip.make_inst(1.0, (2, 1, 0, 1), 0.3, 1.0, 0.0), amrhip.make_inst(1.0, (0, 0, 1, 1), 0.25, 1.0, 0.0),
             amrhip.make_inst(10.0, (0, 0, 0, 0), 0.3, 1.0, 0.0)]
    arrays = amrhip.batch_external(forests)
    nq = arrays[0].size // 3
    log(f"persistent vs paired: scheme {scheme}, {nq} quads")
    assert nq > 3 * 768      # several patches per resident workgroup (768 of them on 256 CUs)
    out = {}
    for tile in [(1, 4, 1)] + [v for v in amrhip.tile_variants(n) if v[2] >= 2]:
        g = amrhip.Gpu()
        g.setup(n, scheme, insts, *arrays[:4], arrays[4])
        g.set_tile(*tile)
        r = g.rhs()
        g.run(5)
        out[tile] = (r, g.soln(), g.time())
        g.close()
    ref = out[(1, 4, 1)]
    assert len(out) >= 2
    # the residual of each instance against the oracle's
    exp, off = [], 0
    for f, ip in zip(forests, insts):
        ci, fi, geom, soln = f.external()
        p = oracle.params(n, bc=tuple(ip.bc), cfl=ip.cfl, dt=1.0, h_base=ip.h_base, scheme=scheme)
        exp.append(oracle.rhs(p, ci, fi, geom, soln))
    for tile, (r, *_rest) in out.items():
        off = 0
        for e in exp:
            assert parity_err(r[off:off + e.size], e) <= RHS_TOL, tile
            off += e.size
    for tile, (r, q, (t, dt)) in out.items():
        assert np.array_equal(r, ref[0]), tile
        assert np.array_equal(q, ref[1]), tile
        assert np.array_equal(t, ref[2][0]) and np.array_equal(dt, ref[2][1]), tile
