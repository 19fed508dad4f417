"""Write golden vectors from the reference's own hot-path build.

Runs only in the development container, where /root/reference exists and
`make -C oracle ref` has compiled the reference's Fortran kernels and C glue
into oracle/_ref/libamrref.so.  The fixtures written here are data (inputs and
the reference's outputs) and are committed; no reference source travels.

    python tests/golden/make_golden.py

Fixtures:
  flux.npz       flux_roe_2d on random state pairs and normals
  bcflux.npz     bc_2d + flux_bc_2d for bc types 0/1/2 and the four sides
  mesh_*.npz     physics_rhs_structured and physics_time_step on host-forest
                 meshes (uniform periodic, hanging faces, every bc type,
                 n = 2, 4, 8)
  ic.npz         initial_condition_2d (free stream and vortex, reference
                 parameters)
  ops.npz        solver_tag_feature, solver_coarsen_operator_2d,
                 solver_refine_operator_2d
  norm.npz       physics_norm_L2_residual on the golden meshes' residuals and
                 on residual vectors with permuted patch offsets and ghost
                 patches after the real ones

    python tests/golden/make_golden.py norm     # norm.npz only (own seed)
"""
import os
import sys

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, os.path.join(ROOT, "amr-demo_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import amrhip  # noqa: E402
import pyoracle  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def rand_states(rng, m, supersonic=False):
    rho = np.exp(rng.uniform(np.log(0.1), np.log(2.0), m))
    vmax = 4.0 if supersonic else 1.0
    u = rng.uniform(-vmax, vmax, m)
    v = rng.uniform(-vmax, vmax, m)
    p = rng.uniform(0.1, 2.0, m)
    return np.stack([rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v)], axis=1)


def gen_flux(ref, rng):
    ql = np.concatenate([rand_states(rng, 300), rand_states(rng, 100, True)])
    qr = np.concatenate([rand_states(rng, 300), rand_states(rng, 100, True)])
    h = 2.0 ** -rng.integers(2, 8, len(ql)) * 5.0
    kinds = rng.integers(0, 5, len(ql))
    nrm = np.zeros((len(ql), 2))
    for k in range(len(ql)):
        if kinds[k] == 0:
            nrm[k] = (h[k], 0.0)
        elif kinds[k] == 1:
            nrm[k] = (0.0, h[k])
        elif kinds[k] == 2:
            nrm[k] = (-h[k], 0.0)
        elif kinds[k] == 3:
            nrm[k] = (0.0, -h[k])
        else:
            nrm[k] = rng.uniform(-1, 1, 2)
    f = np.array([ref.flux(ql[k], qr[k], nrm[k]) for k in range(len(ql))])
    np.savez_compressed(os.path.join(OUT, "flux.npz"), ql=ql, qr=qr, nrm=nrm, flux=f)


def gen_bcflux(ref, rng):
    q = rand_states(rng, 120)
    rows = []
    for k in range(len(q)):
        for bc in (0, 1, 2):
            for side in range(4):
                h = 2.0 ** -int(rng.integers(2, 7))
                nrm = [(-h, 0.0), (h, 0.0), (0.0, -h), (0.0, h)][side]
                qb = ref.bc(bc, nrm, q[k])
                fb = ref.flux_bc(qb, nrm)
                rows.append((k, bc, nrm[0], nrm[1], *qb, *fb))
    a = np.array(rows)
    np.savez_compressed(os.path.join(OUT, "bcflux.npz"), q=q, idx=a[:, 0].astype(int),
                        bc=a[:, 1].astype(int), nrm=a[:, 2:4], qb=a[:, 4:8], flux=a[:, 8:12])


def smooth_state(forest, rng, amp=0.3):
    """positive smooth state plus cell noise from the forest's centres"""
    soln = forest.soln()
    nq, n = forest.nquad, forest.n
    ci, fi, geom, _ = forest.external()
    g = geom.reshape(nq, 4, 3)
    x0, y0 = g[:, 0, 0], g[:, 0, 1]
    dx = (g[:, 1, 0] - g[:, 0, 0]) / n
    dy = (g[:, 2, 1] - g[:, 0, 1]) / n
    ii = np.arange(n) + 0.5
    X = x0[:, None, None] + dx[:, None, None] * ii[None, None, :]
    Y = y0[:, None, None] + dy[:, None, None] * ii[None, :, None]
    k1, k2 = rng.uniform(1, 3, 2)
    rho = 1.0 + amp * np.sin(k1 * X) * np.cos(k2 * Y) + 0.05 * rng.standard_normal(X.shape)
    u = 0.3 * np.cos(k2 * Y) + 0.05 * rng.standard_normal(X.shape)
    v = -0.2 * np.sin(k1 * X) + 0.05 * rng.standard_normal(X.shape)
    p = 1.0 + 0.2 * np.cos(k1 * X + k2 * Y) + 0.05 * rng.standard_normal(X.shape)
    soln[..., 0] = rho
    soln[..., 1] = rho * u
    soln[..., 2] = rho * v
    soln[..., 3] = p / 0.4 + 0.5 * rho * (u * u + v * v)


def refine_some(forest, rng, frac, rounds):
    for _ in range(rounds):
        flags = (rng.uniform(size=forest.nquad) < frac).astype(np.int32)
        forest.refine(flags)
        forest.balance()


def gen_mesh(ref, rng, name, n, ntx, nty, periodic, scale, min_level, max_level, bc, frac, rounds):
    f = amrhip.Forest(ntx, nty, periodic, 0.0, 0.0, scale, n, min_level, max_level)
    refine_some(f, rng, frac, rounds)
    assert f.is_balanced()
    smooth_state(f, rng)
    ci, fi, geom, soln = f.external()
    cfl = 0.3
    ref.configure(n, bc=bc, cfl=cfl, h_base=scale)
    rhs = ref.rhs(n, ci, fi, geom, soln)
    dt = ref.time_step(n, ci, geom, soln, 1.0)
    np.savez_compressed(os.path.join(OUT, f"mesh_{name}.npz"), n=n, bc=np.array(bc), cfl=cfl,
                        h_base=scale, cell_info=ci, face_info=fi, geom=geom, soln=soln, rhs=rhs,
                        dt=dt, levels=f.levels())
    nh = int(np.sum(fi.reshape(-1, 11)[:, 0] == 3)) if fi.size else 0
    nb = int(np.sum(fi.reshape(-1, 11)[:, 0] == 1)) if fi.size else 0
    print(f"mesh_{name}: {f.nquad} quads, {fi.size // 11} faces ({nh} hanging, {nb} boundary)")


def gen_ic(ref, rng):
    f = amrhip.Forest(2, 2, (1, 1), -5.0, -5.0, 5.0, 4, 1, 1)
    ci, fi, geom, _ = f.external()
    rows_free, rows_vort = [], []
    for k in range(f.nquad):
        g = geom[12 * k:12 * k + 12]
        rows_free.append(ref.ic(0, 4, 1.0, g, 0.5, 30.0, 0.0, 1.4, 1.0, 1.0 / 1.4))
        rows_vort.append(ref.ic(1, 4, 1.0, g, 0.5, 0.0, 0.0, 1.4, 1.0, 0.714286))
    np.savez_compressed(os.path.join(OUT, "ic.npz"), geom=geom, n=4, free=np.array(rows_free),
                        vortex=np.array(rows_vort))


def gen_ops(ref, rng):
    n = 4
    qs = [rand_states(rng, n * n).reshape(-1) for _ in range(4)]
    coarse = ref.coarsen(n, *qs)
    tags = []
    cases = []
    for k in range(40):
        q = rand_states(rng, n * n).reshape(-1)
        methods = [int(rng.integers(1, 3)) for _ in range(int(rng.integers(1, 3)))]
        tols = [float(rng.uniform(0.1, 2.0)) for _ in methods]
        cases.append((q, methods, tols))
        tags.append(ref.tag(methods, tols, n, q))
    parent = rand_states(rng, n * n).reshape(-1)     # drawn last: the earlier fixtures stay as they were
    np.savez_compressed(os.path.join(OUT, "ops.npz"), children=np.array(qs), coarse=coarse,
                        tag_q=np.array([c[0] for c in cases]),
                        tag_methods=np.array([c[1] + [0] * (2 - len(c[1])) for c in cases]),
                        tag_tols=np.array([c[2] + [0.0] * (2 - len(c[2])) for c in cases]),
                        tags=np.array(tags), refine_parent=parent, refine_children=np.array(ref.refine(n, parent)))


def gen_norm(ref):
    """sum of squares over the real patches, in cell order at each patch's
    offset (physics_norm_L2_residual, src/physics/src/physics_norm_residual.c:15-56);
    inputs: the committed golden meshes' residuals, and random residuals whose
    magnitudes span many decades (so that the summation order shows in the
    last bits) with cell_info offsets permuted and ghost patches appended"""
    rng = np.random.default_rng(20261203)
    cases = []
    for name in ("uniform_periodic_n4", "uniform_periodic_n8", "hanging_walls_n4",
                 "hanging_periodic_x_n8", "hanging_periodic_n2", "brick_3x2_n16"):
        d = np.load(os.path.join(OUT, f"mesh_{name}.npz"))
        cases.append((f"mesh_{name}", int(d["n"]), d["cell_info"], d["rhs"]))
    for k, (n, ncell, nghost) in enumerate(((2, 7, 0), (4, 13, 3), (8, 21, 5), (16, 5, 2))):
        nt = ncell + nghost
        perm = rng.permutation(nt)[:ncell]
        ci = np.zeros((ncell, 3), np.int32)
        ci[:, 1] = rng.integers(0, 6, ncell)
        ci[:, 2] = perm * 4 * n * n
        res = rng.standard_normal(nt * 4 * n * n) * 10.0 ** rng.uniform(-8, 4, nt * 4 * n * n)
        cases.append((f"random_n{n}_{k}", n, ci.reshape(-1), res))
    out = {}
    for name, n, ci, res in cases:
        ref.configure(n)
        val = ref.norm(n, ci, res)
        out[f"{name}_n"] = np.int32(n)
        out[f"{name}_cell_info"] = np.asarray(ci, np.int32)
        out[f"{name}_res"] = np.asarray(res, float)
        out[f"{name}_norm"] = np.float64(val)
        print(f"norm {name}: n {n}, {len(ci) // 3} cells, sum of squares {val:.17g}")
    np.savez_compressed(os.path.join(OUT, "norm.npz"), names=np.array([c[0] for c in cases]), **out)


def main():
    if not pyoracle.ref_available():
        sys.exit("oracle/_ref/libamrref.so missing: run `make -C oracle ref` where /root/reference exists")
    ref = pyoracle.Reference()
    if sys.argv[1:] == ["norm"]:
        gen_norm(ref)
        return
    rng = np.random.default_rng(20240611)
    gen_flux(ref, rng)
    gen_bcflux(ref, rng)
    gen_mesh(ref, rng, "uniform_periodic_n4", 4, 1, 1, (1, 1), 1.0, 2, 2, (0, 0, 0, 0), 0.0, 0)
    gen_mesh(ref, rng, "uniform_periodic_n8", 8, 1, 1, (1, 1), 10.0, 2, 2, (0, 0, 0, 0), 0.0, 0)
    gen_mesh(ref, rng, "hanging_walls_n4", 4, 1, 1, (0, 0), 1.0, 1, 4, (1, 2, 0, 1), 0.3, 3)
    gen_mesh(ref, rng, "hanging_periodic_x_n8", 8, 2, 1, (1, 0), 1.0, 1, 3, (0, 0, 2, 1), 0.25, 2)
    gen_mesh(ref, rng, "hanging_periodic_n2", 2, 1, 1, (1, 1), 1.0, 2, 4, (0, 0, 0, 0), 0.3, 2)
    gen_mesh(ref, rng, "brick_3x2_n16", 16, 3, 2, (0, 1), 0.5, 0, 2, (2, 0, 1, 1), 0.3, 2)
    gen_ic(ref, rng)
    gen_ops(ref, rng)
    gen_norm(ref)


if __name__ == "__main__":
    main()
