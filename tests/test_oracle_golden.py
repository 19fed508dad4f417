"""The CPU oracle against golden vectors written by the reference's own
compiled hot-path sources (tests/golden/make_golden.py).  Same operation
order, no FMA on either side: the match is bit for bit."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden_meshes


def test_flux_roe_2d_bitwise(oracle):
    d = np.load(os.path.join(GOLDEN, "flux.npz"))
    got = np.array([oracle.flux(d["ql"][k], d["qr"][k], d["nrm"][k]) for k in range(len(d["ql"]))])
    assert np.array_equal(got, d["flux"])


def test_bc_and_boundary_flux_bitwise(oracle):
    d = np.load(os.path.join(GOLDEN, "bcflux.npz"))
    for r in range(len(d["idx"])):
        q = d["q"][d["idx"][r]]
        qb = oracle.bc(int(d["bc"][r]), d["nrm"][r], q)
        assert np.array_equal(qb, d["qb"][r])
        assert np.array_equal(oracle.flux_bc(qb, d["nrm"][r]), d["flux"][r])


@pytest.mark.parametrize("path", golden_meshes(), ids=lambda p: os.path.basename(p)[5:-4])
def test_rhs_and_time_step_bitwise(oracle, path):
    m = np.load(path)
    n = int(m["n"])
    p = oracle.params(n, bc=tuple(int(b) for b in m["bc"]), cfl=float(m["cfl"]), h_base=float(m["h_base"]))
    rhs = oracle.rhs(p, m["cell_info"], m["face_info"], m["geom"], m["soln"])
    assert np.array_equal(rhs, m["rhs"])
    dt = oracle.time_step(p, m["cell_info"], m["geom"], m["soln"], 1.0)
    assert dt == float(m["dt"])


def test_initial_condition_bitwise(oracle):
    d = np.load(os.path.join(GOLDEN, "ic.npz"))
    g = d["geom"]
    for k in range(len(g) // 12):
        gk = g[12 * k:12 * k + 12]
        free = oracle.ic(0, 4, gk, 0.5, 30.0, 0.0, 1.4, 1.0, 1.0 / 1.4, 3.0, 0.0, 0.0)
        vort = oracle.ic(1, 4, gk, 0.5, 0.0, 0.0, 1.4, 1.0, 0.714286, 3.0, 0.0, 0.0)
        assert np.array_equal(free, d["free"][k])
        assert np.array_equal(vort, d["vortex"][k])


def test_coarsen_and_tag_ops(oracle):
    import ctypes as C
    from pyoracle import dpt, ipt
    d = np.load(os.path.join(GOLDEN, "ops.npz"))
    n = 4
    out = np.zeros(4 * n * n)
    ch = [np.ascontiguousarray(c) for c in d["children"]]
    oracle.L.orc_coarsen_2d(n, dpt(ch[0]), dpt(ch[1]), dpt(ch[2]), dpt(ch[3]), dpt(out))
    assert np.array_equal(out, d["coarse"])
    for k in range(len(d["tags"])):
        m = np.ascontiguousarray(d["tag_methods"][k], np.int32)
        t = np.ascontiguousarray(d["tag_tols"][k])
        q = np.ascontiguousarray(d["tag_q"][k])
        nt = int(np.count_nonzero(m))
        assert oracle.L.orc_tag_feature(nt, ipt(m), dpt(t), n, dpt(q)) == int(d["tags"][k])
    # solver_refine_operator_2d: every child a copy of the parent
    parent = np.ascontiguousarray(d["refine_parent"])
    kids = [np.full(4 * n * n, np.nan) for _ in range(4)]
    oracle.L.orc_refine_2d(n, dpt(parent), *[dpt(k) for k in kids])
    for k, r in zip(kids, d["refine_children"]):
        assert np.array_equal(k, r) and np.array_equal(k, parent)
    _ = C


def norm_cases():
    d = np.load(os.path.join(GOLDEN, "norm.npz"))
    return [str(x) for x in d["names"]]


@pytest.mark.parametrize("name", norm_cases())
def test_norm_L2_residual_bitwise(oracle, name):
    """orc_norm_L2_residual against physics_norm_L2_residual: the golden
    meshes' residuals, and random residuals spanning 12 decades with permuted
    patch offsets and ghost patches (so the summation order and the offsets
    both show in the bits)"""
    d = np.load(os.path.join(GOLDEN, "norm.npz"))
    n = int(d[f"{name}_n"])
    got = oracle.norm(n, d[f"{name}_cell_info"], d[f"{name}_res"])
    assert got == float(d[f"{name}_norm"])
