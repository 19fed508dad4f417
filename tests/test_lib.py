"""The C-ABI library loads and exports every entry point its headers declare;
host-only logic (settings parser, interval early exit, threaded CPU baseline)
runs without a GPU."""
import ctypes as C
import glob
import os
import re

import numpy as np

import amrhip
from conftest import ROOT


def header_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\(", txt, flags=re.M):
            if m.group(0).startswith(("typedef", "static")):   # header-only wrappers
                continue
            if m.group(1) not in ("if", "while", "for", "return", "sizeof"):
                names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = amrhip.lib()
    names = header_functions()
    assert "physics_go_gpu" in names and "amr_forest_new" in names
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    assert set(amrhip.EXPORTS) <= names


def test_input_physics_semantics(tmp_path, monkeypatch):
    (tmp_path / "input.physics").write_text(
        "<SIMULATION PROPERTIES>\ncheckpoint_interval: 0\nvisualization_interval: 1000\n"
        "regrid_interval: 10\ntime_scheme: 3\ntime_steps: 200\ndt: 1.0\n"
        "</SIMULATION PROPERTIES>\n<FLUID PROPERTIES>\nbc_xlo: 0\nbc_xhi: 1\nbc_ylo: 2\nbc_yhi: 0\n"
        "cfl: 0.4\n</FLUID PROPERTIES>\n")
    monkeypatch.chdir(tmp_path)
    L = amrhip.lib()
    ints = [C.c_int(v) for v in (0, 2)]
    scale = C.c_double(10.0)
    rest = [C.c_int(v) for v in (4, 4, 2, 8, 3, 11, 0, 0, 2)]
    L.hpc_amr_initialize_solver(C.byref(ints[0]), C.byref(ints[1]), C.byref(scale),
                                *[C.byref(x) for x in rest])
    p = amrhip.get_params()
    assert p.time_scheme == 3 and p.time_steps == 200 and p.dt == 1.0
    assert abs(p.cfl - 0.25 * 0.4 * 3) < 1e-15          # physics_initialize.c:193
    assert p.checkpoint_interval == 2 ** 31 - 1           # <= 0 -> INT_MAX
    assert p.regrid_interval == 10 and p.visualization_interval == 1000
    assert list(p.bc)[:4] == [0, 1, 2, 0]
    assert p.coarse_grid_h == 10.0 and p.nelem_subgrid == 8


def test_physics_go_gpu_early_exit_without_gpu():
    amrhip.set_params(time_steps=5, time_scheme=3)
    L = amrhip.lib()
    z = [C.c_int(0) for _ in range(16)]
    dim, uns, nf, n = C.c_int(2), C.c_int(0), C.c_int(4), C.c_int(2)
    istep, evolve = C.c_int(5), C.c_int(1)
    flags = [C.c_int(0) for _ in range(3)]
    d = [C.c_double(0.0) for _ in range(3)]
    rc = L.physics_go_gpu(C.byref(dim), C.byref(uns), C.byref(nf), C.byref(n),
                          *[C.byref(x) for x in z[:11]], None, None, C.byref(istep),
                          *[C.byref(x) for x in flags], C.byref(evolve), None,
                          *[C.byref(x) for x in d])
    assert rc == 0 and evolve.value == 0     # physics_go.c:66-69


def test_physics_go_gpu_rejects_empty_mesh_without_gpu():
    """one GPU owns the whole forest: an interval with no quadrants is a
    caller error, reported before any device is touched, state untouched"""
    amrhip.set_params(time_steps=5, time_scheme=3)
    L = amrhip.lib()
    z = [C.c_int(0) for _ in range(16)]       # ncell = nghost = nface = dof = 0, ...
    dim, uns, nf, n = C.c_int(2), C.c_int(0), C.c_int(4), C.c_int(8)
    istep, evolve = C.c_int(0), C.c_int(1)
    flags = [C.c_int(0) for _ in range(3)]
    d = [C.c_double(0.0) for _ in range(3)]
    rc = L.physics_go_gpu(C.byref(dim), C.byref(uns), C.byref(nf), C.byref(n),
                          *[C.byref(x) for x in z[:11]], None, None, C.byref(istep),
                          *[C.byref(x) for x in flags], C.byref(evolve), None,
                          *[C.byref(x) for x in d])
    assert rc == -1                           # AMRGPU_ERR_ARG (include/amr_hip.h)
    assert b"empty mesh" in amrhip.lib().amrgpu_last_error()
    assert istep.value == 0 and evolve.value == 1


def test_threaded_cpu_baseline_matches_serial(oracle):
    rng = np.random.default_rng(21)
    f = amrhip.Forest(1, 1, (0, 1), 0.0, 0.0, 1.0, 4, 2, 4)
    for _ in range(2):
        f.refine((rng.uniform(size=f.nquad) < 0.3).astype(np.int32))
        f.balance()
    s = f.soln()
    s[..., 0] = rng.uniform(0.8, 1.2, s.shape[:-1])
    s[..., 1] = 0.1
    s[..., 2] = -0.05
    s[..., 3] = 2.5
    ci, fi, geom, soln = f.external()
    p = oracle.params(4, bc=(1, 2, 0, 0), cfl=0.3, dt=1.0, scheme=3)
    q1, rn1, dt1 = oracle.run(p, ci, fi, geom, soln, 3)
    q3, rn3, dt3 = oracle.run(p, ci, fi, geom, soln, 3, threads=3)
    assert np.array_equal(q1, q3) and dt1 == dt3
    assert abs(rn1 - rn3) <= 1e-13 * rn1
