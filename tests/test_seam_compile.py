"""The seam header binds to the reference's own physics sources: compiled
with -D_GPU_ and this repository's include/ on the path, the reference's
physics_evolve_interface.c (which calls physics_go_gpu,
src/physics/src/physics_evolve_interface.c:42-46) and hpc_amr_evolve_solver.c
(which defines hpc_amr_evolve_solver with an MPI_Comm argument) pass a
syntax-only compile.  This image has no MPI: a stand-in mpi.h of typedefs is
written for the compile check only (nothing is built or run from it).  Skipped
where the reference sources are absent."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

REF_PHYS = "/root/reference/src/physics/src"
FILES = ("physics_evolve_interface.c", "hpc_amr_evolve_solver.c")

pytestmark = pytest.mark.skipif(not os.path.isdir(REF_PHYS) or shutil.which("gcc") is None,
                                reason="reference physics sources or gcc absent")

MPI_TYPEDEFS = """#ifndef MPI_STANDIN_H
#define MPI_STANDIN_H
typedef int MPI_Comm;
int MPI_Abort(MPI_Comm comm, int errorcode);
#endif
"""


def compile_seam(tmp_path, seam_dir, src):
    (tmp_path / "mpi.h").write_text(MPI_TYPEDEFS)
    cmd = ["gcc", "-std=gnu11", "-fsyntax-only", "-Wall", "-Werror=implicit-function-declaration",
           "-D_GPU_", f"-I{seam_dir}", f"-I{tmp_path}", f"-I{REF_PHYS}", os.path.join(REF_PHYS, src)]
    return subprocess.run(cmd, capture_output=True, text=True)


@pytest.mark.parametrize("src", FILES)
def test_reference_seam_compiles_with_gpu_header(tmp_path, src):
    r = compile_seam(tmp_path, os.path.join(ROOT, "include"), src)
    assert r.returncode == 0, r.stderr


def test_compile_check_catches_conflicting_declaration(tmp_path):
    """a seam header that also declares the library's void*-communicator
    hpc_amr_evolve_solver (amr_hip_dropin.h) must fail on the reference's
    MPI_Comm definition: the check above can see the conflict it guards"""
    bad = tmp_path / "bad"
    bad.mkdir()
    txt = open(os.path.join(ROOT, "include", "physics_go_gpu.h")).read()
    txt = txt.replace('#include "amr_hip.h"', '#include "amr_hip_dropin.h"')
    (bad / "physics_go_gpu.h").write_text(txt)
    for h in ("amr_hip.h", "amr_hip_dropin.h"):
        shutil.copy(os.path.join(ROOT, "include", h), bad / h)
    r = compile_seam(tmp_path, str(bad), "hpc_amr_evolve_solver.c")
    assert r.returncode != 0 and "conflicting types" in r.stderr, r.stderr
