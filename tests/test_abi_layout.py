"""The ctypes mirrors of the C ABI's structures (amrhip) have the C layout:
every field at the offset the C compiler gives it and the same total size,
checked by compiling a probe against include/*.h with the host compiler (no
GPU, no library load).  A drifted mirror would pass wrong option values to
the library without any error."""
import ctypes as C
import json
import os
import shutil
import subprocess

import pytest

import amrhip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

STRUCTS = {
    "amrgpu_inst_t": amrhip.InstParams,
    "amr_evolve_stats_t": amrhip.EvolveStats,
    "amr_collect_opts_t": amrhip.CollectOpts,
    "amr_mlp_tagger_t": amrhip.MlpTaggerStruct,
    "amr_feature_tagger_t": amrhip.FeatureTaggerStruct,
    "amr_evolve_opts_t": amrhip.EvolveOpts,
    "physics_gpu_params_t": amrhip.PhysicsParams,
    "amr_quad_t": amrhip.Quad,
    "amr_ic_t": amrhip.IC,
}


def c_layout(tmp_path):
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "amr_hip.h"', '#include "amr_forest.h"',
             '#include "amr_driver.h"', '#include "amr_tagger.h"', 'int main(void) {', 'printf("{");']
    first = True
    for cname, py in STRUCTS.items():
        parts = [f'\\"size\\": %zu']
        args = [f"sizeof({cname})"]
        for f, _ in py._fields_:
            parts.append(f'\\"{f}\\": %zu')
            args.append(f"offsetof({cname}, {f})")
        sep = "" if first else ", "
        first = False
        lines.append(f'printf("{sep}\\"{cname}\\": {{{", ".join(parts)}}}", {", ".join(args)});')
    lines += ['printf("}\\n");', "return 0;", "}"]
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True,
                   capture_output=True, text=True)
    return json.loads(subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs the host C compiler")
def test_ctypes_mirrors_match_the_c_layout(tmp_path):
    lay = c_layout(tmp_path)
    for cname, py in STRUCTS.items():
        got = lay[cname]
        assert got["size"] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[f] == getattr(py, f).offset, (cname, f)
