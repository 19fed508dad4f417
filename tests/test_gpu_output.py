"""Output of the evolve loops on the GPU path: VTK files and checkpoints at
the visualization / checkpoint intervals and at the end of the run (as the
reference evolve loop, src/amr/src/amr_evolve.c:45-66); the final checkpoint
restores the final state bit for bit."""
import os

import numpy as np
import pytest

import amrhip

pytestmark = pytest.mark.gpu


def run(tmp, device):
    amrhip.set_params(nelem_subgrid=8, time_scheme=3, time_steps=20, dt=1.0, cfl=0.3, coarse_grid_h=10.0,
                      regrid_interval=10, checkpoint_interval=10, visualization_interval=5,
                      bc=(0, 0, 0, 0, 1, 1), t_end=0.0, verbose=0)
    ic = amrhip.vortex_ic(5.0, 5.0, 5.0)
    f = amrhip.Forest(1, 1, (1, 1), 0.0, 0.0, 10.0, 8, 2, 3)
    f.apply_ic(ic)
    f.regrid_initial(ic, 0.02)
    amrhip.set_output_dir(str(tmp))
    try:
        st = f.evolve_device(0.02, 10) if device else f.evolve(0.02)
    finally:
        amrhip.set_output_dir(None)
    return f, st


@pytest.mark.parametrize("device", [True, False])
def test_evolve_writes_vtk_and_checkpoints(tmp_path, device):
    print(f"output device={device}", flush=True)
    f, st = run(tmp_path, device)
    assert st.steps == 20
    names = sorted(os.listdir(tmp_path))
    for k in (5, 10, 15, 20):
        assert f"soln_{k:07d}.vtu" in names
    ck = sorted(os.listdir(tmp_path / "checkpoint"))
    assert ck == ["amr_soln_0000010.bin", "amr_soln_0000020.bin"]
    g, istep, t = amrhip.Forest.load(tmp_path / "checkpoint" / "amr_soln_0000020.bin")
    assert istep == 20 and t > 0.0
    assert g.quads() == f.quads()
    assert np.array_equal(g.soln(), f.soln())
