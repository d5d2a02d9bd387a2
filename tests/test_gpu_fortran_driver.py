"""The Fortran driver (openswpc_amd/host/swpc_fdrive.f90) calls the library
through the iso_c_binding module m_swpc_gpu, the way the reference's own
main.f90 would.  Its results must equal, bit for bit, the same model built
through the Python binding."""
import os
import subprocess

import numpy as np
import pytest

from openswpc_amd.host import ModelSpec, build_inputs, make_simulation

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "openswpc_amd", "lib", "swpc_fdrive")

pytestmark = pytest.mark.gpu


def _hex(a):
    return [f"{int(x):08X}" for x in np.ascontiguousarray(a, np.float32).view(np.uint32)]


def test_fortran_driver_matches_python_binding():
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (make -C openswpc_amd/host)")
    nx, ny, nz, nt = 48, 44, 40, 30
    out = subprocess.run([EXE, str(nx), str(ny), str(nz), str(nt)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = [l.split() for l in out.stdout.splitlines() if l.strip()]
    vmax_f = [l for l in lines if l[0] == "vmax"][0][1:]
    seis_f = {int(l[1]): l[2:] for l in lines if l[0] == "seis"}

    spec = ModelSpec(nx=nx, ny=ny, nz=nz, dx=0.5, dy=0.5, dz=0.5, dt=0.02, nt=nt, na=8, nm=3, zbeg=-2.0,
                     layers=[[0.0, 2.7, 5.5, 3.2, 320.0, 160.0], [5.0, 2.9, 6.5, 3.7, 370.0, 185.0]],
                     sources=[dict(x=0.3, y=-0.6, z=5.2, t0=0.0, tr=0.5, mo=1e15, strike=30.0, dip=60.0,
                                   rake=90.0)],
                     stations=[(2.1, 1.3, 0.0, "fsb"), (-3.2, 2.7, 0.0, "fsb")], wav_what=1)
    sim = make_simulation(build_inputs(spec))
    sim.step(1, nt, use_graph=True, block=10)
    vm = sim.kernel_vmax()
    seis = sim.seismograms("vel")
    sim.close()
    assert np.any(vm != 0)
    assert vmax_f == _hex(vm)
    for n in (1, 2):
        assert seis_f[n] == _hex(seis[n - 1, 2, :]), n
