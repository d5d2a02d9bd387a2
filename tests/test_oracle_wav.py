"""Pins of the oracle's stress and strain samples (wav__store,
src/swpc_3d/m_wav.f90:397-625), FP64:

  * strain: for a linear velocity field V = A r on the staggered points, the
    4th-order differences are exact, so each 4-point averaged derivative is
    an entry of A and the accumulated strain after one step is dt times the
    symmetric part of A, in the order xx yy zz yz xz xy, times M0 UC 1e-3;
  * stress: for linear stress fields the normal components sampled at the
    cell and the 4-point averages of the shear components around it both
    equal the field at the cell centre, times M0 UC 1e6.
The samples are single precision (the output arrays and their scale), so
the pins hold to float32 rounding.
A control plants a swapped pair and checks that the pins see it.
"""
import numpy as np
import pytest

from oracle import oracle as O

OFF = O.OFF


def _run():
    cfg = O.Config(nx=20, ny=18, nz=16, na=4, nm=0, zbeg=-2.0, dt=0.02, sources=[], wav_flags=15,
                   layers=[[0.0, 2.7, 5.5, 3.2, 1e6, 1e6]],
                   stations=[(0.6, -0.4, 2.3, "dep"), (-1.3, 1.1, 3.9, "dep"), (1.9, 0.8, 1.2, "dep")])
    return O.OracleRun(O.build_model(cfg), 64)


def _grids(cfg):
    """cell-centre coordinates on the (k,i,j) halo layout, numpy (j, i, k)"""
    ny6, nx6, nz6 = O.n3(cfg)
    j, i, k = np.meshgrid(np.arange(ny6) - OFF, np.arange(nx6) - OFF, np.arange(nz6) - OFF, indexing="ij")
    return (i - 0.5) * cfg.dx, (j - 0.5) * cfg.dy, (k - 0.5) * cfg.dz


def _centre(m, n):
    cfg = m.cfg
    i, j, k = int(m.ist[n]), int(m.jst[n]), int(m.kst[n])
    return np.array([(i - 0.5) * cfg.dx, (j - 0.5) * cfg.dy, (k - 0.5) * cfg.dz])


def test_strain_of_a_linear_velocity_field(orc):
    run = _run()
    m, cfg = run.m, run.m.cfg
    A = np.random.default_rng(4).standard_normal((3, 3))
    x, y, z = _grids(cfg)
    pos = {"Vx": (x + cfg.dx / 2, y, z), "Vy": (x, y + cfg.dy / 2, z), "Vz": (x, y, z + cfg.dz / 2)}
    for c, name in enumerate(("Vx", "Vy", "Vz")):
        px, py, pz = pos[name]
        run.arr[name][...] = A[c, 0] * px + A[c, 1] * py + A[c, 2] * pz
    run.wav_store(1)
    scale = float(np.float32(m.M0)) * float(np.float32(m.UC)) * 1e-3
    dt = float(np.float32(cfg.dt))
    sym = [A[0, 0], A[1, 1], A[2, 2], (A[1, 2] + A[2, 1]) / 2, (A[0, 2] + A[2, 0]) / 2, (A[0, 1] + A[1, 0]) / 2]
    assert m.nst == 3
    for n in range(m.nst):
        got = run.wav["strain"][n, :, 0]
        np.testing.assert_allclose(got, np.array(sym) * dt * scale, rtol=3e-7, atol=1e-12 * scale, err_msg=str(n))
    # control: the same pin fails for a swapped off-diagonal pairing
    wrong = list(sym)
    wrong[3], wrong[4] = sym[4], sym[3]
    assert not np.allclose(run.wav["strain"][0, :, 0], np.array(wrong) * dt * scale, rtol=1e-6, atol=0)


def test_stress_of_linear_stress_fields(orc):
    run = _run()
    m, cfg = run.m, run.m.cfg
    B = np.random.default_rng(5).standard_normal((6, 3))
    x, y, z = _grids(cfg)
    dx, dy, dz = cfg.dx, cfg.dy, cfg.dz
    pos = {"Sxx": (x, y, z), "Syy": (x, y, z), "Szz": (x, y, z), "Syz": (x, y + dy / 2, z + dz / 2),
           "Sxz": (x + dx / 2, y, z + dz / 2), "Sxy": (x + dx / 2, y + dy / 2, z)}
    names = ("Sxx", "Syy", "Szz", "Syz", "Sxz", "Sxy")
    for c, name in enumerate(names):
        px, py, pz = pos[name]
        run.arr[name][...] = B[c, 0] * px + B[c, 1] * py + B[c, 2] * pz
    run.wav_store(1)
    scale = float(np.float32(m.M0)) * float(np.float32(m.UC)) * 1e6
    for n in range(m.nst):
        want = B @ _centre(m, n) * scale
        np.testing.assert_allclose(run.wav["stress"][n, :, 0], want, rtol=3e-7, atol=1e-12 * scale, err_msg=str(n))


def test_samples_only_at_recorded_steps(orc):
    """ntdec_w decimation (m_wav.f90:517): sample itw at step it when
    mod(it - 1, ntdec_w) == 0; the accumulators integrate every step"""
    cfg = O.Config(nx=20, ny=18, nz=16, na=4, nm=0, zbeg=-2.0, dt=0.02, sources=[], wav_flags=10, ntdec_w=3,
                   layers=[[0.0, 2.7, 5.5, 3.2, 1e6, 1e6]], stations=[(0.6, -0.4, 2.3, "dep")])
    run = O.OracleRun(O.build_model(cfg), 64)
    m = run.m
    run.arr["Vx"][...] = 1.0
    for it in range(1, 8):
        run.wav_store(it)
    scale = float(np.float32(m.M0)) * float(np.float32(m.UC)) * 1e9
    dt = float(np.float32(cfg.dt))
    # U = integral of V dt: samples at it = 1, 4, 7 after 1, 4, 7 accumulations
    np.testing.assert_allclose(run.wav["disp"][0, 0, :3], np.array([1, 4, 7]) * dt * scale, rtol=3e-7)
