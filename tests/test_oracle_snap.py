"""Pins of the oracle's section sampling (snap__write, m_snap.f90:919-2189).

  * a linear velocity field V = A r has exact staggered differences, so every
    ps section must give div = trace(A) and rot = the antisymmetric part of A
    (component pairing and 1/dx factors), times the output scale UC M0 1e-3;
  * a field holding its own indices shows where each section samples
    (ii -> i = ii*idec - idec/2, the plane indices, kfs+1 / kob+1);
  * u sections integrate V dt, and the fs/ob maps keep running maxima.
"""
import numpy as np
import pytest

from oracle import oracle as O

OFF = O.OFF


def _run(nx=20, ny=18, nz=16, prec=64):
    cfg = O.Config(nx=nx, ny=ny, nz=nz, na=4, nm=0, zbeg=-2.0, dt=0.02, sources=[],
                   layers=[[0.0, 2.7, 5.5, 3.2, 1e6, 1e6]])
    return O.OracleRun(O.build_model(cfg), prec)


ALL = [f"{p}_{q}" for p in ("xy", "yz", "xz", "fs", "ob") for q in ("ps", "v", "u")]


def _coords(cfg):
    """staggered positions (k,i,j)-indexed numpy grids of Vx, Vy, Vz points"""
    ny6, nx6, nz6 = O.n3(cfg)
    j, i, k = np.meshgrid(np.arange(ny6) - OFF, np.arange(nx6) - OFF, np.arange(nz6) - OFF, indexing="ij")
    dx, dy, dz = cfg.dx, cfg.dy, cfg.dz
    xc, yc, zc = (i - 0.5) * dx, (j - 0.5) * dy, (k - 0.5) * dz
    return (xc + dx / 2, yc, zc), (xc, yc + dy / 2, zc), (xc, yc, zc + dz / 2)


def test_ps_of_linear_field(orc):
    run = _run()
    cfg, m = run.m.cfg, run.m
    rng = np.random.default_rng(3)
    A = rng.standard_normal((3, 3))
    (px, py, pz) = _coords(cfg)
    for n, (x, y, z) in zip(("Vx", "Vy", "Vz"), (px, py, pz)):
        comp = "xyz".index(n[1])
        run.arr[n][...] = A[comp, 0] * x + A[comp, 1] * y + A[comp, 2] * z
    for n in ("Sxy", "Sxz", "Syz"):
        run.arr[n][...] = 1.0
    run.snap_setup(ALL, k0_xy=7, i0_yz=9, j0_xz=5, idec=2, jdec=3, kdec=2, ntdec_s=1, nframe=1)
    run.snap_write(1)
    eps = float(np.float32(1.1920928955078125e-07))
    lam = 1.0  # solid everywhere below the air
    scale = float(np.float32(m.UC)) * float(np.float32(m.M0)) * float(np.float32(1e-3))
    want = np.array([np.trace(A) * lam / abs(lam + eps), (A[2, 1] - A[1, 2]) / (1 + eps),
                     (A[0, 2] - A[2, 0]) / (1 + eps), (A[1, 0] - A[0, 1]) / (1 + eps)]) * scale
    for pl in ("xy", "yz", "xz", "fs", "ob"):
        fr = run.snap_frames[f"{pl}_ps"][0]
        for c in range(1, 4):
            np.testing.assert_allclose(fr[c], want[c], rtol=1e-10, err_msg=f"{pl} rot {c}")
        if pl in ("yz", "xz"):
            # depth rows: the 2 km of air (k <= 4 at kdec = 2: rows 1, 2) carry lam = 0
            np.testing.assert_array_equal(fr[0][:2], 0.0)
            np.testing.assert_allclose(fr[0][2:], want[0], rtol=1e-6, err_msg=pl)
        else:
            np.testing.assert_allclose(fr[0], want[0], rtol=1e-6, err_msg=pl)


def test_section_sampling_points(orc):
    run = _run(nx=21, ny=17, nz=15)
    cfg, m = run.m.cfg, run.m
    ny6, nx6, nz6 = O.n3(cfg)
    j, i, k = np.meshgrid(np.arange(ny6) - OFF, np.arange(nx6) - OFF, np.arange(nz6) - OFF, indexing="ij")
    code = 10000.0 * i + 100.0 * j + k
    run.arr["Vx"][...] = code
    run.arr["Vy"][...] = -code
    run.arr["Vz"][...] = 2 * code
    idec, jdec, kdec, k0, i0, j0 = 2, 3, 2, 6, 8, 4
    run.snap_setup(ALL, k0_xy=k0, i0_yz=i0, j0_xz=j0, idec=idec, jdec=jdec, kdec=kdec, ntdec_s=1, nframe=1)
    run.snap_write(1)
    nxs, nys, nzs = (21 + 1) // 2, (17 + 1) // 3, (15 + 1) // 2
    ii = np.arange(1, nxs + 1) * idec - idec // 2
    jj = np.arange(1, nys + 1) * jdec - jdec // 2
    kk = np.arange(1, nzs + 1) * kdec - kdec // 2
    sc = float(np.float32(m.UC)) * float(np.float32(m.M0))
    kfs, kob = m.kfs, m.kob
    expect = {
        "xy": 10000.0 * ii[None, :] + 100.0 * jj[:, None] + k0,
        "yz": 10000.0 * i0 + 100.0 * jj[None, :] + kk[:, None],
        "xz": 10000.0 * ii[None, :] + 100.0 * j0 + kk[:, None],
        "fs": 10000.0 * ii[None, :] + 100.0 * jj[:, None] + kfs[jj[:, None] + OFF, ii[None, :] + OFF] + 1,
        "ob": 10000.0 * ii[None, :] + 100.0 * jj[:, None] + kob[jj[:, None] + OFF, ii[None, :] + OFF] + 1,
    }
    assert np.all(kfs[OFF + 1:OFF + 18, OFF + 1:OFF + 22] == 4)   # 2 km of air over dz = 0.5
    for pl, e in expect.items():
        fr = run.snap_frames[f"{pl}_v"][0]
        np.testing.assert_allclose(fr[0], e * sc, rtol=1e-12, err_msg=pl)
        np.testing.assert_allclose(fr[1], -e * sc, rtol=1e-12, err_msg=pl)
        np.testing.assert_allclose(fr[2], 2 * e * sc, rtol=1e-12, err_msg=pl)


def test_u_integrates_and_maps_keep_maxima(orc):
    run = _run()
    m = run.m
    run.snap_setup(["fs_v", "fs_u", "xy_u"], k0_xy=5, ntdec_s=2, nframe=3)
    vals = [1.0, -3.0, 2.0, 0.5, -1.0]
    for it, v in enumerate(vals, start=1):
        for n in ("Vx", "Vy", "Vz"):
            run.arr[n][...] = v
        run.snap_write(it)
    sc = float(np.float32(m.UC)) * float(np.float32(m.M0))
    dt = float(np.float32(m.cfg.dt))
    u = np.cumsum(vals) * sc * dt
    fr = run.snap_frames["xy_u"]
    # frames at it = 1, 3, 5 hold the accumulated displacement
    for f, it in enumerate((1, 3, 5)):
        np.testing.assert_allclose(fr[f], u[it - 1], rtol=1e-10)
    vmax = run.snap_vmax["fs"]
    np.testing.assert_allclose(vmax[0], 3.0 * sc, rtol=1e-12)                # max |Vz|
    np.testing.assert_allclose(vmax[1], 3.0 * np.sqrt(2) * sc, rtol=1e-12)   # max horizontal
    np.testing.assert_allclose(vmax[2], np.hypot(vmax[0], vmax[1]), rtol=1e-12)
    umax = run.snap_umax["fs"]
    np.testing.assert_allclose(umax[0], np.abs(u).max(), rtol=1e-10)


@pytest.mark.parametrize("shear", [-0.5, -2.0, 0.25])
def test_ps_masks_near_epsilon(orc, shear):
    """Shear stresses of a few epsilon, negative or positive, weight the
    rotation by |S|/|S + eps| on the yz section (m_snap.f90:1028-1030) and by
    |S|/(|S| + eps) on every other plane (:1121-1123): closed forms that tell
    the two masks apart (for S = -eps/2 they are 1 and 1/3)."""
    run = _run(prec=32)
    cfg, m = run.m.cfg, run.m
    rng = np.random.default_rng(5)
    A = rng.standard_normal((3, 3))
    (px, py, pz) = _coords(cfg)
    for n, (x, y, z) in zip(("Vx", "Vy", "Vz"), (px, py, pz)):
        comp = "xyz".index(n[1])
        run.arr[n][...] = A[comp, 0] * x + A[comp, 1] * y + A[comp, 2] * z
    eps = float(np.float32(1.1920928955078125e-07))
    s = np.float32(shear * eps)
    for n in ("Sxy", "Sxz", "Syz"):
        run.arr[n][...] = s
    run.snap_setup(ALL, k0_xy=7, i0_yz=9, j0_xz=5, idec=2, jdec=3, kdec=2, ntdec_s=1, nframe=1)
    run.snap_write(1)
    scale = float(np.float32(m.UC)) * float(np.float32(m.M0)) * 1e-3
    rot = np.array([A[2, 1] - A[1, 2], A[0, 2] - A[2, 0], A[1, 0] - A[0, 1]])
    w_yz = abs(float(s)) / abs(float(s) + eps)
    w_other = abs(float(s)) / (abs(float(s)) + eps)
    # the two forms differ only for negative shear stresses
    assert (abs(w_yz - w_other) > 0.1) == (shear < 0)
    for pl in ("xy", "yz", "xz", "fs", "ob"):
        fr = run.snap_frames[f"{pl}_ps"][0]
        w = w_yz if pl == "yz" else w_other
        for c in range(1, 4):
            np.testing.assert_allclose(fr[c], rot[c - 1] * w * scale, rtol=2e-5, err_msg=f"{pl} rot {c}")
