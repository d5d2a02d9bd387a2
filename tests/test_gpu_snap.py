"""Sections (snap__write) on the GPU against the FP32 oracle, bit for bit:
all 15 sections, decimation, graph-replayed steps, maximum maps."""
import numpy as np
import pytest

from oracle import oracle as O
from gpu_helpers import gpu_from_model, small_config, ulp_diff

pytestmark = pytest.mark.gpu

ALL = [f"{p}_{q}" for p in ("xy", "yz", "xz", "fs", "ob") for q in ("ps", "v", "u")]


@pytest.mark.parametrize("dec,graph", [((1, 1, 1), False), ((2, 3, 2), True), ((3, 2, 1), True)])
def test_sections_match_oracle(orc, dec, graph):
    nt, ntdec_s = 20, 3
    nframe = (nt - 1) // ntdec_s + 1
    cfg = small_config(nt=nt)
    m = O.build_model(cfg)
    par = dict(k0_xy=9, i0_yz=17, j0_xz=15, idec=dec[0], jdec=dec[1], kdec=dec[2], ntdec_s=ntdec_s, nframe=nframe)
    run = O.OracleRun(m, 32)
    run.snap_setup(ALL, **par)
    run.run(nt)
    g = gpu_from_model(m)
    g.set_snapshots(ALL, m.kfs, M0=m.M0, UC=m.UC, **par)
    if graph:
        g.step(1, nt, use_graph=True, block=5)
    else:
        g.step(1, nt)
    g.synchronize()
    nonzero = 0
    for name in ALL:
        for f in range(nframe):
            got = g.snapshot(name, f)
            want = run.snap_frames[name][f]
            assert got.shape == want.shape, name
            assert ulp_diff(got, want) == 0, (name, f)
            nonzero += int(np.any(want != 0))
    for pl in ("fs", "ob"):
        assert ulp_diff(g.snap_max(f"{pl}_v"), run.snap_vmax[pl]) == 0, pl
        assert ulp_diff(g.snap_max(f"{pl}_u"), run.snap_umax[pl]) == 0, pl
    assert np.any(run.snap_vmax["fs"] != 0)
    assert nonzero > 0


def test_frame_store_bound_is_checked(orc):
    cfg = small_config(nt=10)
    m = O.build_model(cfg)
    g = gpu_from_model(m)
    g.set_snapshots(["xy_v"], m.kfs, k0_xy=5, ntdec_s=2, nframe=2, M0=m.M0, UC=m.UC)
    g.step(1, 4)        # frames 0, 1
    from openswpc_amd.gpu import SwpcError
    with pytest.raises(SwpcError):
        g.step(5, 1)    # would take frame 2


def test_section_files_from_the_device(orc, tmp_path):
    """snpio.write_sections over the frames a GPU run holds: read back, the
    frames are the oracle's, bit for bit, after the header and medium"""
    from openswpc_amd import snpio
    nt, ntdec_s = 12, 4
    nframe = (nt - 1) // ntdec_s + 1
    cfg = small_config(nt=nt)
    m = O.build_model(cfg)
    par = dict(k0_xy=9, i0_yz=17, j0_xz=15, idec=2, jdec=1, kdec=3, ntdec_s=ntdec_s, nframe=nframe)
    run = O.OracleRun(m, 32)
    run.snap_setup(ALL, **par)
    run.run(nt)
    g = gpu_from_model(m)
    g.set_snapshots(ALL, m.kfs, M0=m.M0, UC=m.UC, **par)
    g.step(1, nt, use_graph=True, block=4)
    g.synchronize()
    geo = snpio.SnapGeometry(nx=cfg.nx, ny=cfg.ny, nz=cfg.nz, na=cfg.na, dx=cfg.dx, dy=cfg.dy, dz=cfg.dz,
                             dt=cfg.dt, xbeg=cfg.xbeg, ybeg=cfg.ybeg, zbeg=cfg.zbeg, idec=2, jdec=1, kdec=3,
                             ntdec_s=ntdec_s, k0_xy=9, i0_yz=17, j0_xz=15)
    paths = snpio.write_sections(g, geo, ALL, nframe, str(tmp_path), "gpu", 0, m.rho, m.lam, m.mu, m.kfs, m.kob)
    g.close()
    assert len(paths) == len(ALL)
    for name, path in zip(ALL, paths):
        r = snpio.read_snp(path)
        assert r["frames"].shape[0] == nframe
        for f in range(nframe):
            assert ulp_diff(r["frames"][f], run.snap_frames[name][f]) == 0, (name, f)


@pytest.mark.parametrize("shear", [-0.5, -2.0, 0.25])
def test_ps_masks_near_epsilon(orc, shear):
    """Shear stresses a few epsilon from zero, of either sign: the yz section
    masks the rotation with |S|/|S + eps| (m_snap.f90:1028-1030), the other
    planes with |S|/(|S| + eps); the device equals the oracle bit for bit."""
    cfg = small_config(nt=1)
    m = O.build_model(cfg)
    par = dict(k0_xy=9, i0_yz=17, j0_xz=15, idec=2, jdec=3, kdec=2, ntdec_s=1, nframe=1)
    run = O.OracleRun(m, 32)
    rng = np.random.default_rng(11)
    eps = np.float32(1.1920928955078125e-07)
    g = gpu_from_model(m)
    for n in ("Vx", "Vy", "Vz"):
        run.arr[n][...] = rng.standard_normal(run.arr[n].shape).astype(np.float32)
    for n in ("Syz", "Sxz", "Sxy"):
        run.arr[n][...] = np.float32(shear) * eps * rng.uniform(0.5, 1.5, run.arr[n].shape).astype(np.float32)
    for n in ("Vx", "Vy", "Vz", "Syz", "Sxz", "Sxy"):
        g.set_field(n, run.arr[n])
    run.snap_setup(ALL, **par)
    run.snap_write(1)
    g.set_snapshots(ALL, m.kfs, M0=m.M0, UC=m.UC, **par)
    g.snap_write(1)
    g.synchronize()
    for pl in ("xy", "yz", "xz", "fs", "ob"):
        got, want = g.snapshot(f"{pl}_ps", 0), run.snap_frames[f"{pl}_ps"][0]
        assert np.any(want[1:] != 0), pl
        assert ulp_diff(got, want) == 0, pl
    g.close()
