"""Native section files (openswpc_amd/snpio.py) on the CPU, with the
oracle's section frames as data.

The reference writer (src/swpc_3d/m_snap.f90) and its reader
(src/shared/m_fdsnap.f90, src/tools/read_snp.f90) all `use netcdf`, which
this image lacks, so neither builds here: the byte layout is pinned by
reading, through
  * the header read back in the field order of the reference reader
    (fdsnap__readhdr, m_fdsnap.f90:195-282) and compared with values worked
    out from the reference's formulas (section size, i2x coordinates,
    absorber widths na/dec, frame interval dt*ntdec_s);
  * the medium block against a per-point loop over the reference's index
    arithmetic (newfile_yz/xz/xy, m_snap.f90:324-473);
  * the frames byte for byte, and the file size.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from openswpc_amd import snpio

ALL = [f"{p}_{q}" for p in ("xy", "yz", "xz", "fs", "ob") for q in ("ps", "v", "u")]


def _cfg():
    return O.Config(nx=30, ny=26, nz=22, na=6, nt=12, dx=0.5, dy=0.4, dz=0.3, dt=0.02, zbeg=-1.2,
                    sources=[dict(x=0.3, y=-0.6, z=2.2, t0=0.0, tr=0.5, mo=1e15, strike=30.0, dip=60.0, rake=90.0)],
                    layers=[[-0.3, 2.7, 5.5, 3.2, 320.0, 160.0], [2.0, 2.9, 6.5, 3.7, 370.0, 185.0]])


@pytest.fixture(scope="module", params=[(1, 1, 1), (2, 3, 2)], ids=["dec1", "dec232"])
def case(orc, request):
    cfg = _cfg()
    m = O.build_model(cfg)
    idec, jdec, kdec = request.param
    nt, ntdec_s = 12, 4
    nframe = (nt - 1) // ntdec_s + 1
    par = dict(k0_xy=7, i0_yz=13, j0_xz=11, idec=idec, jdec=jdec, kdec=kdec, ntdec_s=ntdec_s, nframe=nframe)
    run = O.OracleRun(m, 32)
    run.snap_setup(ALL, **par)
    run.run(nt)
    geo = snpio.SnapGeometry(nx=cfg.nx, ny=cfg.ny, nz=cfg.nz, na=cfg.na, dx=cfg.dx, dy=cfg.dy, dz=cfg.dz, dt=cfg.dt,
                             xbeg=cfg.xbeg, ybeg=cfg.ybeg, zbeg=cfg.zbeg, idec=idec, jdec=jdec, kdec=kdec,
                             ntdec_s=ntdec_s, k0_xy=7, i0_yz=13, j0_xz=11)
    return cfg, m, run, geo, nframe


def _write(tmp_path, cfg, m, run, geo, nframe, name):
    plane, kind = name.split("_")
    path = snpio.file_name(str(tmp_path), "swpc", plane, kind)
    med = snpio.medium_block(geo, plane, m.rho, m.lam, m.mu, m.kfs, m.kob, topo_depth=cfg.layers[0][0])
    with snpio.SnpFile(path, geo, plane, kind, med, "swpc", 1234567890) as f:
        for fr in range(nframe):
            f.write_frame(run.snap_frames[name][fr])
    return path


def _expect_axis(n, dec, beg, d):
    """i2x at cells i*dec - dec/2, one point at a time in single precision"""
    f32 = np.float32
    return [f32(f32(beg) + f32(f32(i * dec - dec // 2) - f32(0.5)) * f32(d)) for i in range(1, n + 1)]


@pytest.mark.parametrize("name", ALL)
def test_header_medium_and_frames(tmp_path, case, name):
    cfg, m, run, geo, nframe = case
    path = _write(tmp_path, cfg, m, run, geo, nframe, name)
    assert os.path.basename(path) == f"swpc.3d.{name.replace('_', '.')}.snp"
    r = snpio.read_snp(path)
    h = r["header"]
    plane, kind = name.split("_")
    idec, jdec, kdec = geo.idec, geo.jdec, geo.kdec
    nxs, nys, nzs = (cfg.nx + idec // 2) // idec, (cfg.ny + jdec // 2) // jdec, (cfg.nz + kdec // 2) // kdec
    ax = {"x": (nxs, idec, cfg.xbeg, cfg.dx, cfg.na // idec), "y": (nys, jdec, cfg.ybeg, cfg.dy, cfg.na // jdec),
          "z": (nzs, kdec, cfg.zbeg, cfg.dz, cfg.na // kdec)}
    a1, a2 = {"yz": "yz", "xz": "xz"}.get(plane, "xy")
    c1, c2 = _expect_axis(*ax[a1][:4]), _expect_axis(*ax[a2][:4])
    assert h["bintype"] == b"STREAMIO" and h["codetype"] == b"SWPC_3D " and h["hdrver"] == 6
    assert h["title"] == "swpc" and h["exedate"] == 1234567890
    assert h["coordinate"] == plane and h["datatype"] == {"ps": "ps", "v": "v3", "u": "u3"}[kind]
    assert (h["ns1"], h["ns2"]) == (ax[a1][0], ax[a2][0])
    assert h["beg1"] == c1[0] and h["beg2"] == c2[0]
    assert h["ds1"] == np.float32(c1[1] - c1[0]) and h["ds2"] == np.float32(c2[1] - c2[0])
    assert h["dt"] == np.float32(np.float32(cfg.dt) * np.float32(4))
    assert (h["na1"], h["na2"]) == (ax[a1][4], ax[a2][4])
    assert h["nmed"] == (3 if plane in ("yz", "xz") else 6)
    assert h["nsnp"] == (4 if kind == "ps" else 3)
    assert h["clon"] == np.float32(139.7604) and h["clat"] == np.float32(35.7182) and h["phi"] == 0.0
    assert h["dum"] == (1.0, 1.0, 1.0)
    # medium: per point, the reference's index arithmetic
    ns1, ns2 = h["ns1"], h["ns2"]
    O2 = O.OFF
    for q, arr in enumerate((m.rho, m.lam, m.mu)):
        want = np.zeros((ns2, ns1), np.float32)
        for b in range(1, ns2 + 1):
            for a in range(1, ns1 + 1):
                if plane == "yz":
                    ii, jj, kk = geo.i0_yz, a * jdec - jdec // 2, b * kdec - kdec // 2
                elif plane == "xz":
                    ii, jj, kk = a * idec - idec // 2, geo.j0_xz, b * kdec - kdec // 2
                else:
                    ii, jj = a * idec - idec // 2, b * jdec - jdec // 2
                    kk = {"fs": int(m.kfs[jj + O2, ii + O2]) + 1, "ob": int(m.kob[jj + O2, ii + O2]) + 1}.get(
                        plane, geo.k0_xy)
                want[b - 1, a - 1] = arr[jj + O2, ii + O2, kk + O2]
        assert np.array_equal(r["medium"][q], want), (name, q)
    if plane not in ("yz", "xz"):
        assert np.all(r["medium"][3] == np.float32(300.0))          # -(-0.3 km) * 1000, positive up
    # frames byte for byte, and nothing else in the file
    assert r["frames"].shape == (nframe, h["nsnp"], ns2, ns1)
    for fr in range(nframe):
        assert r["frames"][fr].tobytes() == np.ascontiguousarray(run.snap_frames[name][fr], np.float32).tobytes()
    nmw = 3 if plane in ("yz", "xz") else 4
    assert os.path.getsize(path) == 176 + 4 * ns1 * ns2 * (nmw + nframe * h["nsnp"])


def test_frames_are_not_all_zero(case):
    _, _, run, _, _ = case
    assert any(np.any(run.snap_frames[n] != 0) for n in ALL)


def test_wrong_frame_shape_is_rejected(tmp_path, case):
    cfg, m, run, geo, nframe = case
    med = snpio.medium_block(geo, "xy", m.rho, m.lam, m.mu, m.kfs, m.kob)
    with snpio.SnpFile(str(tmp_path / "x.snp"), geo, "xy", "v", med, "t", 0) as f:
        with pytest.raises(ValueError):
            f.write_frame(np.zeros((4, 2, 2), np.float32))
