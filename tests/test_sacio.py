"""SAC, CSF and tar writers (openswpc_amd/sacio.py) byte for byte against
files written by the reference's own m_sac.f90/m_tar.f90
(tests/golden/sac/, made by tests/golden/make_sac_golden.py), plus readers
independent of both: numpy for the SAC words and Python's tarfile for the
archives."""
import io
import os
import tarfile

import numpy as np
import pytest

from openswpc_amd import sacio

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sac")


def _data():
    i = np.arange(1, 38, dtype=np.float32)
    d1 = i * np.float32(0.25) - np.float32(3.0)
    d3 = np.stack([d1, i * np.float32(1.0e-3), -(i * i) * np.float32(7.5e-5)])
    d98 = np.arange(1, 99, dtype=np.float32) * np.float32(0.5) - np.float32(20.0)
    return d1, d3, d98


def _full():
    """the header of sac_driver.f90's full.sac"""
    h = sacio.SacHeader(
        delta=0.15000000000000002, depmin=-3.0, depmax=6.25, b=-0.1, e=5.3, o=1.0 / 3.0, a=2.7182818284590452,
        t0=0.1, t1=0.2, t2=0.3, t3=0.4, t4=0.5, t5=0.6, t6=0.7, t7=0.8, t8=0.9, t9=1.1,
        stla=35.123456789, stlo=139.87654321, stel=12.5, stdp=1500.0,
        evla=41.5, evlo=142.0, evel=-3.0, evdp=30.0, mag=6.3333333333,
        user0=1.0e-1, user1=-2.0e-2, user2=3.0e-3, user3=4.0e4, user4=-5.5, user5=6.6, user6=141.4, user7=41.1,
        user8=0.0, user9=9.9, dist=123.456, az=45.6789, baz=225.6789, gcarc=1.2345,
        depmen=0.125, cmpaz=90.0, cmpinc=90.0,
        user10=10.0, user11=11.0, user12=12.0, user13=13.0, user14=14.0, user15=15.0, user16=16.0,
        nzyear=2025, nzjday=123, nzhour=4, nzmin=5, nzsec=6, nzmsec=7, npts=37, idep=7, ievtyp=40,
        iuser0=100, iuser1=101, iuser2=102, iuser3=103, iuser4=104, iuser5=105, iuser6=106, iuser7=107,
        leven=True, lpspol=True, lovrok=False, lcalda=True, luser0=True,
        kstnm="N.TMRF", kevnm="swpc_example_run", khole="H1", ko="ORIGIN", ka="P", kt0="T0", kt1="T1",
        kt2="T2", kt3="T3", kt4="T4", kt5="T5", kt6="T6", kt7="T7", kt8="T8", kt9="T9", kf="FIN",
        kuser0="U0", kuser1="U1", kuser2="U2", kcmpnm="Vx", knetwk="NET", kdatrd="DATRD", kinst="SWPC",
        tim=1735657200)
    return h


def _gold(name):
    with open(os.path.join(GOLD, name), "rb") as f:
        return f.read()


def test_full_header_and_samples(tmp_path):
    d1, _, _ = _data()
    p = tmp_path / "full.sac"
    sacio.write_sac(str(p), _full(), d1)
    assert p.read_bytes() == _gold("full.sac")


def test_default_header(tmp_path):
    d1, _, _ = _data()
    p = tmp_path / "default.sac"
    sacio.write_sac(str(p), sacio.SacHeader(delta=0.05, npts=37, b=0.0), d1)
    assert p.read_bytes() == _gold("default.sac")


def test_csf(tmp_path):
    _, d3, _ = _data()
    h = _full()
    hs = [h, h.copy(kcmpnm="Vy", cmpaz=180.0), h.copy(kcmpnm="Vz", cmpinc=0.0, idep=6)]
    p = tmp_path / "three.csf"
    sacio.write_csf(str(p), hs, d3)
    assert p.read_bytes() == _gold("three.csf")


def test_tar_with_a_whole_block_member(tmp_path):
    d1, _, d98 = _data()
    body = (sacio.tar_member("swpc.N.TMRF.3d.Vx.sac", _full(), d1)
            + sacio.tar_member("swpc.ST02.3d.Uz.sac", sacio.SacHeader(delta=0.05, npts=98, b=0.0, tim=1700000123),
                               d98)
            + sacio.TAR_END)
    assert body == _gold("two.tar")


def test_sac_words_read_back_independently():
    """the golden file itself, read with numpy: word positions of sac__whdr"""
    b = _gold("full.sac")
    f = np.frombuffer(b[:280], np.float32)
    i = np.frombuffer(b[280:420], np.int32)
    lg = np.frombuffer(b[420:440], np.int32)
    assert f[0] == np.float32(np.float32(1500000) / np.float32(1e7))     # delta cut to 1e-7
    assert f[31] == np.float32(35.123456789) and f[57] == np.float32(90.0)  # stla (32), cmpaz (58)
    assert f[3] == np.float32(-12345.0)                                     # scale is never written
    assert list(i[:7]) == [2025, 123, 4, 5, 6, 7, 6] and i[9] == 37        # nz*, nvhdr, npts (80)
    assert list(lg) == [1, 1, 0, 1, 1]
    assert b[440:448] == b"N.TMRF  " and b[448:464] == b"swpc_example_run"
    assert np.array_equal(np.frombuffer(b[632:], np.float32), _data()[0])


def test_tar_is_readable_by_tarfile():
    t = tarfile.open(fileobj=io.BytesIO(_gold("two.tar")))
    names = t.getnames()
    assert names == ["swpc.N.TMRF.3d.Vx.sac", "swpc.ST02.3d.Uz.sac"]
    m = t.getmember(names[1])
    assert m.size == (158 + 98) * 4 and m.mtime == 1700000123 and m.mode == 0o644
    raw = t.extractfile(m).read()
    assert np.array_equal(np.frombuffer(raw[632:], np.float32), _data()[2])


def test_driver_regenerates_the_fixtures(tmp_path):
    """where the reference is mounted, its writer still produces the committed files"""
    if not os.path.isdir("/root/reference/src/shared"):
        pytest.skip("reference sources not mounted")
    from tests.golden.make_sac_golden import FILES, run_driver
    run_driver(str(tmp_path))
    for fn in FILES:
        assert (tmp_path / fn).read_bytes() == _gold(fn), fn


def test_station_files_round_trip(tmp_path):
    run = sacio.RunInfo(title="swpc", exedate=1735657200, tbeg=0.0, dt=0.02, ntdec_w=5, ntw=11, M0=1e15,
                        evlo=142.0, evla=41.5, evdp=30.0, sx0=0.5, sy0=-0.25, mij0=(1, 0, -1, 0.5, 0, 0.25))
    stations = [("ST01", 10.0, 0.0, 0.0), ("ST02", 0.5, 7.0, 1.5)]
    rng = np.random.default_rng(3)
    tr = {"vel": rng.standard_normal((2, 3, 11)).astype(np.float32),
          "disp": rng.standard_normal((2, 3, 11)).astype(np.float32)}
    files = sacio.write_wav(str(tmp_path), run, stations, tr, "sac")
    assert len(files) == 12
    b = open(os.path.join(tmp_path, "wav", "swpc.3d.ST02.Uy.sac"), "rb").read()
    assert np.array_equal(np.frombuffer(b[632:], np.float32), tr["disp"][1, 1])
    f = np.frombuffer(b[:280], np.float32)
    # delta = ntdec_w * dt in single precision (5 * 0.02 ties down to
    # 0.099999994), then cut to 1e-7 by sac__whdr: 0.0999999
    d = np.float32(5) * np.float32(0.02)
    assert f[0] == np.float32(np.float32(int(float(d) * 1e7)) / np.float32(1e7)) == np.float32(0.0999999)
    assert abs(f[50] - np.hypot(0.0, 7.25)) < 1e-5               # dist from the source
    assert abs(f[51] - 90.0) < 1e-4                              # due +y: azimuth 90 in x-east axes
    assert np.frombuffer(b[280:420], np.int32)[16] == 6          # idep: displacement
    # tar per station holds the six traces in kind order
    sacio.write_wav(str(tmp_path), run, stations, tr, "tar_st")
    t = tarfile.open(os.path.join(tmp_path, "wav", "swpc.3d.ST01.sac.tar"))
    assert t.getnames() == [f"swpc.ST01.3d.{c}.sac" for c in ("Vx", "Vy", "Vz", "Ux", "Uy", "Uz")]


def test_green_files_flip_z(tmp_path):
    run = sacio.RunInfo(title="rcpr", exedate=1735657200, tbeg=0.0, dt=0.02, ntdec_w=5, ntw=7, M0=1.0)
    gf = np.arange(2 * 9 * 7, dtype=np.float32).reshape(18, 7)
    files = sacio.write_green(str(tmp_path), run, "N.TMRF", "z", gf, [1, 42], [(1, 2, 3), (4, 5, 6)],
                              [(1.25, 2.25, 3.25), (4.25, 5.25, 6.25)])
    assert len(files) == 18
    fn = os.path.join(tmp_path, "green", "N.TMRF", "rcpr__00000042__N.TMRF__z__fy___.sac")
    b = open(fn, "rb").read()
    assert np.array_equal(np.frombuffer(b[632:], np.float32), -gf[9 + 7])
    assert b[600:608] == b"G_Vz_fy_"                                  # kcmpnm (words 151-152)
    assert b[448:456] == b"00000042"                                  # kevnm = GF id
