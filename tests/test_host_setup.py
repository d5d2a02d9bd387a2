"""The product's Fortran set-up (libswpc_host) against the oracle's C
restatement: two independent restatements of the reference set-up must
produce the same kernel inputs.  Integer maps and everything outside the
complex arithmetic of relaxed_medium are compared bit for bit; the relaxed
moduli (complex sqrt/division in the two runtimes) to a few ulp."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_helpers import ulp_diff


@pytest.fixture(scope="module")
def hostmod():
    import subprocess, os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    subprocess.run(["make", "-s", "-C", os.path.join(root, "openswpc_amd", "host")], check=True)
    from openswpc_amd import host
    return host


def _pair(orc, hostmod, **kw):
    base = dict(nx=40, ny=36, nz=32, na=8, nt=50, dt=0.02, zbeg=-2.0,
                layers=[[0.0, 2.7, 5.5, 3.2, 320.0, 160.0], [5.0, 2.9, 6.5, 3.7, 370.0, 185.0]],
                sources=[dict(x=0.3, y=-0.6, z=5.2, t0=0.05, tr=0.8, mo=1e15, strike=30.0, dip=60.0, rake=90.0),
                         dict(x=-1.3, y=1.1, z=7.7, t0=0.2, tr=1.3, mo=4e14, strike=200.0, dip=20.0, rake=-40.0)],
                stations=[(2.1, 1.3, 0.0, 'fsb'), (-3.2, 2.7, 0.0, 'fsb'), (0.7, -2.9, 6.0, 'dep'), (99.0, 0.0, 0.0, 'fsb')])
    base.update(kw)
    cfg = O.Config(**base)
    m = O.build_model(cfg)
    hk = {k: v for k, v in base.items() if k not in ("wav_flags",)}
    spec = hostmod.ModelSpec(**hk)
    inp = hostmod.build_inputs(spec)
    return m, inp


@pytest.mark.parametrize("kw", [dict(), dict(nm=2, vmodel="lgm", layers=[[0.0, 2.3, 3.5, 2.0, 200, 100],
                                                                          [12.0, 3.0, 7.8, 4.5, 450, 225]]),
                                dict(abc="cerjan"), dict(nm=0),
                                # an ocean column: 1.5 km of air and 3 km of seawater over the solid
                                dict(zbeg=-1.5, layers=[[3.0, 2.7, 5.5, 3.2, 320.0, 160.0],
                                                        [6.0, 2.9, 6.5, 3.7, 370.0, 185.0]],
                                     stations=[(2.1, 1.3, 0.0, 'obb'), (-3.2, 2.7, 0.0, 'fsb'),
                                               (0.7, -2.9, 6.0, 'dep')]),
                                dict(zbeg=-0.5, munk=True, vmodel="lgm",
                                     layers=[[2.5, 2.3, 3.5, 2.0, 200, 100], [12.0, 3.0, 7.8, 4.5, 450, 225]],
                                     stations=[(2.1, 1.3, 0.0, 'obb'), (0.7, -2.9, 0.0, 'obb')])])
def test_host_matches_oracle(orc, hostmod, kw):
    m, inp = _pair(orc, hostmod, **kw)
    a = inp.arrays
    for n in ("kfs", "kob", "kfs_top", "kfs_bot", "kob_top", "kob_bot", "isrc", "jsrc", "ksrc", "ist", "jst", "kst"):
        assert np.array_equal(a[n], getattr(m, n)), n
    for n in ("rho", "taup", "taus", "ts", "c1", "c2", "d1", "stf", "mo"):
        assert ulp_diff(a[n], getattr(m, n)[: a[n].size].reshape(a[n].shape)) == 0, n
    assert np.array_equal(a["comp"], m.mij)
    for n in ("lam", "mu"):
        want = getattr(m, n).astype(np.float64)
        rel = np.abs(a[n] - want) / np.maximum(np.abs(want), 1e-30)
        assert float(rel.max()) <= 2e-6, (n, float(rel.max()))
    sc = inp.scalars
    assert np.float32(sc["zeta"]) == np.float32(m.zeta)
    assert np.float32(sc["d2"]) == np.float32(m.d2)
    assert sc["M0"] == m.M0 and sc["UC"] == m.UC and sc["ntw"] == m.ntw
    if m.cfg.abc == "pml":
        for n in ("gxc", "gxe", "gyc", "gye", "gzc", "gze"):
            assert ulp_diff(a[n], getattr(m, n)) == 0, n
    else:
        for n in ("cgx_c", "cgx_b", "cgy_c", "cgy_b", "cgz_c", "cgz_b"):
            assert ulp_diff(a[n], getattr(m, n)) == 0, n


def test_host_seawater_velocity(orc, hostmod):
    """both restatements of seawater__vel agree bit for bit (the oracle's is
    pinned by the reference's own m_seawater.f90)"""
    L = hostmod.host_lib()
    for munk in (0, 1):
        for z in np.linspace(0.0, 7.0, 57, dtype=np.float32):
            assert np.float32(L.swpch_seawater_vel(z, munk)) == np.float32(O.lib(32).orc_seawater_vel(z, munk))
