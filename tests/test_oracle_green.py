"""Pins of the oracle's reciprocity Green's-function mode (orc_green_store,
orc_green_source; src/swpc_3d/m_green.f90) by discrete reciprocity, FP64.

The velocity-stress scheme is reciprocal as a discrete system: with the
stress update sigma += dt C D v and the velocity update v += dt B (-D^T) sigma,
the response from a stress source P^T m at one cell to the velocity sample
R v at another equals, for the same moment-rate sequence, the strain sample
P D v at the first cell of a run driven by the force B R^T f at the second.
So for one station and one GF cell, with unit moment rate s(it):

  * six forward runs with a unit moment-tensor component M_pq at the GF cell
    give at the station v_n(it) = gf_pq^(n)(it), the strain Green's function
    of the reciprocal run with a unit force along n at the station
    (the example/green/input_green_fwd.M*.inf against input_green_rcpr.*.inf
    check of the reference, made exact);
  * three forward runs with a unit force f_p at the GF cell give the
    station displacement u_n(it) = gf_(6+p)^(n)(it) (green_bforce).

This ties green__store's nine staggered derivatives and its four-point shear
means to source__stressglut's /4 shear split (their transposes), and
green__source to the receiver sampling of wav__store.  Two conditions make
the identity exact rather than approximate: both runs use the same moment-rate
sequence (the forward stress source is normally sampled half a step earlier,
m_source.f90:798), and the window ends before any wave reaches the absorber
or the free surface, whose discrete operators are not symmetric.
"""
import numpy as np
import pytest

from oracle import oracle as O

NT = 70
TRISE = 0.3
GF_XYZ = (1.5, 1.0, 8.5)       # GF cell = forward source cell
ST_XYZ = (-1.0, -0.5, 7.0)     # station = virtual source of the reciprocal runs
MT = ("mxx", "myy", "mzz", "myz", "mxz", "mxy")
UC = np.float32(1.0000001095e-15)
UC_BF = np.float32(9.999999960041972e-13)


def _cfg(**kw):
    base = dict(nx=40, ny=40, nz=40, na=8, nm=0, nt=NT, dt=0.02, dx=0.5, dy=0.5, dz=0.5, zbeg=-1.5,
                layers=[[0.0, 2.5, 5.0, 2.9, 1e6, 1e6], [7.8, 2.8, 6.0, 3.5, 1e6, 1e6]],
                stations=[(*ST_XYZ, "dep")], wav_flags=3, ntdec_w=1)
    base.update(kw)
    return O.Config(**base)


def _forward(source, bf, shift_i=0):
    src = dict(x=GF_XYZ[0] + 0.5 * shift_i, y=GF_XYZ[1], z=GF_XYZ[2], t0=0.0, tr=TRISE)
    src.update(source)
    cfg = _cfg(sources=[src], bf_mode=bf)
    m = O.build_model(cfg)
    m.stf = np.ascontiguousarray(O.green_stf(cfg, TRISE)[:, None])   # the reciprocal runs' sequence
    run = O.OracleRun(m, 64)
    run.run(NT)
    # station samples in grid axes (wav__store flips z to positive up)
    sign = np.array([1.0, 1.0, -1.0])[:, None]
    v = run.wav["vel"][0] * sign / (np.float64(m.M0) * np.float64(m.UC) * 1e9)
    u = run.wav["disp"][0] * sign / (np.float64(m.M0) * np.float64(m.UC) * 1e9)   # m.UC is x1000 for forces
    return m, v, u


def _reciprocal(cmp):
    cfg = _cfg()
    m = O.build_model(cfg)
    run = O.OracleRun(m, 64)
    i, j, k = (O.lib(32).orc_x2i(c, b, np.float32(0.5)) for c, b in zip(GF_XYZ, (cfg.xbeg, cfg.ybeg, cfg.zbeg)))
    run.green_setup([i], [j], [k], m.ist[0], m.jst[0], m.kst[0], cmp, True, 1, m.ntw, O.green_stf(cfg, TRISE))
    run.run(NT)
    gf = run.gf.copy()
    gf[:6] /= np.float64(UC) * 1e9
    gf[6:] /= np.float64(UC_BF) * 1e9
    return m, gf


@pytest.fixture(scope="module")
def rcpr(orc):
    return {c: _reciprocal(c)[1] for c in "xyz"}


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("c", range(6))
def test_moment_tensor_reciprocity(rcpr, c):
    unit = {n: 0.0 for n in MT}
    unit[MT[c]] = 1.0
    _, v, _ = _forward(dict(mo=1.0, **unit), bf=False)
    for n, cmp in enumerate("xyz"):
        g = rcpr[cmp][c]
        assert np.max(np.abs(g)) > 0
        rel = _rel(v[n], g)
        print(MT[c], cmp, rel, flush=True)
        assert rel <= 1e-9, (MT[c], cmp, rel)


@pytest.mark.parametrize("p", range(3))
def test_body_force_reciprocity(rcpr, p):
    f = [0.0, 0.0, 0.0]
    f[p] = 1.0
    _, _, u = _forward(dict(fx=f[0], fy=f[1], fz=f[2]), bf=True)
    for n, cmp in enumerate("xyz"):
        g = rcpr[cmp][6 + p]
        rel = _rel(u[n], g)
        print("f" + "xyz"[p], cmp, rel, flush=True)
        assert rel <= 1e-9, ("xyz"[p], cmp, rel)


def test_reciprocity_detects_a_misplaced_shear_split(rcpr):
    """control: the shear source one cell off in x breaks the identity"""
    unit = {n: 0.0 for n in MT}
    unit["mxy"] = 1.0
    _, v, _ = _forward(dict(mo=1.0, **unit), bf=False, shift_i=1)
    assert _rel(v[0], rcpr["x"][5]) > 1e-2


def test_store_decimation_and_body_force_columns(orc):
    """gf columns: ncmp = 6 without body force; with ntdec_w = 4 column itw
    holds the accumulators after steps 1 + 4*itw (m_green.f90:497-543)"""
    cfg = _cfg(nt=20)
    m = O.build_model(cfg)
    st = O.green_stf(cfg, TRISE)
    ig, jg, kg = [24, 20], [23, 21], [20, 18]
    full = O.OracleRun(m, 32)
    full.green_setup(ig, jg, kg, m.ist[0], m.jst[0], m.kst[0], "y", True, 1, 20, st)
    full.run(20)
    dec = O.OracleRun(m, 32)
    ntw = (20 - 1) // 4 + 1
    dec.green_setup(ig, jg, kg, m.ist[0], m.jst[0], m.kst[0], "y", False, 4, ntw, st)
    dec.run(20)
    assert dec.gf.shape == (2 * 6, ntw)
    for n in range(2):
        for c in range(6):
            assert np.array_equal(dec.gf[n * 6 + c], full.gf[n * 9 + c, ::4]), (n, c)
