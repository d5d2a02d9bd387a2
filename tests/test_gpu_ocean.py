"""Ocean columns on the GPU against the FP32 oracle, bit for bit.

A seawater layer between sea level and the first layer top (rho 1, mu 0,
lam = 1.5^2, m_vmodel_lhm.f90:111-133) puts the ocean bottom kob below the
free surface kfs (m_medium.f90:339-394), so the second term of the band
select (m_kernel.f90:103-104) switches the stencils to 2nd order in a band
of its own.  The runs cover that band in both phases, 'obb' stations one
cell under the sea floor (m_wav.f90:209), the ob sections and their maximum
maps (m_snap.f90:62,85), kernel__vmax (sampled at kob + 1) and the adjoint.

The select is symmetric in a band's top and bottom ((k-a)(b-k) = (k-b)(a-k)),
so the control replaces the ocean-bottom band by the free-surface band and
must be told apart from the oracle."""
import dataclasses

import numpy as np
import pytest

from oracle import oracle as O
from gpu_helpers import gpu_from_model, ulp_diff

pytestmark = pytest.mark.gpu

OFF = O.OFF
ALL = [f"{p}_{q}" for p in ("xy", "yz", "xz", "fs", "ob") for q in ("ps", "v", "u")]


def ocean_config(**kw):
    # z(k) = -1.5 + (k - 0.5) 0.5: air k <= 3, seawater k = 4..9, solid from k = 10
    base = dict(nx=40, ny=36, nz=32, na=8, nt=40, dx=0.5, dy=0.5, dz=0.5, dt=0.02, zbeg=-1.5,
                layers=[[3.0, 2.7, 5.5, 3.2, 320.0, 160.0], [6.0, 2.9, 6.5, 3.7, 370.0, 185.0]],
                sources=[dict(x=0.3, y=-0.6, z=4.4, t0=0.0, tr=0.6, mo=1e15, strike=30.0, dip=60.0, rake=90.0)],
                stations=[(2.1, 1.3, 0.0, "obb"), (-3.2, 2.7, 0.0, "fsb"), (0.7, -2.9, 6.0, "dep"),
                          (1.1, -0.4, 0.0, "obb")],
                wav_flags=15)
    base.update(kw)
    return O.Config(**base)


def _check_fields(g, run, label):
    for n in O.FIELDS_V + O.FIELDS_S:
        got, want = g.get_field(n), run.field(n)
        assert ulp_diff(got, want) == 0, (label, n, O.rel_l2(got, want))


def test_ocean_bands_are_apart(orc):
    m = O.build_model(ocean_config())
    inner = (slice(OFF + 3, OFF + 34), slice(OFF + 3, OFF + 38))
    assert np.all(m.kfs[inner] == 3) and np.all(m.kob[inner] == 9)
    assert np.all(m.kob_top[inner] == 7) and np.all(m.kfs_bot[inner] == 5)
    assert list(m.kst[[0, 3]]) == [10, 10] and m.kst[1] == 4


@pytest.mark.parametrize("nm,abc,graph", [(3, "pml", False), (0, "pml", True), (2, "cerjan", False)])
def test_ocean_steps_match_oracle(orc, nm, abc, graph):
    cfg = ocean_config(nm=nm, abc=abc)
    m = O.build_model(cfg)
    run = O.OracleRun(m, 32)
    run.run(cfg.nt)
    g = gpu_from_model(m)
    if graph:
        g.step(1, cfg.nt, use_graph=True, block=7)
    else:
        g.step(1, cfg.nt)
    g.synchronize()
    _check_fields(g, run, f"ocean nm={nm} {abc}")
    for kind in ("vel", "disp", "stress", "strain"):
        got, want = g.seismograms(kind), run.wav[kind]
        assert ulp_diff(got, want) == 0, (kind, O.rel_l2(got, want))
    # the wave has crossed the sea floor: water cells and obb stations move
    vz = run.field("Vz")[OFF + 3:OFF + 34, OFF + 3:OFF + 38, OFF + 4:OFF + 10]
    assert np.any(vz != 0)
    assert np.any(run.wav["vel"][0] != 0)
    assert np.array_equal(g.kernel_vmax(), run.vmax())
    assert np.any(run.vmax() != 0)
    g.close()


def test_ocean_sections_match_oracle(orc):
    nt, ntdec_s = 30, 4
    nframe = (nt - 1) // ntdec_s + 1
    cfg = ocean_config(nt=nt)
    m = O.build_model(cfg)
    par = dict(k0_xy=7, i0_yz=19, j0_xz=17, idec=2, jdec=1, kdec=1, ntdec_s=ntdec_s, nframe=nframe)
    run = O.OracleRun(m, 32)
    run.snap_setup(ALL, **par)
    run.run(nt)
    g = gpu_from_model(m)
    g.set_snapshots(ALL, m.kfs, M0=m.M0, UC=m.UC, **par)
    g.step(1, nt, use_graph=True, block=5)
    g.synchronize()
    for name in ALL:
        for f in range(nframe):
            assert ulp_diff(g.snapshot(name, f), run.snap_frames[name][f]) == 0, (name, f)
    for pl in ("fs", "ob"):
        assert ulp_diff(g.snap_max(f"{pl}_v"), run.snap_vmax[pl]) == 0, pl
        assert ulp_diff(g.snap_max(f"{pl}_u"), run.snap_umax[pl]) == 0, pl
    # the ob maps sample below the sea floor, the fs maps the water surface
    assert np.any(run.snap_vmax["ob"] != 0) and np.any(run.snap_vmax["fs"] != 0)
    assert not np.array_equal(run.snap_vmax["ob"], run.snap_vmax["fs"])
    g.close()


def test_dropped_ocean_bottom_band_is_detected(orc):
    """control: the GPU given the free-surface band in place of the
    ocean-bottom band differs from the oracle, and only at the sea floor's
    depth at first"""
    cfg = ocean_config(nt=30, nm=0)
    m = O.build_model(cfg)
    run = O.OracleRun(m, 32)
    run.run(cfg.nt)
    wrong = dataclasses.replace(m, kob_top=m.kfs_top.copy(), kob_bot=m.kfs_bot.copy())
    g = gpu_from_model(wrong)
    g.step(1, cfg.nt)
    g.synchronize()
    diff = [O.rel_l2(g.get_field(n), run.field(n)) for n in O.FIELDS_V + O.FIELDS_S]
    print("rel L2 with the band dropped", diff, flush=True)
    assert max(diff) > 1e-5
    g.close()
    g = gpu_from_model(m)
    g.step(1, cfg.nt)
    g.synchronize()
    _check_fields(g, run, "ocean elastic")
    g.close()


def test_ocean_adjoint_matches_fp64_oracle(orc):
    """lam/mu/rho gradients through the ocean-bottom band, against the FP64
    oracle adjoint (itself pinned by the dot-product test with an ocean)"""
    nt = 40
    cfg = O.Config(nx=22, ny=20, nz=18, na=5, nm=0, dt=0.02, dx=0.5, dy=0.5, dz=0.5, zbeg=-1.0, nt=nt,
                   layers=[[1.5, 2.5, 5.0, 2.9, 1e6, 1e6], [4.0, 2.8, 6.2, 3.6, 1e6, 1e6]],
                   sources=[dict(x=0.2, y=-0.4, z=3.6, t0=0.0, tr=0.5, mo=1.0, strike=30.0, dip=60.0, rake=90.0)],
                   stations=[(1.6, 1.1, 0.0, "obb"), (-2.1, 0.7, 0.0, "fsb"), (0.4, -1.9, 0.0, "obb"),
                             (-0.9, -0.6, 2.0, "dep")], wav_flags=1)
    m = O.build_model(cfg)
    assert np.any(m.kob != m.kfs)
    rng = np.random.default_rng(5)
    adj = O.OracleAdjoint(m, 64)
    d = adj.forward(nt, keep=False)
    d_obs = d * (1 + 0.3 * rng.standard_normal(d.shape))
    J, (gl, gm, gr), _ = O.misfit_and_gradient(m, nt, d_obs)
    obs = np.ascontiguousarray(d_obs.transpose(1, 2, 0), np.float32)
    gpu = gpu_from_model(m)
    Jg, hl, hm, hr = gpu.adjoint_gradient(nt, obs, ckp=9)
    gpu.close()
    assert abs(Jg - J) <= 1e-4 * abs(J), (Jg, J)
    sl = (slice(OFF + 1, -OFF - 3), slice(OFF + 1, -OFF - 3), slice(OFF + 1, -OFF - 3))
    for name, got, want in (("lam", hl, gl), ("mu", hm, gm), ("rho", hr, gr)):
        w = want[sl].astype(np.float64)
        rel = np.linalg.norm(got[sl].astype(np.float64) - w) / np.linalg.norm(w)
        print(name, rel, flush=True)
        assert rel <= 1e-4, (name, rel)
