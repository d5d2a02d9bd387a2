"""GPU parity of the reciprocity Green's-function mode: libswpc_gpu's
green__store and green__source (m_green.f90) against the FP32 oracle, bit
for bit, in direct steps, graph replays, routine-by-routine calls and a batch
of reciprocal runs (the three force directions at one station)."""
import numpy as np
import pytest

from oracle import oracle as O
from gpu_helpers import compare_fields, gpu_from_model, small_config, ulp_diff
from openswpc_amd.gpu import SwpcBatch

pytestmark = pytest.mark.gpu

TRISE = 0.6


def _green_cfg(**kw):
    base = dict(sources=[], nt=36)
    base.update(kw)
    return small_config(**base)


def _points(cfg, m):
    """GF cells: interior, the first solid cell under the free surface, in the
    x and bottom absorbers, and next to the grid edges"""
    nx, ny, nz, na = cfg.nx, cfg.ny, cfg.nz, cfg.na
    ij = [(20, 17), (11, 25), (3, 9), (nx - 1, ny - 2), (1, ny), (nx, 1), (16, 14)]
    ks = [12, None, 9, 5, 6, nz, nz - 2]
    ig, jg, kg = [], [], []
    for (i, j), k in zip(ij, ks):
        ig.append(i); jg.append(j)
        kg.append(int(m.kob[j + O.OFF, i + O.OFF]) if k is None else k)
    return np.array(ig, np.int32), np.array(jg, np.int32), np.array(kg, np.int32)


def _pair(cfg, cmp, bforce, ntdec_w, station=0):
    m = O.build_model(cfg)
    ig, jg, kg = _points(cfg, m)
    ntw = (cfg.nt - 1) // ntdec_w + 1
    stf = O.green_stf(cfg, TRISE)
    run = O.OracleRun(m, 32)
    src = (m.ist[station], m.jst[station], m.kst[station])
    run.green_setup(ig, jg, kg, *src, cmp, bforce, ntdec_w, ntw, stf)
    g = gpu_from_model(m)
    g.set_green(ig, jg, kg, *src, cmp, bforce, ntdec_w, ntw, stf)
    return m, run, g


def _check(g, run, label):
    for n, (ulp, rel) in compare_fields(g, run).items():
        assert ulp == 0, (label, n, ulp, rel)
    got, want = g.green(), run.gf
    assert got.shape == want.shape, label
    assert np.any(want != 0), label
    assert ulp_diff(got, want) == 0, (label, O.rel_l2(got, want))


@pytest.mark.parametrize("cmp", ["x", "y", "z"])
def test_green_direct_steps(orc, cmp):
    cfg = _green_cfg()
    m, run, g = _pair(cfg, cmp, True, 1)
    run.run(cfg.nt)
    g.step(1, cfg.nt)
    _check(g, run, f"green {cmp}")
    g.close()


def test_green_graph_replay_decimated_no_body_force(orc):
    cfg = _green_cfg(nm=2)
    m, run, g = _pair(cfg, "z", False, 4, station=2)
    run.run(cfg.nt)
    g.step(1, 20, use_graph=True, block=6)
    g.step(21, cfg.nt - 20, use_graph=True, block=5)
    _check(g, run, "graph z")
    g.close()


def test_green_routine_by_routine_cerjan(orc):
    cfg = _green_cfg(abc="cerjan", nt=20)
    m, run, g = _pair(cfg, "x", True, 2)
    run.run(cfg.nt)
    for it in range(1, cfg.nt + 1):
        g.green_store(it)
        g.wav_store(it)
        g.kernel_update_stress()
        g.absorb_update_stress()
        g.source_stressglut(it)
        g.kernel_update_vel()
        g.absorb_update_vel()
        g.source_bodyforce(it)
        g.green_source(it)
    g.synchronize()
    _check(g, run, "cerjan x")
    g.close()


def test_green_mode_off_and_reset(orc):
    """cmp = None switches the mode off: steps then match a plain run; setting
    it again restarts the accumulators and outputs from zero"""
    cfg = _green_cfg(nt=12)
    m, run, g = _pair(cfg, "y", True, 1)
    g.set_green(None, None, None, 1, 1, 1, None, False, 1, 1, [])
    plain = O.OracleRun(m, 32)
    plain.run(6)
    g.step(1, 6)
    for n, (ulp, rel) in compare_fields(g, plain).items():
        assert ulp == 0, (n, ulp)
    # on again from step 7: zero accumulators, outputs sampled from step 7
    ig, jg, kg = _points(cfg, m)
    stf = O.green_stf(cfg, TRISE)
    plain.green_setup(ig, jg, kg, m.ist[1], m.jst[1], m.kst[1], "z", True, 1, cfg.nt, stf)
    plain.run(6)
    g.set_green(ig, jg, kg, m.ist[1], m.jst[1], m.kst[1], "z", True, 1, cfg.nt, stf)
    g.step(7, 6)
    _check(g, plain, "restart z")
    g.close()


def test_batch_of_reciprocal_runs(orc):
    """the three force directions at one station as one batch, plus a member
    with a different grid and no Green's functions"""
    nt = 16
    cfgs = [_green_cfg(nt=nt), _green_cfg(nt=nt), _green_cfg(nt=nt),
            small_config(nx=37, ny=41, nz=29, na=6, nt=nt)]
    runs, gs = [], []
    for cmp, cfg in zip(("x", "y", "z", None), cfgs):
        if cmp is None:
            m = O.build_model(cfg)
            run, g = O.OracleRun(m, 32), gpu_from_model(m)
        else:
            m, run, g = _pair(cfg, cmp, True, 3)
        run.run(nt)
        runs.append(run)
        gs.append(g)
    b = SwpcBatch(gs)
    b.step(1, nt, use_graph=True, block=5)
    b.synchronize()
    for n, (g, run) in enumerate(zip(gs, runs)):
        if n < 3:
            _check(g, run, f"batch member {n}")
        else:
            for name, (ulp, rel) in compare_fields(g, run).items():
                assert ulp == 0, (n, name, ulp)
    b.close()
    for g in gs:
        g.close()
