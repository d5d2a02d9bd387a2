"""Batches of uneven simulations stepped together (swpc_batch_*): every
member must end bit-identical to its own oracle run, with and without graph
replays, with Cerjan and PML members mixed, under the three member packings,
and after the autotuner's dry trials."""
import numpy as np
import pytest

from oracle import oracle as O
from gpu_helpers import compare_fields, gpu_from_model, small_config, ulp_diff
from openswpc_amd.gpu import SwpcBatch

pytestmark = pytest.mark.gpu


def _members():
    mt = dict(x=0.3, y=-0.6, z=4.2, t0=0.05, tr=0.8, mo=1e15, strike=30.0, dip=60.0, rake=90.0)
    bf = dict(x=-0.7, y=0.4, z=3.3, t0=0.0, tr=0.6, fx=0.3, fy=-1.0, fz=0.5)
    return [
        small_config(nx=40, ny=36, nz=32, na=8, nm=2, sources=[mt]),
        small_config(nx=37, ny=41, nz=29, na=6, nm=2, sources=[bf], bf_mode=True),
        small_config(nx=48, ny=34, nz=36, na=10, nm=2, sources=[mt, dict(mt, x=-1.1, strike=120.0)]),
        small_config(nx=33, ny=30, nz=27, na=5, nm=2, abc="cerjan", sources=[bf], bf_mode=True),
    ]


def _check_members(gs, runs):
    for n, (g, run) in enumerate(zip(gs, runs)):
        res = compare_fields(g, run)
        for name, (ulp, rel) in res.items():
            assert ulp == 0, (n, name, ulp, rel)
        assert ulp_diff(g.seismograms("vel"), run.wav["vel"]) == 0, n
        assert np.any(run.field("Vz") != 0), n


@pytest.mark.parametrize("graph,variant,packing", [(False, 1, 0), (True, 1, 0), (True, 5, 0), (True, 13, 0),
                                                   (True, 9, 1), (True, 9, 2), (False, 13, 2)])
def test_batch_members_match_their_oracles(orc, graph, variant, packing):
    nt = 18
    cfgs = _members()
    models = [O.build_model(c) for c in cfgs]
    runs = []
    for m in models:
        r = O.OracleRun(m, 32)
        r.run(nt)
        runs.append(r)
    gs = [gpu_from_model(m, variant=variant) for m in models]
    b = SwpcBatch(gs)
    b.set_packing(packing)
    if graph:
        b.step(1, 12, use_graph=True, block=4)
        b.step(13, nt - 12, use_graph=True, block=3)
    else:
        b.step(1, nt)
    b.synchronize()
    _check_members(gs, runs)
    b.close()


def test_tuner_dry_trials_leave_state_untouched(orc):
    nt = 10
    cfgs = _members()[:3]
    models = [O.build_model(c) for c in cfgs]
    runs = []
    for m in models:
        r = O.OracleRun(m, 32)
        r.run(nt)
        runs.append(r)
    gs = [gpu_from_model(m) for m in models]
    b = SwpcBatch(gs)
    kc, by, width, ms = b.tune(trial_steps=2)
    assert kc > 0 and by in (2, 4) and width in (32, 64) and ms > 0
    for g in gs:
        for name in ("Vx", "Sxx"):
            assert not np.any(g.get_field(name))
    b.step(1, nt, use_graph=True, block=5)
    b.synchronize()
    _check_members(gs, runs)
    b.close()


def test_full_greens_batch_matches_members_stepped_alone():
    """The benchmark's 64-member Green's-function batch at full size
    (302.6 M cells, grids 96..256 x 96..256 x 96..192, NM = 2, 10-cell PML,
    body forces, 8 receivers each), autotuned, as graph replays: every
    member's fields and traces equal, bit for bit, those of the same member
    stepped alone (the single-context path the tests above pin to the
    oracle).  20 steps carry each force's field into its free surface and,
    on the narrow members, into the side PML."""
    from openswpc_amd import workloads as bench
    from openswpc_amd.host import build_inputs, make_simulation

    nt = 20
    specs = bench.batch_specs(nt=nt + 1)
    batched = [make_simulation(build_inputs(sp)) for sp in specs]
    alone = [make_simulation(build_inputs(sp)) for sp in specs]
    b = SwpcBatch(batched)
    try:
        kc, by, width, ms = b.tune(trial_steps=1)
        print("tuned", kc, by, width, ms, flush=True)
        b.step(1, nt, use_graph=True, block=5)
        b.synchronize()
        for s in alone:
            s.step(1, nt, use_graph=True, block=5)
        for n, (x, y) in enumerate(zip(batched, alone)):
            y.synchronize()
            for name in O.FIELDS_V + O.FIELDS_S:
                fx, fy = x.get_field(name), y.get_field(name)
                assert ulp_diff(fx, fy) == 0, (n, name)
                assert np.array_equal(np.signbit(fx), np.signbit(fy)), (n, name)
            assert np.any(x.get_field("Vz") != 0), n
            tx, ty = x.seismograms("vel"), y.seismograms("vel")
            assert tx.shape == ty.shape and np.array_equal(tx, ty), n
    finally:
        b.close()
        for s in batched + alone:
            s.close()


@pytest.mark.parametrize("packing", [0, 1, 2])
def test_batch_sections_and_packing(orc, packing):
    """Sections of every member taken by the batch's one section launch per
    step (different sections, sizes and decimations per member), under both
    member packings: each member's frames and maximum maps equal its own
    oracle run's, bit for bit."""
    nt, ntdec_s = 14, 3
    nframe = (nt - 1) // ntdec_s + 1
    cfgs = _members()[:3]
    secs = [["xy_v", "fs_v", "yz_ps"], ["xz_u", "ob_v"], ["fs_u", "xy_ps", "xz_v", "ob_u"]]
    pars = [dict(k0_xy=7, i0_yz=12, j0_xz=11, idec=1, jdec=1, kdec=1),
            dict(k0_xy=5, i0_yz=9, j0_xz=10, idec=2, jdec=1, kdec=2),
            dict(k0_xy=6, i0_yz=8, j0_xz=9, idec=1, jdec=3, kdec=1)]
    models = [O.build_model(c) for c in cfgs]
    runs, gs = [], []
    for m, sec, par in zip(models, secs, pars):
        par = dict(par, ntdec_s=ntdec_s, nframe=nframe)
        r = O.OracleRun(m, 32)
        r.snap_setup(sec, **par)
        r.run(nt)
        runs.append(r)
        g = gpu_from_model(m)
        g.set_snapshots(sec, m.kfs, M0=m.M0, UC=m.UC, **par)
        gs.append(g)
    b = SwpcBatch(gs)
    b.set_packing(packing)
    assert b.packing() == packing
    b.step(1, nt, use_graph=True, block=7)
    b.synchronize()
    _check_members(gs, runs)
    nonzero = 0
    for g, r, sec in zip(gs, runs, secs):
        for name in sec:
            for f in range(nframe):
                want = r.snap_frames[name][f]
                assert ulp_diff(g.snapshot(name, f), want) == 0, (name, f)
                nonzero += int(np.any(want != 0))
            pl, q = name.split("_")
            if pl in ("fs", "ob") and q in ("v", "u"):
                want = (r.snap_vmax if q == "v" else r.snap_umax)[pl]
                assert ulp_diff(g.snap_max(name), want) == 0, name
    assert nonzero > 0
    b.close()
