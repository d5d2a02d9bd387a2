"""GPU parity: the HIP library against the FP32 oracle on identical inputs.

The library evaluates every expression in the reference's FP32 order with
IEEE division and no FMA contraction, so the target is bit-for-bit equality
(0 ulp) with the oracle's FP32 variant.  The accuracy bar of
relative L2 <= 1e-5 is asserted as well, so a failure of the stronger
property still reports how far apart the results are.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_helpers import compare_fields, gpu_from_model, small_config, ulp_diff

pytestmark = pytest.mark.gpu
REL_TOL = 1e-5


def _check(g, run, label):
    res = compare_fields(g, run)
    print(label, res, flush=True)
    for n, (ulp, rel) in res.items():
        assert rel <= REL_TOL, (label, n, rel)
        assert ulp == 0, (label, n, ulp, rel)


def _run_both(cfg, nsteps, graph=False):
    m = O.build_model(cfg)
    run = O.OracleRun(m, 32)
    run.run(nsteps)
    g = gpu_from_model(m)
    if graph:
        g.step(1, nsteps, use_graph=True, block=7)
    else:
        g.step(1, nsteps)
    g.synchronize()
    return m, run, g


@pytest.mark.parametrize("nm", [3, 2, 0, 1, 5, 8])
def test_pml_moment_tensor_steps(orc, nm):
    cfg = small_config(nm=nm)
    m, run, g = _run_both(cfg, 40)
    _check(g, run, f"pml nm={nm}")
    for kind in ("vel", "disp", "stress", "strain"):
        got = g.seismograms(kind)
        want = run.wav[kind]
        assert ulp_diff(got, want) == 0, (kind, O.rel_l2(got, want))


def test_cerjan_moment_tensor_steps(orc):
    cfg = small_config(abc="cerjan", nm=3)
    m, run, g = _run_both(cfg, 40)
    _check(g, run, "cerjan")


def test_body_force_steps(orc):
    cfg = small_config(bf_mode=True, sources=[dict(x=0.3, y=-0.6, z=5.2, t0=0.0, tr=0.6, fx=0.3, fy=-1.0, fz=0.7)])
    m, run, g = _run_both(cfg, 40)
    _check(g, run, "body force")
    assert ulp_diff(g.seismograms("vel"), run.wav["vel"]) == 0


def test_graph_replay_matches_direct(orc):
    cfg = small_config()
    m, run, g = _run_both(cfg, 30, graph=True)
    _check(g, run, "graph block 7")
    assert ulp_diff(g.seismograms("vel"), run.wav["vel"]) == 0


@pytest.mark.parametrize("abc", ["pml", "cerjan"])
def test_step_timed_matches_oracle(orc, abc):
    """swpc_step_timed runs the same launches as swpc_step, with events between them."""
    cfg = small_config(abc=abc)
    m = O.build_model(cfg)
    run = O.OracleRun(m, 32)
    run.run(9)
    g = gpu_from_model(m)
    ms = g.step_timed(1, 9)
    _check(g, run, f"timed steps {abc}")
    assert ulp_diff(g.seismograms("vel"), run.wav["vel"]) == 0
    assert all(v >= 0.0 for v in ms.values())
    assert ms["stress"] > 0.0 and ms["vel"] > 0.0
    # an empty slot still shows the few microseconds between two events
    assert ms["stress_side" if abc == "pml" else "stress_abc"] > 0.0


def test_random_fields_one_step_each_routine(orc):
    """Random V/S everywhere (band, interior, PML) exercise every stencil term."""
    cfg = small_config(nm=3, sources=[], stations=[])
    m = O.build_model(cfg)
    run = O.OracleRun(m, 32)
    g = gpu_from_model(m)
    rng = np.random.default_rng(11)
    for n in O.FIELDS_V + O.FIELDS_S:
        a = np.zeros(O.n3(cfg), np.float32)
        a[3:-3, 3:-3, 3:-3] = rng.standard_normal(a[3:-3, 3:-3, 3:-3].shape).astype(np.float32)
        run.arr[n][...] = a
        g.set_field(n, a)
    L = run.L
    L.orc_kernel_update_stress(run.ptr)
    g.kernel_update_stress()
    _check(g, run, "kernel stress")
    L.orc_absorb_update_stress(run.ptr)
    g.absorb_update_stress()
    _check(g, run, "absorb stress")
    L.orc_kernel_update_vel(run.ptr)
    g.kernel_update_vel()
    _check(g, run, "kernel vel")
    L.orc_absorb_update_vel(run.ptr)
    g.absorb_update_vel()
    _check(g, run, "absorb vel")
    for _ in range(3):
        run.step()
    g.step(1, 3)
    _check(g, run, "3 more steps")


def test_colocated_sources_keep_order(orc):
    srcs = [dict(x=0.3, y=-0.6, z=5.2, t0=0.0, tr=0.7, mo=1e15 * (n + 1), strike=10.0 * n, dip=45.0, rake=30.0 * n)
            for n in range(4)]
    srcs.append(dict(x=0.8, y=-0.6, z=5.2, t0=0.1, tr=0.9, mo=3e14, strike=200.0, dip=80.0, rake=-90.0))
    cfg = small_config(sources=srcs)
    m, run, g = _run_both(cfg, 25)
    _check(g, run, "colocated sources")


def test_vmax_matches(orc):
    cfg = small_config(nx=48, ny=48, na=8)
    m, run, g = _run_both(cfg, 40)
    v_g = g.kernel_vmax()
    v_o = run.vmax()
    assert np.array_equal(v_g, v_o), (v_g, v_o)


@pytest.mark.parametrize("kchunk,by,variant,graph", [(7, 1, 0, False), (16, 2, 1, True), (200, 4, 3, True),
                                                     (64, 4, 3, False), (32, 2, 2, True), (64, 4, 2, False),
                                                     (7, 1, 4, False), (16, 2, 5, True), (64, 4, 7, False),
                                                     (7, 1, 8, False), (16, 2, 9, True), (1, 4, 9, False),
                                                     (64, 4, 13, True), (16, 4, 41, True), (7, 2, 33, False),
                                                     (32, 4, 45, True), (5, 1, 43, True)])
def test_launch_shapes_same_result(orc, kchunk, by, variant, graph):
    cfg = small_config(nx=44, ny=40, nz=36)
    m = O.build_model(cfg)
    run = O.OracleRun(m, 32)
    run.run(20)
    g = gpu_from_model(m, kchunk=kchunk, by=by, variant=variant)
    g.step(1, 20, use_graph=graph, block=6)
    _check(g, run, f"kchunk={kchunk} by={by} variant={variant} graph={graph}")


@pytest.mark.parametrize("na,by,variant", [(7, 4, 1), (5, 2, 1), (13, 1, 1), (7, 2, 5), (13, 4, 5), (7, 2, 9),
                                           (13, 4, 13), (7, 4, 41), (5, 2, 45), (13, 1, 41)])
def test_odd_sizes_and_absorber_widths(orc, na, by, variant):
    """PML tiles where na is not a multiple of the tile height and the grid
    is not a multiple of the tile width: every PML cell updated exactly once."""
    cfg = small_config(nx=37, ny=41, nz=29, na=na, nt=25,
                       sources=[dict(x=0.3, y=-0.6, z=4.2, t0=0.0, tr=0.6, mo=1e15, strike=30.0, dip=60.0, rake=90.0)],
                       stations=[(1.1, 1.3, 0.0, 'fsb')])
    m = O.build_model(cfg)
    run = O.OracleRun(m, 32)
    run.run(25)
    g = gpu_from_model(m, kchunk=8, by=by, variant=variant)
    g.step(1, 25)
    _check(g, run, f"na={na} by={by} variant={variant}")
