"""The fused velocity-stress launch (variant bit 7) against the FP32 oracle.

With bit 7 set, one interior launch runs the velocity update of a step and
the stress update of the next.  Fields alternate between two buffer sets,
and each tile recomputes the velocities of the cells around it.  The step
is the same arithmetic in the same order, so fields, seismograms and
sections must stay bit-identical to the oracle.  The cases cover: direct
calls and graph blocks (even and odd block lengths); 4- and 8-row tiles
(bit 8); k chunks from one plane to the whole interior; odd grids and
absorber widths; an ocean column; sections with decimation.  A body-force
source keeps the per-phase launches, and that run must match too."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_helpers import compare_fields, gpu_from_model, small_config, ulp_diff

pytestmark = pytest.mark.gpu

FUSED = 128 | 9          # fused steps on top of the default variant
FUSED8 = FUSED | 256     # 8-row fused tiles


def _check(g, run, label, kinds=("vel", "disp", "stress", "strain")):
    res = compare_fields(g, run)
    print(label, res, flush=True)
    for n, (ulp, rel) in res.items():
        assert ulp == 0, (label, n, ulp, rel)
    for kind in kinds:
        got, want = g.seismograms(kind), run.wav[kind]
        assert ulp_diff(got, want) == 0, (label, kind, O.rel_l2(got, want))


@pytest.mark.parametrize("nm,variant,kchunk,graph,block,nsteps", [
    (3, FUSED, 8, False, 0, 40),
    (3, FUSED, 16, True, 7, 40),
    (0, FUSED8, 5, True, 4, 30),
    (1, FUSED8, 64, False, 0, 21),
    (3, FUSED, 1, True, 6, 20),
    (2, FUSED8, 3, True, 3, 17),
    (3, FUSED, 8, False, 0, 3),
    (5, FUSED, 8, False, 0, 2),
])
def test_fused_steps_match_oracle(orc, nm, variant, kchunk, graph, block, nsteps):
    cfg = small_config(nm=nm)
    m = O.build_model(cfg)
    run = O.OracleRun(m, 32)
    run.run(nsteps)
    g = gpu_from_model(m, kchunk=kchunk, by=4, variant=variant)
    g.step(1, nsteps, use_graph=graph, block=block or 100)
    g.synchronize()
    _check(g, run, f"nm={nm} variant={variant} kchunk={kchunk} graph={graph} block={block} nsteps={nsteps}")


@pytest.mark.parametrize("nx,ny,nz,na,variant,kchunk", [
    (37, 41, 29, 7, FUSED, 8), (37, 41, 29, 5, FUSED8, 6), (37, 41, 29, 13, FUSED, 4),
    (130, 70, 40, 10, FUSED, 16), (130, 70, 40, 10, FUSED8, 16), (67, 45, 30, 3, FUSED8, 7),
])
def test_fused_odd_grids_and_absorbers(orc, nx, ny, nz, na, variant, kchunk):
    """tiles that overhang the interior box in x and y, chunks of uneven
    depth, interior boxes narrower than one tile"""
    cfg = small_config(nx=nx, ny=ny, nz=nz, na=na, nt=25,
                       sources=[dict(x=0.3, y=-0.6, z=4.2, t0=0.0, tr=0.6, mo=1e15, strike=30.0, dip=60.0, rake=90.0)],
                       stations=[(1.1, 1.3, 0.0, 'fsb')])
    m = O.build_model(cfg)
    run = O.OracleRun(m, 32)
    run.run(25)
    g = gpu_from_model(m, kchunk=kchunk, by=4, variant=variant)
    g.step(1, 25, use_graph=True, block=8)
    g.synchronize()
    _check(g, run, f"{nx}x{ny}x{nz} na={na} variant={variant} kchunk={kchunk}")


def test_fused_ocean_column(orc):
    from tests.test_gpu_ocean import ocean_config
    cfg = ocean_config(nm=3, abc="pml")
    m = O.build_model(cfg)
    run = O.OracleRun(m, 32)
    run.run(30)
    g = gpu_from_model(m, kchunk=8, by=4, variant=FUSED8)
    g.step(1, 30, use_graph=True, block=10)
    g.synchronize()
    res = compare_fields(g, run)
    for n, (ulp, rel) in res.items():
        assert ulp == 0, ("ocean", n, ulp, rel)


@pytest.mark.parametrize("variant,graph", [(FUSED, False), (FUSED8, True)])
def test_fused_sections(orc, variant, graph):
    """sections read the velocities and stresses of the step before, which
    the two field sets hold while the fused launches run"""
    all_sections = [f"{p}_{q}" for p in ("xy", "yz", "xz", "fs", "ob") for q in ("ps", "v", "u")]
    nt, ntdec_s = 20, 3
    nframe = (nt - 1) // ntdec_s + 1
    m = O.build_model(small_config(nt=nt))
    par = dict(k0_xy=9, i0_yz=17, j0_xz=15, idec=2, jdec=3, kdec=2, ntdec_s=ntdec_s, nframe=nframe)
    run = O.OracleRun(m, 32)
    run.snap_setup(all_sections, **par)
    run.run(nt)
    g = gpu_from_model(m, kchunk=8, by=4, variant=variant)
    g.set_snapshots(all_sections, m.kfs, M0=m.M0, UC=m.UC, **par)
    g.step(1, nt, use_graph=graph, block=5)
    g.synchronize()
    for name in all_sections:
        for f in range(nframe):
            assert ulp_diff(g.snapshot(name, f), run.snap_frames[name][f]) == 0, (name, f)
    for pl in ("fs", "ob"):
        assert ulp_diff(g.snap_max(f"{pl}_v"), run.snap_vmax[pl]) == 0, pl
        assert ulp_diff(g.snap_max(f"{pl}_u"), run.snap_umax[pl]) == 0, pl


def test_fused_body_force_keeps_the_phase_launches(orc):
    """a body force acts between the two updates: the library keeps the
    per-phase launches for it, with the same results"""
    cfg = small_config(bf_mode=True, nm=2, sources=[dict(x=0.3, y=-0.6, z=5.2, t0=0.0, tr=0.6, fx=0.3, fy=-1.0, fz=0.7)])
    m = O.build_model(cfg)
    run = O.OracleRun(m, 32)
    run.run(20)
    g = gpu_from_model(m, kchunk=8, by=4, variant=FUSED)
    g.step(1, 20, use_graph=True, block=5)
    g.synchronize()
    _check(g, run, "body force, variant bit 7")


def test_fused_equals_phase_launches_on_a_larger_grid():
    """the fused and the per-phase step give the same bits where every tile
    has interior neighbours on all sides (192 x 160 x 96, 60 steps)"""
    cfg = small_config(nx=192, ny=160, nz=96, na=12, nm=3, nt=60, dx=0.25, dy=0.25, dz=0.25, dt=0.01,
                       sources=[dict(x=0.3, y=-0.6, z=8.2, t0=0.05, tr=0.4, mo=1e15, strike=30.0, dip=60.0,
                                     rake=90.0)])
    m = O.build_model(cfg)
    out = []
    for variant in (9, FUSED, FUSED8):
        g = gpu_from_model(m, kchunk=16, by=4, variant=variant)
        g.step(1, 60, use_graph=True, block=15)
        g.synchronize()
        out.append({n: g.get_field(n) for n in O.FIELDS_V + O.FIELDS_S})
        out[-1]["vel"] = g.seismograms("vel")
        g.close()
    assert np.any(out[0]["Vz"])
    for o in out[1:]:
        for n, a in out[0].items():
            assert ulp_diff(a, o[n]) == 0, n


def test_fused_cerjan_keeps_the_phase_launches(orc):
    """the fused launch needs the PML: a Cerjan run with bit 7 holds no second
    field set and steps as before"""
    m = O.build_model(small_config(abc="cerjan", nm=2))
    run = O.OracleRun(m, 32)
    run.run(15)
    g = gpu_from_model(m, kchunk=8, by=4, variant=9)
    base = g.device_bytes()
    g.set_tuning(8, 4, FUSED)
    assert g.device_bytes() == base
    g.step(1, 15, use_graph=True, block=5)
    g.synchronize()
    _check(g, run, "cerjan, variant bit 7", kinds=("vel",))


def test_second_field_set_comes_and_goes(orc):
    """bit 7 allocates nine more field arrays, clearing it frees them, and
    the state survives both changes"""
    m = O.build_model(small_config(nm=3))
    run = O.OracleRun(m, 32)
    run.run(24)
    g = gpu_from_model(m, kchunk=8, by=4, variant=9)
    base = g.device_bytes()
    g.step(1, 8)
    g.set_tuning(0, 0, FUSED)
    grown = g.device_bytes() - base
    cfg = m.cfg
    pitch = -(-(32 + cfg.nx + 3) // 32) * 32
    assert grown == 9 * pitch * (cfg.ny + 6) * (cfg.nz + 6) * 4, grown
    g.step(9, 8, use_graph=True, block=4)
    g.set_tuning(0, 0, 9)
    assert g.device_bytes() == base
    g.step(17, 8)
    g.synchronize()
    _check(g, run, "bit 7 set and cleared mid-run")
