"""Per-cell arrays striped over the HBM regions (swpc_stripe_reserve).

The arena changes where the arrays live, not what the kernels compute: a
context whose arrays come from the arena must step to the same bits as the
FP32 oracle (small grid: every array sub-allocated from the regions' slabs)
and as a context with hipMalloc placement (larger grids: fields and medium
arrays in whole 256 MiB granules, the memory-variable packs and the PML
auxiliaries split over the three regions; and an arena too small for the
grid, so that some arrays fall back to hipMalloc).  The bookkeeping: arena
memory is taken while a context holds it and comes back when the context is
destroyed, and the arena refuses to be released while in use."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_helpers import compare_fields, gpu_from_model, small_config

pytestmark = pytest.mark.gpu
GRAN = 256 << 20


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _free():
    from openswpc_amd.gpu import stripe_free_bytes
    return [stripe_free_bytes(0, r) for r in range(3)]


def test_small_grid_in_the_arena_equals_the_oracle(orc):
    from openswpc_amd.gpu import SwpcError, stripe_release, stripe_reserve
    cfg = small_config(nm=3)            # 18 memory variables: four float4 packs and a float2 pack
    m = O.build_model(cfg)
    run = O.OracleRun(m, 32)
    run.run(30)
    stripe_reserve(GRAN)
    try:
        free0 = _free()
        assert all(f == GRAN for f in free0), free0
        g = gpu_from_model(m)
        try:
            taken = [a - b for a, b in zip(free0, _free())]
            print("arena bytes taken per region", taken, flush=True)
            assert all(t > 0 for t in taken), taken
            with pytest.raises(SwpcError):
                stripe_release()
            g.step(1, 30, use_graph=True, block=10)
            g.synchronize()
            res = compare_fields(g, run)
            print(res, flush=True)
            for n, (ulp, rel) in res.items():
                assert ulp == 0, (n, ulp, rel)
        finally:
            g.close()
        assert _free() == free0
    finally:
        stripe_release()


def _spec(nx, ny, nz):
    from openswpc_amd.host import ModelSpec
    src = [dict(x=0.4, y=-0.7, z=6.3, t0=0.0, tr=0.8, mo=1e15, strike=30.0, dip=60.0, rake=90.0)]
    st = [(2.1, 1.3, 0.0, "fsb"), (-3.2, 2.7, 0.0, "fsb"), (1.5, 0.5, 3.0, "dep")]
    return ModelSpec(nx=nx, ny=ny, nz=nz, dx=0.25, dy=0.25, dz=0.25, dt=0.01, nt=40, na=20, nm=3, zbeg=-1.0,
                     vmodel="lhm", sources=src, stations=st, wav_what=3, fq_min=0.05, fq_max=5.0, fq_ref=1.0,
                     layers=[[0.0, 2.4, 4.5, 2.6, 200.0, 100.0], [4.0, 2.6, 5.6, 3.2, 300.0, 150.0]])


# (grid, arena bytes per region): 512x512x256 fits (fields of 312 MB take
# whole granules, the 4.1 GB of memory variables five packs over the three
# regions); 256x256x320 in one granule per region does not (fields of 109 MB
# share slabs until they are full, the rest falls back to hipMalloc)
@pytest.mark.parametrize("shape,per_region", [((512, 512, 256), None), ((256, 256, 320), GRAN)])
def test_striped_context_equals_hipmalloc_placement(shape, per_region, monkeypatch):
    from openswpc_amd.gpu import stripe_bytes_per_region, stripe_release, stripe_reserve
    from openswpc_amd.host import build_inputs, make_simulation
    spec = _spec(*shape)
    inp = build_inputs(spec)
    monkeypatch.setenv("SWPC_STRIPE", "0")      # no private arena: hipMalloc placement
    plain = make_simulation(inp)
    monkeypatch.delenv("SWPC_STRIPE")
    sims = [plain]
    per = per_region or stripe_bytes_per_region(spec.nx, spec.ny, spec.nz, spec.nm, spec.na) + GRAN
    stripe_reserve(per)
    try:
        free0 = _free()
        striped = make_simulation(inp)
        sims.append(striped)
        taken = sum(a - b for a, b in zip(free0, _free()))
        print(shape, "arena bytes taken", taken, "of", striped.device_bytes(), flush=True)
        if per_region is None:
            assert taken >= 0.9 * striped.device_bytes()
        else:
            assert 0 < taken <= 3 * GRAN
        for s in sims:
            s.step(1, 40, use_graph=True, block=20)
        for n in ("Vx", "Vy", "Vz", "Sxx", "Syy", "Szz", "Syz", "Sxz", "Sxy"):
            a, b = striped.get_field(n), plain.get_field(n)
            assert np.array_equal(_bits(a), _bits(b)), (shape, n)
        assert np.any(plain.get_field("Vz"))
        assert np.array_equal(_bits(striped.seismograms("vel")), _bits(plain.seismograms("vel")))
    finally:
        for s in sims:
            s.close()
        stripe_release()


def test_adjoint_gradient_in_the_arena_equals_hipmalloc_placement():
    """The adjoint workspace and its checkpoint stores come from the arena
    too (in segments rotating over the regions): the misfit and the three
    gradients are the same bits as with hipMalloc placement."""
    import dataclasses
    from openswpc_amd.gpu import stripe_release, stripe_reserve
    from openswpc_amd.host import build_inputs, make_simulation
    spec = dataclasses.replace(_spec(96, 88, 72), nt=60)
    true = make_simulation(build_inputs(dataclasses.replace(spec, layers=[[0.0, 2.4, 4.7, 2.7, 200.0, 100.0],
                                                                           [4.0, 2.6, 5.6, 3.2, 300.0, 150.0]])))
    true.step(1, 60, use_graph=True, block=20)
    obs = true.seismograms("vel")
    true.close()
    inp = build_inputs(spec)
    plain = make_simulation(inp)
    J0, *g0 = plain.adjoint_gradient(60, obs, ckp=8)
    plain.close()
    stripe_reserve(2 << 30)
    try:
        striped = make_simulation(inp)
        try:
            J1, *g1 = striped.adjoint_gradient(60, obs, ckp=8)
        finally:
            striped.close()
    finally:
        stripe_release()
    print("misfit", J0, J1, flush=True)
    assert J0 == J1 and J0 > 0
    for a, b in zip(g0, g1):
        assert np.array_equal(_bits(a), _bits(b))
        assert np.any(a)


def test_batch_of_arena_members_matches_their_oracles(orc):
    """Batch members whose arrays come from the arena's slabs (PML and Cerjan
    members mixed, packing 2, graph replays) end bit-identical to their own
    oracle runs."""
    from openswpc_amd.gpu import SwpcBatch, stripe_release, stripe_reserve
    from tests.test_gpu_batch import _check_members, _members
    nt = 18
    models = [O.build_model(c) for c in _members()]
    runs = []
    for m in models:
        r = O.OracleRun(m, 32)
        r.run(nt)
        runs.append(r)
    stripe_reserve(GRAN)
    gs = []
    try:
        free0 = _free()
        gs = [gpu_from_model(m) for m in models]
        assert all(a > b for a, b in zip(free0, _free()))
        b = SwpcBatch(gs)
        b.set_packing(2)
        b.step(1, 12, use_graph=True, block=4)
        b.step(13, nt - 12, use_graph=True, block=3)
        b.synchronize()
        _check_members(gs, runs)
        b.close()
    finally:
        for g in gs:
            g.close()
        stripe_release()


def test_large_context_stripes_itself(monkeypatch):
    """A context of 4 GiB or more created on a device without an arena
    reserves a private one for itself: another context does not draw from it
    and cannot reserve over it, it goes with its owner, and the placement
    changes no result.  SWPC_STRIPE=0 turns it off."""
    from openswpc_amd.gpu import SwpcError, stripe_reserve
    from openswpc_amd.host import build_inputs, make_simulation
    inp = build_inputs(_spec(512, 512, 256))
    assert _free() == [0, 0, 0]
    own = make_simulation(inp)
    sims = [own]
    try:
        slack = _free()
        print("private arena, free bytes per region", slack, flush=True)
        assert all(f > 0 for f in slack)
        monkeypatch.setenv("SWPC_STRIPE", "0")
        off = make_simulation(inp)              # hipMalloc placement
        sims.append(off)
        monkeypatch.delenv("SWPC_STRIPE")
        other = make_simulation(inp)            # the private arena is not its to use
        sims.append(other)
        assert _free() == slack
        with pytest.raises(SwpcError):
            stripe_reserve(GRAN)
        for s in sims:
            s.step(1, 30, use_graph=True, block=15)
        for n in ("Vx", "Vz", "Sxx", "Syz"):
            a = own.get_field(n)
            assert np.any(a) or n != "Vz"
            for s in sims[1:]:
                assert np.array_equal(_bits(a), _bits(s.get_field(n))), n
        own.close()
        sims.remove(own)
        assert _free() == [0, 0, 0]
    finally:
        for s in sims:
            s.close()
