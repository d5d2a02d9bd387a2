"""Full production size (1024 x 1024 x 512, dx = 0.25 km, NM = 3, 20-cell
PML) against the oracle, through translation invariance.

In a laterally uniform medium every interior cell runs the same arithmetic
(m_kernel.f90:75-348), and a cell's update reads only cells within two of
it along each axis, so a point source's field stays inside a box of half
width 4 per step around it.  While that box stays clear of the PML, the
free-surface band and the other sources, the field around a source does
not depend on where the source sits: the field around each of three
sources on the full grid must equal, bit for bit, the oracle's field around
one source in the middle of a 160^3 grid of the same medium.  The three
positions put the source at different offsets in the kernels' tiles and
k chunks, near the low corner, mid-grid and near the high corner (byte
offsets past 2^31 in every array), and every cell outside the three boxes
must still be exactly zero.
"""
import gc

import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_helpers import ulp_diff

pytestmark = pytest.mark.gpu

NX, NY, NZ, NA = 1024, 1024, 512, 20
DX = 0.25
DT = 0.0117
NSTEP = 12
W = 4 * NSTEP + 8                     # half width of the compared boxes
SMALL = 160
# source cells (i, j, k) on the full grid: clear of the PML (cells 21..1004
# laterally, above 493 in depth) and the free-surface band (k <= 8) by W
PLACES = [(77, 78, 66), (517, 262, 141), (946, 941, 433)]


def _small_model():
    c = -SMALL / 2 * DX + (SMALL / 2 - 0.5) * DX       # centre of cell 80
    cfg = O.Config(nx=SMALL, ny=SMALL, nz=SMALL, dx=DX, dy=DX, dz=DX, dt=DT, nt=NSTEP + 1, na=NA, nm=3,
                   zbeg=-1.0, layers=[[0.0, 2.7, 5.5, 3.2, 320.0, 160.0]],
                   sources=[dict(x=c, y=c, z=-1.0 + (SMALL / 2 - 0.5) * DX, t0=0.0, tr=1.0, mo=1e15, strike=30.0,
                                 dip=60.0, rake=90.0)])
    return O.build_model(cfg)


def _axis_map(nbig, nsmall, mid):
    """full-grid halo index -> small-grid halo index: both edges (halo, PML
    and a margin) map to the same distance from the small grid's edge, the
    rest to the small grid's middle"""
    b = np.arange(nbig + 2 * O.HALO)
    edge = O.OFF + NA + 8
    s = np.full(b.shape, mid)
    s[b < edge] = b[b < edge]
    hi = b >= nbig + 2 * O.HALO - edge
    s[hi] = b[hi] - (nbig - nsmall)
    return s


def test_full_grid_sources_match_the_oracle_at_three_offsets(orc):
    from openswpc_amd.gpu import SwpcGPU

    m = _small_model()
    assert (m.isrc[0], m.jsrc[0], m.ksrc[0]) == (80, 80, 80)
    small = O.OracleRun(m, 32)
    small.run(NSTEP)

    # full-grid medium: the small model is uniform laterally (halo included)
    # and below its band, so one column, stretched in depth, fills the grid
    for name in ("rho", "lam", "mu", "taup", "taus"):
        a = getattr(m, name)
        assert np.array_equal(a, np.broadcast_to(a[:1, :1, :], a.shape)), name
        assert np.all(a[0, 0, 16:] == a[0, 0, 16]), name
    mid = O.OFF + SMALL // 2                 # numpy index of cell 80
    kmap = np.minimum(np.arange(NZ + 2 * O.HALO), mid)
    jmap, imap = _axis_map(NY, SMALL, mid), _axis_map(NX, SMALL, mid)
    g = SwpcGPU(NX, NY, NZ, NA, 3, "pml", DX, DX, DX, DT)
    try:
        big = {}
        for name in ("rho", "lam", "mu", "taup", "taus"):
            col = getattr(m, name)[0, 0, kmap]
            arr = np.empty((NY + 2 * O.HALO, NX + 2 * O.HALO, NZ + 2 * O.HALO), np.float32)
            arr[...] = col[None, None, :]
            big[name] = arr
        maps = {n: np.ascontiguousarray(getattr(m, n)[np.ix_(jmap, imap)])
                for n in ("kfs_top", "kfs_bot", "kob_top", "kob_bot", "kob")}
        g.set_medium(big["rho"], big["lam"], big["mu"], big["taup"], big["taus"], maps["kfs_top"],
                     maps["kfs_bot"], maps["kob_top"], maps["kob_bot"], maps["kob"])
        del big
        gc.collect()
        g.set_visco(m.c1, m.c2, m.d1, m.d2)
        f32 = np.float32
        prof = {}
        for ax, n, beg in (("x", NX, -NX / 2 * DX), ("y", NY, -NY / 2 * DX), ("z", NZ, -1.0)):
            gc_, ge_ = np.zeros(4 * n, f32), np.zeros(4 * n, f32)
            O.lib(32).orc_pml_profile_axis(n, NA, beg, f32(DX), f32(m.fcut), DT, gc_, ge_)
            prof[ax] = (gc_, ge_)
        g.set_pml(*prof["x"], *prof["y"], *prof["z"])
        ns = len(PLACES)
        g.set_sources(np.array([p[0] for p in PLACES], np.int32), np.array([p[1] for p in PLACES], np.int32),
                      np.array([p[2] for p in PLACES], np.int32), np.repeat(m.mo, ns),
                      np.ascontiguousarray(np.repeat(m.mij, ns, axis=1)), m.M0,
                      np.ascontiguousarray(np.repeat(m.stf, ns, axis=1)))
        g.step(1, NSTEP, use_graph=True, block=4)
        g.synchronize()

        H = O.OFF
        sw = slice(80 - W + H, 80 + W + 1 + H)
        for name in O.FIELDS_V + O.FIELDS_S:
            ref = small.field(name)[sw, sw, sw]
            assert np.count_nonzero(ref) > 20_000, name
            assert np.count_nonzero(small.field(name)) == np.count_nonzero(ref), name   # box holds the cone
            f = g.get_field(name)
            inside = 0
            for (i, j, k) in PLACES:
                got = f[j - W + H:j + W + 1 + H, i - W + H:i + W + 1 + H, k - W + H:k + W + 1 + H]
                d = ulp_diff(got, ref)
                print(name, (i, j, k), "ulp", d, flush=True)
                assert d == 0, (name, (i, j, k), d)
                assert np.array_equal(np.signbit(got), np.signbit(ref)), (name, (i, j, k))
                inside += np.count_nonzero(got)
            assert np.count_nonzero(f) == inside, (name, "values outside the source boxes")
            del f
            gc.collect()
    finally:
        g.close()


@pytest.mark.timeout(900)
def test_production_medium_on_the_device_equals_the_host_library():
    """The benchmark's own set-up at full size: the production workload's
    medium built on the device (the von Karman field generated into the
    context's staging buffer, swpc_medium_layered) equals the host library's
    set-up of the same workload (the same field copied to the host,
    swpch_perturb and swpch_medium_setup), bit for bit in every medium array,
    band map and kernel coefficient."""
    from openswpc_amd.host import build_inputs, make_simulation_device
    from openswpc_amd.workloads import production_spec
    spec = production_spec()
    gd, inp_d = make_simulation_device(spec)
    inp = build_inputs(spec)
    a = inp.arrays
    for n in ("kfs", "kob", "ist", "jst", "kst", "c1", "c2", "d1", "ts"):
        assert np.array_equal(inp_d.arrays[n], a[n]), n
    for n in ("vmin", "vmax", "zeta", "d2"):
        assert np.float32(inp_d.scalars[n]) == np.float32(inp.scalars[n]), n
    for n in ("kfs_top", "kfs_bot", "kob_top", "kob_bot", "kob"):
        assert np.array_equal(gd.get_medium(n), a[n]), n
    for n in ("rho", "lam", "mu", "taup", "taus"):
        got = gd.get_medium(n)
        want = a.pop(n)
        nbad = int(np.count_nonzero(got.view(np.uint32) != want.view(np.uint32)))
        assert nbad == 0, (n, nbad)
        del got, want
        gc.collect()
    gd.close()
