"""The plain 128^3 workload end to end through the product path, with the
FP32 error against an FP64 run of the same discrete scheme stated.

Product path: the Fortran host set-up (openswpc_amd/host/swpc_host.f90,
through host.build_inputs) -> the C ABI (host.make_simulation) -> graph
replays of 100-step blocks (swpc_step), sections taken on the device.
Workload: 128 x 128 x 128 cells, dx = 0.5 km, dt = 0.02 s, 1000 steps,
20-cell PML, two layers (5 km top of the second), NM = 3, one seeded moment
tensor at 10 km (Kupper, 1 s), 16 seeded surface receivers (velocity and
displacement every step), xy/xz/yz/fs velocity sections every 100 steps and
the 3D velocity field every 100 steps (src/swpc_3d/main.f90:119-139,
m_wav.f90:397-625, m_snap.f90:919-2189).

Against the FP32 oracle (the reference's FP32 build restated) on the same
inputs every trace, section frame and 3D field is bit-identical, and so
within the relative L2 of 1e-5.  Against the FP64 oracle the relative L2 of
each trace component, each section frame and each 3D field is printed: the
error of the FP32 scheme itself."""
import numpy as np
import pytest

from oracle import oracle as O
from gpu_helpers import oracle_config, ulp_diff

pytestmark = pytest.mark.gpu

REL_TOL = 1e-5
SECTIONS = ["xy_v", "xz_v", "yz_v", "fs_v"]


def _rel(a, b):
    return O.rel_l2(a, b)


@pytest.mark.timeout(1200)
def test_plain_128_product_path_1000_steps(orc):
    from openswpc_amd import workloads
    from openswpc_amd.host import build_inputs, make_simulation

    spec = workloads.plain_spec()
    spec.wav_what = 3                                    # velocity and displacement
    nt, ntdec_s = spec.nt, 100
    nframe = (nt - 1) // ntdec_s + 1
    cfg = oracle_config(spec)
    m = O.build_model(cfg)
    inp = build_inputs(spec)
    a = inp.arrays
    # the product set-up gives the oracle's kernel inputs for this medium, bit for bit
    for n in ("rho", "lam", "mu", "taup", "taus", "kfs_top", "kfs_bot", "kob_top", "kob_bot", "ist", "jst", "kst",
              "isrc", "jsrc", "ksrc", "stf", "c1", "c2", "d1", "gxc", "gzc", "gze"):
        want = getattr(m, n if n not in ("stf",) else n)
        assert np.array_equal(a[n].reshape(-1), np.asarray(want).reshape(-1)[: a[n].size]), n
    assert len(a["ist"]) == 16
    par = dict(k0_xy=24, i0_yz=64, j0_xz=64, idec=1, jdec=1, kdec=1, ntdec_s=ntdec_s, nframe=nframe)
    g = make_simulation(inp)
    g.set_snapshots(SECTIONS, a["kfs"], M0=inp.scalars["M0"], UC=inp.scalars["UC"], **par)
    runs = {}
    for prec in (32, 64):
        r = O.OracleRun(m, prec)
        r.snap_setup(SECTIONS, **par)
        runs[prec] = r
    field_err = []
    for b in range(nframe):
        it0 = 1 + b * ntdec_s
        g.step(it0, ntdec_s, use_graph=True, block=ntdec_s)
        for r in runs.values():
            r.run(ntdec_s)
        g.synchronize()
        row = []
        for n in O.FIELDS_V:
            got = g.get_field(n)
            assert ulp_diff(got, runs[32].field(n)) == 0, (it0 + ntdec_s - 1, n)
            row.append(_rel(got, runs[64].field(n)))
        field_err.append(row)
        print(f"step {it0 + ntdec_s - 1}: V fields bit-identical to FP32; rel L2 vs FP64 "
              + " ".join(f"{n}={e:.2e}" for n, e in zip(O.FIELDS_V, row)), flush=True)
    # seismograms: (nst, 3, ntw)
    for kind in ("vel", "disp"):
        got = g.seismograms(kind)
        w32, w64 = runs[32].wav[kind], runs[64].wav[kind]
        assert got.shape == w32.shape
        errs = np.zeros(got.shape[:2])
        for s in range(got.shape[0]):
            for c in range(3):
                assert _rel(got[s, c], w32[s, c]) <= REL_TOL, (kind, s, c)
                errs[s, c] = _rel(got[s, c], w64[s, c])
        assert ulp_diff(got, w32) == 0, kind
        assert np.all(np.abs(w64).max(axis=2) > 0), kind
        print(f"{kind} traces (16 stations x 3 components), rel L2 vs FP64: median {np.median(errs):.2e} "
              f"max {errs.max():.2e} (station {np.unravel_index(errs.argmax(), errs.shape)})", flush=True)
        for s in range(got.shape[0]):
            print(f"  {kind} station {s:2d}: " + " ".join(f"{errs[s, c]:.2e}" for c in range(3)), flush=True)
    for name in SECTIONS:
        errs = []
        for f in range(nframe):
            got = g.snapshot(name, f)
            want = runs[32].snap_frames[name][f]
            assert _rel(got, want) <= REL_TOL and ulp_diff(got, want) == 0, (name, f)
            errs.append(_rel(got, runs[64].snap_frames[name][f]) if np.any(runs[64].snap_frames[name][f]) else 0.0)
        print(f"{name} frames 0..{nframe - 1} rel L2 vs FP64: " + " ".join(f"{e:.2e}" for e in errs), flush=True)
    g.close()
