"""The reference's own known answer: example/example.out, the run of the
default 384^3 example (tests/example_case.py).

  * the set-up summary it prints (stability c, relaxed vmin / vmax, fmax, r;
    F15.3) from the oracle's set-up;
  * the max-|V| trace it prints every 50 steps (ES9.2) from the oracle's
    time loop, mixed variant (the reference default build) and FP32 variant
    (fixture tests/golden/example_vmax_oracle.txt, written by
    tests/golden/make_example_trace.py; SWPC_SLOW=1 re-runs the oracle and
    requires the same bit patterns).
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from example_case import NT, es92, example_config, report_steps

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def reference_trace():
    rows = {}
    for line in open(os.path.join(GOLD, "example_out_vmax.txt")):
        if line.startswith("#") or not line.strip():
            continue
        it, a, b, c = line.split()
        rows[int(it)] = (a, b, c)
    return rows


def oracle_trace(prec):
    rows = {}
    for line in open(os.path.join(GOLD, "example_vmax_oracle.txt")):
        if line.startswith("#") or not line.strip():
            continue
        p, it, h1, h2, h3, a, b, c = line.split()
        if int(p) == prec:
            bits = np.array([int(h, 16) for h in (h1, h2, h3)], np.uint32).view(np.float32)
            rows[int(it)] = (bits, (a, b, c))
    return rows


def test_reference_trace_fixture_is_complete():
    ref = reference_trace()
    assert sorted(ref) == report_steps(NT)


@pytest.mark.parametrize("prec", [16, 32])
def test_oracle_reproduces_example_out(prec):
    ref, orc = reference_trace(), oracle_trace(prec)
    assert sorted(orc) == sorted(ref)
    for it, (bits, printed) in orc.items():
        assert printed == ref[it], (prec, it, printed, ref[it])


def test_setup_summary_matches_example_out(orc):
    want = dict(l.split() for l in open(os.path.join(GOLD, "example_medium.txt")) if not l.startswith("#"))
    m = O.build_model(example_config())
    L = O.lib(32)
    f32 = np.float32
    dt2 = L.orc_fdm_stable_dt(f32(0.5), f32(0.5), f32(0.5), f32(m.vmax))
    c = f32(0.02) / f32(dt2)
    fmax = f32(2) * f32(m.fcut)
    r = (f32(m.vmin) / fmax) / f32(0.5)
    got = dict(c=c, vmin=m.vmin, vmax=m.vmax, fmax=fmax, r=r)
    for k, v in got.items():
        assert f"{float(v):.3f}" == want[k], (k, v, want[k])


@pytest.mark.skipif(os.environ.get("SWPC_SLOW") != "1", reason="25 min oracle run; set SWPC_SLOW=1")
def test_rerun_oracle_trace_bit_for_bit(orc):
    fix = oracle_trace(16)
    m = O.build_model(example_config())
    run = O.OracleRun(m, 16)
    for it in range(1, NT + 1):
        if it in fix:
            v = run.vmax()
            assert np.array_equal(v, fix[it][0]), it
            va = v * np.float32(m.UC) * np.float32(m.M0)
            assert tuple(es92(x) for x in va) == fix[it][1]
        if it < NT:
            run.step()


def test_vectorised_build_is_the_same_scheme(orc):
    """The AVX2 build of the FP32 oracle (the CPU baseline's) steps to the
    same bits as the parity build: vectorisation without contraction or
    fast-math keeps every rounding."""
    cfg = O.Config(nx=40, ny=36, nz=32, na=8, nt=30, dx=0.5, dy=0.5, dz=0.5, dt=0.02, zbeg=-1.0,
                   sources=[dict(x=0.3, y=-0.6, z=4.2, t0=0.0, tr=0.6, mo=1e15, strike=30.0, dip=60.0, rake=90.0)],
                   stations=[(2.1, 1.3, 0.0, "fsb")])
    m = O.build_model(cfg)
    a, b = O.OracleRun(m, 32), O.OracleRun(m, 32, vectorised=True)
    a.run(30)
    b.run(30)
    for n in O.FIELDS_V + O.FIELDS_S:
        assert np.array_equal(a.field(n).view(np.uint32), b.field(n).view(np.uint32)), n
    assert np.any(a.field("Vz"))
