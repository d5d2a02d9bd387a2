"""Gradients of several sources summed on the device (swpc_adj_reset /
swpc_adj_accumulate / swpc_adj_get_gradient) with two-level checkpointing,
three Zener mechanisms:

  * two-level schemes, and step records taken out of place, give the bits
    of the single-level scheme with copied records;
  * the device sum over two sources, each with its own source and receivers,
    matches the sum of the FP64 oracle's two gradients, and is bit for bit
    the float32 sum of the two single-source GPU gradients (each source's
    gradient is summed over its steps apart and added once);
  * the planner's schemes fit their budget, its memory model is what the
    library really allocates, and a smaller budget never costs fewer
    forward steps.
"""
import numpy as np
import pytest

from oracle import oracle as O
from gpu_helpers import gpu_from_model
from test_oracle_adjoint import _visco

pytestmark = pytest.mark.gpu

OFF = O.OFF
NT = 50

# the same rise time as the first source: the PML profiles depend on it
SRC_B = [dict(x=-1.1, y=0.9, z=4.4, t0=0.05, tr=0.5, mo=1.0, strike=140.0, dip=35.0, rake=-70.0)]
STA_B = [(-1.4, -1.8, 0.0, "fsb"), (2.2, -0.3, 0.0, "fsb"), (0.1, 1.9, 1.5, "dep")]


def _obs(m, seed):
    adj = O.OracleAdjoint(m, 64)
    d = adj.forward(NT, keep=False)
    rng = np.random.default_rng(seed)
    d_obs = d * (1 + 0.3 * rng.standard_normal(d.shape))
    J, (gl, gm, gr), _ = O.misfit_and_gradient(m, NT, d_obs)
    return np.ascontiguousarray(d_obs.transpose(1, 2, 0), np.float32), J, {"lam": gl, "mu": gm, "rho": gr}


@pytest.fixture(scope="module")
def two(orc):
    ma = O.build_model(_visco(nt=NT, wav_flags=1))
    mb = O.build_model(_visco(nt=NT, wav_flags=1, sources=SRC_B, stations=STA_B))
    for k in ("lam", "mu", "rho", "gxc", "gxe", "gyc", "gye", "gzc", "gze"):
        assert np.array_equal(getattr(ma, k), getattr(mb, k)), k
    oa, Ja, ga = _obs(ma, 5)
    ob, Jb, gb = _obs(mb, 6)
    return ma, mb, (oa, Ja, ga), (ob, Jb, gb)


def _inner(a):
    return a[OFF + 1:-OFF - 3, OFF + 1:-OFF - 3, OFF + 1:-OFF - 3].astype(np.float64)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _switch(g, m):
    g.set_sources(m.isrc, m.jsrc, m.ksrc, m.mo, m.mij, m.M0, m.stf, bf_mode=m.cfg.bf_mode)
    g.set_receivers(m.ist, m.jst, m.kst, m.cfg.wav_flags, m.cfg.ntdec_w, m.ntw, m.M0, m.UC)


def test_two_level_schemes_are_bit_identical(two):
    ma, _, (oa, _, _), _ = two
    # reference: single level, step records by copies (variant bit 4)
    h = gpu_from_model(ma, variant=1 | 16)
    J0, *g0 = h.adjoint_gradient(NT, oa, ckp=7)
    h.close()
    g = gpu_from_model(ma)
    for c1, c2 in ((20, 6), (50, 7), (13, 13), (16, 1), (1, 0)):
        g.adjoint_reset()
        J = g.adjoint_accumulate(NT, oa, c1, c2)
        got = g.adjoint_sums()
        print(c1, c2, J, flush=True)
        assert J == J0
        for a, b in zip(got, g0):
            assert np.array_equal(a, b), (c1, c2)
    g.adjoint_release()
    g.close()


def test_two_sources_sum_on_the_device(two):
    ma, mb, (oa, Ja, ga), (ob, Jb, gb) = two
    g = gpu_from_model(ma)
    g.adjoint_reset()
    Jga = g.adjoint_accumulate(NT, oa, 20, 6)
    _switch(g, mb)
    Jgb = g.adjoint_accumulate(NT, ob, 20, 6)
    sums = g.adjoint_sums()
    g.adjoint_release()
    g.close()
    assert abs(Jga - Ja) <= 1e-4 * Ja and abs(Jgb - Jb) <= 1e-4 * Jb, (Jga, Ja, Jgb, Jb)
    # each source alone, single level
    singles = []
    for m, o in ((ma, oa), (mb, ob)):
        h = gpu_from_model(m)
        singles.append(h.adjoint_gradient(NT, o, ckp=9)[1:])
        h.close()
    for q, name in enumerate(("lam", "mu", "rho")):
        want = _inner(ga[name]) + _inner(gb[name])
        got = _inner(sums[q])
        r64 = _rel(got, want)
        print(name, r64, flush=True)
        assert np.linalg.norm(_inner(gb[name])) > 0.02 * np.linalg.norm(_inner(ga[name]))
        assert r64 <= 1e-5, (name, r64)
        assert np.array_equal(sums[q], singles[0][q] + singles[1][q]), name
        # control: one source alone is far from the sum
        assert _rel(_inner(singles[0][q]), want) > 1e-2


def test_get_before_any_source_is_an_error(two):
    from openswpc_amd.gpu import SwpcError
    ma = two[0]
    g = gpu_from_model(ma)
    with pytest.raises(SwpcError):
        g.adjoint_sums()
    g.adjoint_reset()
    with pytest.raises(SwpcError):
        g.adjoint_sums()
    g.close()


def test_planner_fits_its_budget_and_matches_the_allocations(two):
    from openswpc_amd.gpu import SwpcError
    ma, _, (oa, _, _), _ = two
    g = gpu_from_model(ma)
    big = g.adjoint_plan(NT, 1 << 40)
    c1, c2, nbytes, nfwd = big
    # room for every step's record: the one forward pass records them all
    assert c1 == c2 == NT and nfwd == NT
    prev = nfwd
    plans = []
    b = nbytes - 1
    for _ in range(64):
        try:
            p = g.adjoint_plan(NT, b)
        except SwpcError:
            break
        assert p[2] <= b and p[1] <= p[0], (b, p)
        assert p[3] >= prev
        prev = p[3]
        plans.append(p)
        b = p[2] * 9 // 10
    print(plans, flush=True)
    assert plans and plans[0][3] > NT
    assert any(p[1] < p[0] for p in plans), "no two-level scheme before the budget ran out"
    with pytest.raises(SwpcError):
        g.adjoint_plan(NT, 1 << 20)
    # what the library allocates for a scheme is what the planner counted
    for c1, c2, nb, _ in (big, plans[-1]):
        g.adjoint_release()
        before = g.device_bytes()
        g.adjoint_reset()
        g.adjoint_accumulate(NT, oa, c1, c2)
        held = g.device_bytes() - before
        # besides the plan: the data, and the record table (24 nst + 1 words)
        assert held == nb + oa.size * 4 + (24 * ma.ist.size + 1) * 4, (c1, c2, held, nb)
    g.adjoint_release()
    assert g.device_bytes() == before
    g.close()
