"""The GPU discrete adjoint (swpc_adj_gradient) against the FP64 oracle
adjoint (itself pinned by the dot-product test and finite differences):
misfit and lam/mu/rho gradients on the same model and data, elastic and with
one, three and eight Zener mechanisms, and checkpoint spacing that leaves the result
bit-identical."""
import numpy as np
import pytest

from oracle import oracle as O
from gpu_helpers import gpu_from_model
from test_oracle_adjoint import _cfg, _visco

pytestmark = pytest.mark.gpu

OFF = O.OFF


@pytest.fixture(scope="module", params=[0, 3, 1, 8], ids=["elastic", "nm3", "nm1", "nm8"])
def case(orc, request):
    nt = 50
    cfg = _cfg(nt=nt, wav_flags=1) if request.param == 0 else _visco(nt=nt, wav_flags=1, nm=request.param)
    m = O.build_model(cfg)
    rng = np.random.default_rng(5)
    adj = O.OracleAdjoint(m, 64)
    d = adj.forward(nt, keep=False)
    d_obs = d * (1 + 0.3 * rng.standard_normal(d.shape))
    J, (gl, gm, gr), _ = O.misfit_and_gradient(m, nt, d_obs)
    obs = np.ascontiguousarray(d_obs.transpose(1, 2, 0), np.float32)   # (nst, 3, ntw)
    return m, nt, obs, J, {"lam": gl, "mu": gm, "rho": gr}


def _inner(a):
    return a[OFF + 1:-OFF - 3 or None, OFF + 1:-OFF - 3 or None, OFF + 1:-OFF - 3 or None].astype(np.float64)


def test_adjoint_gradient_matches_fp64_oracle(case):
    m, nt, obs, J, g = case
    gpu = gpu_from_model(m)
    Jg, gl, gm, gr = gpu.adjoint_gradient(nt, obs, ckp=9)
    gpu.close()
    assert abs(Jg - J) <= 1e-4 * abs(J), (Jg, J)
    for name, got in (("lam", gl), ("mu", gm), ("rho", gr)):
        want = _inner(g[name])
        got = _inner(got)
        rel = np.linalg.norm(got - want) / np.linalg.norm(want)
        print(name, rel, flush=True)
        assert np.linalg.norm(want) > 0
        assert rel <= 1e-4, (name, rel)


def test_checkpoint_spacing_does_not_change_the_result(case):
    m, nt, obs, J, g = case
    out = []
    for ckp in (7, 50):
        gpu = gpu_from_model(m)
        out.append(gpu.adjoint_gradient(nt, obs, ckp=ckp))
        gpu.close()
    (J1, *g1), (J2, *g2) = out
    assert J1 == J2
    for a, b in zip(g1, g2):
        assert np.array_equal(a, b)
