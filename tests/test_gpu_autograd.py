"""The GPU discrete adjoint wrapped as a torch.autograd.Function (test
harness only): the misfit of the velocity traces as a differentiable function
of lam, mu and rho, its backward being swpc_adj_gradient.  Checked by the
chain rule through a torch reparametrisation and by central differences of
the GPU forward misfit along random directions."""
import dataclasses

import numpy as np
import pytest

from oracle import oracle as O
from gpu_helpers import gpu_from_model
from test_oracle_adjoint import _cfg

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

NT = 50


class SwpcMisfit(torch.autograd.Function):
    """J(lam, mu, rho) = 1/2 sum (d - obs)^2 over the recorded velocity
    samples of NT steps from rest; gradients from the discrete adjoint"""

    @staticmethod
    def forward(ctx, lam, mu, rho, model, obs):
        m = dataclasses.replace(model, lam=lam.detach().numpy(), mu=mu.detach().numpy(), rho=rho.detach().numpy())
        g = gpu_from_model(m)
        J, gl, gm, gr = g.adjoint_gradient(NT, obs, ckp=9)
        g.close()
        ctx.save_for_backward(torch.from_numpy(gl), torch.from_numpy(gm), torch.from_numpy(gr))
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        gl, gm, gr = ctx.saved_tensors
        s = gJ.to(torch.float32)
        return s * gl, s * gm, s * gr, None, None


def _misfit(m, obs):
    """the GPU forward alone: NT steps, traces, misfit in double"""
    g = gpu_from_model(m)
    g.step(1, NT)
    d = g.seismograms("vel").astype(np.float64)
    g.close()
    r = d - obs.astype(np.float64)
    return 0.5 * float(np.sum(r * r))


@pytest.fixture(scope="module")
def case(orc):
    m = O.build_model(_cfg(nt=NT, wav_flags=1))
    g = gpu_from_model(m)
    g.step(1, NT)
    d = g.seismograms("vel")
    g.close()
    rng = np.random.default_rng(9)
    obs = np.ascontiguousarray(d * (1 + 0.3 * rng.standard_normal(d.shape)), np.float32)
    return m, obs


def _params(m):
    return [torch.tensor(a, requires_grad=True) for a in (m.lam, m.mu, m.rho)]


def test_backward_is_the_adjoint_gradient_and_chains(case):
    m, obs = case
    lam, mu, rho = _params(m)
    J = SwpcMisfit.apply(lam, mu, rho, m, obs)
    J.backward()
    g = gpu_from_model(m)
    J2, gl, gm, gr = g.adjoint_gradient(NT, obs, ckp=9)
    g.close()
    assert float(J) == J2
    for t, want in ((lam, gl), (mu, gm), (rho, gr)):
        assert np.array_equal(t.grad.numpy(), want)
    assert np.linalg.norm(gm) > 0
    # a reparametrisation mu = mu0 * exp(s): dJ/ds = mu0 * dJ/dmu
    s = torch.zeros_like(mu, requires_grad=True)
    lam2, rho2 = (torch.tensor(a) for a in (m.lam, m.rho))
    mu2 = torch.tensor(m.mu) * torch.exp(s)
    SwpcMisfit.apply(lam2, mu2, rho2, m, obs).backward()
    assert np.allclose(s.grad.numpy(), m.mu * gm, rtol=1e-6, atol=0)


@pytest.mark.parametrize("name", ["lam", "mu", "rho"])
def test_directional_derivative_matches_central_differences(case, name):
    m, obs = case
    lam, mu, rho = _params(m)
    SwpcMisfit.apply(lam, mu, rho, m, obs).backward()
    grad = {"lam": lam, "mu": mu, "rho": rho}[name].grad.numpy().astype(np.float64)
    base = getattr(m, name)
    rng = np.random.default_rng(17)
    # a relative perturbation of the solid cells of the region the GPU adjoint
    # test compares (interior, free-surface band and PML; no halo)
    inner = (slice(O.OFF + 1, -O.OFF - 3),) * 3
    dirn = np.zeros(base.shape)
    dirn[inner] = np.where(m.mu[inner] > 0, base[inner] * rng.standard_normal(base[inner].shape), 0.0)
    h = 2e-3
    vals = []
    for sgn in (+1, -1):
        pert = (base.astype(np.float64) + sgn * h * dirn).astype(np.float32)
        step = pert.astype(np.float64) - base      # what the float32 medium really moved
        vals.append((_misfit(dataclasses.replace(m, **{name: pert}), obs), step))
    (jp, sp), (jm, sm) = vals
    fd = jp - jm
    ad = float(np.sum(grad * (sp - sm)))
    print(name, fd, ad, flush=True)
    assert abs(ad) > 0
    assert abs(fd - ad) <= 2e-3 * abs(ad), (name, fd, ad)
