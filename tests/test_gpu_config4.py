"""The elastic adjoint-gradient workload (256 x 256 x 192 cells, dx 0.2 km,
2000 steps, 20-cell PML, layered model, data from the layers plus a seeded
Gaussian anomaly of 8 % and 3 km radius, one seeded moment tensor, 64 seeded
receivers, checkpoints every 45 steps), through the product path
(host.build_inputs -> the C ABI).

  * At full size: the gradient is the same bits under single-level
    checkpoints every 45 steps and under two levels (outer 450, inner 45),
    nonzero, and its derivative along the anomaly's shape agrees with a
    central difference of the GPU forward misfit.
  * On the same workload subsampled to 64 x 64 x 48 cells of 0.8 km
    (10-cell PML, 200 steps of 0.05 s): the GPU gradient against the FP64
    oracle adjoint, and the oracle's gradient at 20 seeded cells against
    central differences of its own misfit."""
import dataclasses

import numpy as np
import pytest

from oracle import oracle as O
from gpu_helpers import oracle_config

pytestmark = pytest.mark.gpu

OFF = O.OFF
NT = 2000
CKP = 45


def _forward_misfit(inp, obs, nt, **arrays):
    """J = 1/2 sum (d - obs)^2 of the GPU forward with some medium arrays replaced"""
    from openswpc_amd.host import make_simulation
    sim = make_simulation(dataclasses.replace(inp, arrays={**inp.arrays, **arrays}))
    sim.step(1, nt, use_graph=True, block=100)
    d = sim.seismograms("vel").astype(np.float64)
    sim.close()
    r = d - obs.astype(np.float64)
    return 0.5 * float(np.sum(r * r))


@pytest.mark.timeout(1200)
def test_full_grid_checkpoint_schemes_and_anomaly_derivative(orc):
    from openswpc_amd import workloads as W
    from openswpc_amd.host import build_inputs, make_simulation

    true_spec, init_spec, info = W.adjoint_specs(nt=NT)
    sim = make_simulation(build_inputs(true_spec))
    sim.step(1, NT, use_graph=True, block=100)
    obs = sim.seismograms("vel")
    sim.close()
    assert obs.shape == (64, 3, NT) and np.abs(obs).max() > 0
    inp = build_inputs(init_spec)
    sim = make_simulation(inp)
    J1, *g1 = sim.adjoint_gradient(NT, obs, ckp=CKP)
    sim.adjoint_reset()
    J2 = sim.adjoint_accumulate(NT, obs, 10 * CKP, CKP)
    g2 = sim.adjoint_sums()
    sim.adjoint_release()
    sim.close()
    print("misfit", J1, J2, "gradient norms", [float(np.linalg.norm(x)) for x in g1], flush=True)
    assert J1 == J2 and J1 > 0
    inner = (slice(3, -3),) * 3
    for name, x, y in zip(("lam", "mu", "rho"), g1, g2):
        assert np.linalg.norm(x[inner]) > 0, name
        assert np.array_equal(x, y), name

    # the anomaly's shape as a relative perturbation of the solid cells
    s = W.gauss_xi(init_spec.nx, init_spec.ny, init_spec.nz, init_spec.dx, init_spec.dy, init_spec.dz,
                   init_spec.zbeg, info["centre"], 1.0, 3.0).astype(np.float64)
    a = inp.arrays
    solid = a["mu"] > 0
    h = 2e-3
    steps, vals = [], []
    for sgn in (+1, -1):
        pert = {n: np.where(solid, a[n] * (1 + sgn * h * s), a[n]).astype(np.float32) for n in ("lam", "mu", "rho")}
        steps.append({n: pert[n].astype(np.float64) - a[n] for n in pert})
        vals.append(_forward_misfit(inp, obs, NT, **pert))
    fd = vals[0] - vals[1]
    ad = sum(float(np.sum(g.astype(np.float64) * (steps[0][n] - steps[1][n])))
             for n, g in zip(("lam", "mu", "rho"), g1))
    print("directional derivative along the anomaly: central difference", fd, "adjoint", ad, flush=True)
    assert abs(ad) > 0
    assert abs(fd - ad) <= 5e-3 * abs(ad), (fd, ad)


SUB = dict(nx=64, ny=64, nz=48, dx=0.8, dy=0.8, dz=0.8, dt=0.05, nt=200, na=10)


@pytest.fixture(scope="module")
def subsampled(orc):
    from openswpc_amd import workloads as W
    true_spec, init_spec, info = W.adjoint_specs(nt=SUB["nt"])
    xi = W.gauss_xi(SUB["nx"], SUB["ny"], SUB["nz"], SUB["dx"], SUB["dy"], SUB["dz"], true_spec.zbeg,
                    info["centre"], info["sign"] * 0.08, 3.0)
    true_spec = dataclasses.replace(true_spec, xi=xi, **SUB)
    init_spec = dataclasses.replace(init_spec, **SUB)
    mt = O.build_model(oracle_config(true_spec))
    mi = O.build_model(oracle_config(init_spec))
    nt = SUB["nt"]
    d_obs = O.OracleAdjoint(mt, 64).forward(nt, keep=False)          # (nt, nst, 3)
    obs = np.ascontiguousarray(d_obs.transpose(1, 2, 0), np.float32)   # (nst, 3, ntw)
    d_obs = obs.transpose(2, 0, 1).astype(np.float64)                  # the data both sides use
    J, (gl, gm, gr), _ = O.misfit_and_gradient(mi, nt, d_obs)
    return init_spec, mi, obs, d_obs, J, {"lam": gl, "mu": gm, "rho": gr}


@pytest.mark.timeout(900)
def test_subsampled_gpu_gradient_matches_fp64_oracle(subsampled):
    from openswpc_amd.host import build_inputs, make_simulation
    spec, m, obs, d_obs, J, g = subsampled
    inp = build_inputs(spec)
    for n in ("lam", "mu", "rho"):
        assert np.array_equal(inp.arrays[n], getattr(m, n)), n
    sim = make_simulation(inp)
    Jg, gl, gm, gr = sim.adjoint_gradient(SUB["nt"], obs, ckp=CKP)
    sim.close()
    print("misfit GPU", Jg, "FP64 oracle", J, flush=True)
    assert abs(Jg - J) <= 1e-4 * abs(J)
    sl = (slice(OFF + 1, -OFF - 3),) * 3
    for name, got in (("lam", gl), ("mu", gm), ("rho", gr)):
        want = g[name][sl].astype(np.float64)
        rel = np.linalg.norm(got[sl].astype(np.float64) - want) / np.linalg.norm(want)
        print(name, "rel L2 vs FP64 oracle", rel, flush=True)
        assert np.linalg.norm(want) > 0
        assert rel <= 1e-4, (name, rel)


def _cells(n=20, seed=8):
    """20 seeded cells of the subsampled grid (between the absorbers, below
    the free surface), each with the parameter it is differentiated by"""
    rng = np.random.default_rng(seed)
    na = SUB["na"]
    out = []
    for c in range(n):
        i = int(rng.integers(na + 1, SUB["nx"] - na + 1))
        j = int(rng.integers(na + 1, SUB["ny"] - na + 1))
        k = int(rng.integers(4, SUB["nz"] - na + 1))
        out.append((("lam", "mu", "rho")[c % 3], i, j, k))
    return out


@pytest.mark.timeout(900)
def test_subsampled_oracle_gradient_matches_central_differences_at_20_cells(subsampled):
    spec, m, obs, d_obs, J, g = subsampled
    nt = SUB["nt"]
    for par, i, j, k in _cells():
        arr = getattr(m, par)
        idx = (j + OFF, i + OFF, k + OFF)
        v0 = float(arr[idx])
        vals = []
        for sgn in (+1, -1):
            arr[idx] = np.float32(v0 + sgn * 1e-3 * v0)
            hh = float(arr[idx]) - v0          # the step the float32 medium really takes
            r = O.OracleAdjoint(m, 64).forward(nt, keep=False) - d_obs
            vals.append((0.5 * float(np.sum(r * r)), hh))
        arr[idx] = np.float32(v0)
        (jp, hp), (jm, hm) = vals
        fd = (jp - jm) / (hp - hm)
        ad = float(g[par][idx])
        print(par, i, j, k, "central difference", fd, "adjoint", ad, flush=True)
        assert abs(ad) > 1e-7 * np.abs(g[par]).max(), (par, i, j, k)
        assert abs(fd - ad) <= 1e-4 * abs(ad), (par, i, j, k, fd, ad)
