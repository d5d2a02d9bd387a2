"""Pins of the oracle's discrete adjoint (oracle/adjoint.c), FP64 variant:

  * dot-product test: for one whole step M (stress then velocity update, PML
    auxiliaries and Zener memory variables included), <M x, y> = <x, M^T y>
    for random x, y, to rounding;
  * the misfit gradient with respect to lam, mu and rho against central
    finite differences at seeded cells, interior and PML, elastic and with
    three memory variables.
"""
import numpy as np
import pytest

from oracle import oracle as O

OFF = O.OFF


def _cfg(**kw):
    base = dict(nx=22, ny=20, nz=18, na=5, nm=0, dt=0.02, dx=0.5, dy=0.5, dz=0.5, zbeg=-1.5,
                layers=[[0.0, 2.5, 5.0, 2.9, 1e6, 1e6], [3.0, 2.8, 6.2, 3.6, 1e6, 1e6]],
                sources=[dict(x=0.2, y=-0.4, z=3.6, t0=0.0, tr=0.5, mo=1.0, strike=30.0, dip=60.0, rake=90.0)],
                stations=[(1.6, 1.1, 0.0, "fsb"), (-2.1, 0.7, 0.0, "fsb"), (0.4, -1.9, 0.0, "fsb"),
                          (-0.9, -0.6, 2.0, "dep")])
    base.update(kw)
    return O.Config(**base)


def _inner(a, b, keys):
    return sum(float(np.sum(a[k].astype(np.float64) * b[k].astype(np.float64))) for k in keys)


def _state_keys(nm):
    return O.FIELDS_V + O.FIELDS_S + (O.FIELDS_R if nm else ())


def _visco(**kw):
    """three Zener mechanisms with finite Q (m_kernel.f90:222-238)"""
    base = dict(nm=3, layers=[[0.0, 2.5, 5.0, 2.9, 150.0, 75.0], [3.0, 2.8, 6.2, 3.6, 250.0, 125.0]])
    base.update(kw)
    return _cfg(**base)


OCEAN = dict(zbeg=-1.0, layers=[[1.5, 2.5, 5.0, 2.9, 150.0, 75.0], [4.0, 2.8, 6.2, 3.6, 250.0, 125.0]])


@pytest.mark.parametrize("nm,ocean", [(0, False), (3, False), (1, False), (8, False), (0, True), (3, True)])
def test_dot_product_one_step(orc, nm, ocean):
    """ocean: a seawater layer puts the ocean-bottom band (kob_top..kob_bot)
    below the free-surface band, so its 2nd-order stencils are transposed too"""
    kw = OCEAN if ocean else {}
    cfg = _cfg(sources=[], **kw) if nm == 0 else _visco(sources=[], nm=nm, **kw)
    m = O.build_model(cfg)
    if ocean:
        assert np.all(m.kob[OFF + 1:-OFF - 3, OFF + 1:-OFF - 3] > m.kfs[OFF + 1:-OFF - 3, OFF + 1:-OFF - 3])
    adj = O.OracleAdjoint(m, 64)
    run = adj.run
    rng = np.random.default_rng(21)
    x = {}
    for k in O.FIELDS_V + O.FIELDS_S:
        a = np.zeros_like(run.arr[k])
        a[OFF + 1:-(OFF + 4) + 3 or None, OFF + 1:-(OFF + 4) + 3 or None, OFF + 1:-(OFF + 4) + 3 or None] = 0.0
        a[3:-3, 3:-3, 3:-3] = rng.standard_normal(a[3:-3, 3:-3, 3:-3].shape)
        x[k] = a
    for k in O.FIELDS_ADE + (O.FIELDS_R if nm else ()):
        x[k] = rng.standard_normal(run.arr[k].shape)
    y = {}
    for k in O.FIELDS_V + O.FIELDS_S:
        a = np.zeros_like(run.arr[k])
        a[3:-3, 3:-3, 3:-3] = rng.standard_normal(a[3:-3, 3:-3, 3:-3].shape)
        y[k] = a
    for k in O.FIELDS_ADE + (O.FIELDS_R if nm else ()):
        y[k] = rng.standard_normal(run.arr[k].shape)
    # forward: M x
    for k, v in x.items():
        run.arr[k][...] = v
    L, p = run.L, run.ptr
    L.orc_kernel_update_stress(p)
    L.orc_absorb_update_stress(p)
    L.orc_kernel_update_vel(p)
    L.orc_absorb_update_vel(p)
    Mx = {k: run.arr[k].copy() for k in x}
    # adjoint: M^T y (the forward state only feeds the gradients)
    adj.new_adjoint()
    for k in O.FIELDS_V + O.FIELDS_S:
        adj.adj[k][...] = y[k]
    for k in O.FIELDS_ADE + (O.FIELDS_R if nm else ()):
        adj.adj["a" + k][...] = y[k]
    adj.adj_vel()
    adj.adj_stress()
    MTy = {k: adj.adj[k] for k in O.FIELDS_V + O.FIELDS_S}
    MTy.update({k: adj.adj["a" + k] for k in O.FIELDS_ADE + (O.FIELDS_R if nm else ())})
    keys = list(x)
    lhs, rhs = _inner(Mx, y, keys), _inner(x, MTy, keys)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), abs(rhs)), (lhs, rhs)
    if nm:
        # control: without the memory variables' adjoints the identity fails
        MTy.update({k: np.zeros_like(MTy[k]) for k in O.FIELDS_R})
        assert abs(lhs - _inner(x, MTy, keys)) > 1e-3 * abs(lhs)


def test_dot_product_detects_a_wrong_transpose(orc):
    """control: the same identity fails when one coefficient of the
    adjoint is planted wrong (here the PML auxiliaries' adjoints dropped)"""
    cfg = _cfg(sources=[])
    m = O.build_model(cfg)
    adj = O.OracleAdjoint(m, 64)
    run = adj.run
    rng = np.random.default_rng(22)
    x = {k: np.zeros_like(run.arr[k]) for k in O.FIELDS_V + O.FIELDS_S}
    for k in x:
        x[k][3:-3, 3:-3, 3:-3] = rng.standard_normal(x[k][3:-3, 3:-3, 3:-3].shape)
    for k in O.FIELDS_ADE:
        x[k] = rng.standard_normal(run.arr[k].shape)
    y = {k: v.copy() for k, v in x.items()}
    for k, v in x.items():
        run.arr[k][...] = v
    L, p = run.L, run.ptr
    L.orc_kernel_update_stress(p)
    L.orc_absorb_update_stress(p)
    L.orc_kernel_update_vel(p)
    L.orc_absorb_update_vel(p)
    Mx = {k: run.arr[k].copy() for k in x}
    adj.new_adjoint()
    for k in O.FIELDS_V + O.FIELDS_S:
        adj.adj[k][...] = y[k]
    for k in O.FIELDS_ADE:
        adj.adj["a" + k][...] = y[k]
    adj.adj_vel()
    adj.adj_stress()
    MTy = {k: adj.adj[k] for k in O.FIELDS_V + O.FIELDS_S}
    MTy.update({k: np.zeros_like(adj.adj["a" + k]) for k in O.FIELDS_ADE})   # the planted mistake
    keys = list(x)
    lhs, rhs = _inner(Mx, y, keys), _inner(x, MTy, keys)
    assert abs(lhs - rhs) > 1e-6 * max(abs(lhs), abs(rhs))


@pytest.fixture(scope="module", params=[0, 3], ids=["elastic", "nm3"])
def fd_case(orc, request):
    nt = 50
    cfg = _cfg(nt=nt) if request.param == 0 else _visco(nt=nt)
    m = O.build_model(cfg)
    rng = np.random.default_rng(5)
    d_obs = np.zeros((nt, m.nst, 3))
    J, (gl, gm, gr), d = O.misfit_and_gradient(m, nt, d_obs)
    # perturbed-data misfit so the residual is not the data itself
    d_obs = d * (1 + 0.3 * rng.standard_normal(d.shape))
    J, (gl, gm, gr), _ = O.misfit_and_gradient(m, nt, d_obs)
    return cfg, m, nt, d_obs, J, {"lam": gl, "mu": gm, "rho": gr}


def _misfit(m, nt, d_obs):
    adj = O.OracleAdjoint(m, 64)
    d = adj.forward(nt, keep=False)
    r = d - d_obs
    return 0.5 * float(np.sum(r * r))


# (parameter, i, j, k): interior cells near the source, one under the free
# surface, one in the x-slab PML and one in the bottom PML
CELLS = [("lam", 12, 10, 7), ("mu", 11, 9, 8), ("rho", 12, 9, 7), ("mu", 10, 11, 5), ("lam", 5, 10, 8),
         ("rho", 11, 10, 15), ("mu", 5, 9, 9), ("rho", 4, 10, 7)]


@pytest.mark.parametrize("par,i,j,k", CELLS)
def test_gradient_matches_central_differences(fd_case, par, i, j, k):
    cfg, m, nt, d_obs, J, g = fd_case
    arr = getattr(m, par)
    idx = (j + OFF, i + OFF, k + OFF)
    v0 = float(arr[idx])
    h = 1e-3 * v0
    vals = []
    for sgn in (+1, -1):
        arr[idx] = np.float32(v0 + sgn * h)
        hh = float(arr[idx]) - v0      # the step the float32 medium really takes
        vals.append((_misfit(m, nt, d_obs), hh))
    arr[idx] = np.float32(v0)
    (jp, hp), (jm, hm) = vals
    fd = (jp - jm) / (hp - hm)
    ad = g[par][idx]
    # the waves have reached the cell: its gradient is far above rounding
    assert abs(ad) > 1e-9 * np.abs(g[par]).max(), (par, i, j, k, ad)
    assert abs(fd - ad) <= 1e-4 * abs(ad), (par, i, j, k, fd, ad)
