"""medium__setup on the device (swpc_medium_layered) against the host library.

The device builds rho, lam, mu, taup, taus, the free-surface and ocean-bottom
maps and their 2nd-order bands from the per-plane profile of the layered
model, the tau-method scalars and an optional random perturbation
(src/swpc_3d/m_medium.f90:124-427, m_vmodel_lhm.f90:20-158,
m_vmodel_lgm.f90:20-174, m_vmodel_lgm_rmed.f90:195-203).  The host library
(openswpc_amd/host/swpc_host.f90, through host.build_inputs) computes the
same arrays on the CPU; the two must agree bit for bit in every array, map,
velocity extreme, station cell and kernel coefficient, and a few steps of
the two simulations must give the same bits.  Cases: linear gradients with a
von Karman perturbation (NM = 3), homogeneous layers under seawater so the
ocean bottom lies below the free surface, an elastic model with a Gaussian
anomaly uploaded from the host, and a perturbation strong enough to clip at
the stability velocity, the cut-off velocity and the minimum density.  The
random media are generated on the device: into a host array for the host
library, into the context's staging buffer for the device path."""
import numpy as np
import pytest

from openswpc_amd import workloads as W

pytestmark = pytest.mark.gpu

MED = ("rho", "lam", "mu", "taup", "taus")
MAPS = ("kfs_top", "kfs_bot", "kob_top", "kob_bot", "kob")


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _spec(case):
    from openswpc_amd.host import ModelSpec
    src = [dict(x=0.4, y=-0.7, z=4.3, t0=0.0, tr=0.8, mo=1e15, strike=30.0, dip=60.0, rake=90.0)]
    st = [(2.1, 1.3, 0.0, "fsb"), (-3.2, 2.7, 0.0, "fsb"), (0.7, -2.9, 0.0, "obb"), (1.5, 0.5, 3.0, "dep")]
    nx, ny, nz = 44, 40, 36
    base = dict(nx=nx, ny=ny, nz=nz, dx=0.5, dy=0.5, dz=0.5, dt=0.02, nt=20, na=8, sources=src, stations=st,
                wav_what=3, fq_min=0.05, fq_max=5.0, fq_ref=1.0)
    if case == "lgm_vonkarman":
        return ModelSpec(**base, nm=3, zbeg=-1.0, vmodel="lgm",
                         rmedia=dict(ptype="vonkarman", a=(2.0, 2.5, 1.5), epsil=0.05, kappa=0.5, seed=3),
                         layers=[[0.0, 2.3, 3.464, 2.0, 200.0, 100.0], [12.0, 3.2, 7.794, 4.5, 450.0, 225.0]])
    if case == "lhm_ocean":
        return ModelSpec(**base, nm=3, zbeg=-1.5, vmodel="lhm",
                         layers=[[2.0, 2.5, 5.0, 2.9, 300.0, 150.0], [6.0, 2.8, 6.2, 3.5, 350.0, 175.0]])
    if case == "elastic_anomaly":
        xi = W.gauss_xi(nx, ny, nz, 0.5, 0.5, 0.5, -1.0, (1.0, -1.5, 6.0), -0.08, 3.0)
        return ModelSpec(**{**base, "na": 6}, nm=0, zbeg=-1.0, vmodel="lhm", xi=xi,
                         layers=[[0.0, 2.4, 4.5, 2.6, 1e6, 1e6], [4.0, 2.6, 5.6, 3.2, 1e6, 1e6]])
    if case == "clipped":
        return ModelSpec(**base, nm=2, zbeg=-1.0, vmodel="lgm", vcut=1.8,
                         rmedia=dict(ptype="exp", a=(1.5, 1.5, 1.5), epsil=0.6, seed=5),
                         layers=[[0.0, 1.2, 3.464, 2.0, 200.0, 100.0], [12.0, 3.2, 7.794, 4.5, 450.0, 225.0]])
    raise ValueError(case)


@pytest.mark.parametrize("case", ["lgm_vonkarman", "lhm_ocean", "elastic_anomaly", "clipped"])
def test_device_medium_equals_host_library(case):
    from openswpc_amd.host import build_inputs, make_simulation, make_simulation_device
    spec = _spec(case)
    inp = build_inputs(spec)
    a = inp.arrays
    gd, inp_d = make_simulation_device(spec)
    for n in MED:
        got = gd.get_medium(n)
        ne = int(np.count_nonzero(_bits(got) != _bits(a[n])))
        assert ne == 0, (case, n, ne, float(np.max(np.abs(got - a[n]))))
    for n in MAPS:
        assert np.array_equal(gd.get_medium(n), a[n]), (case, n)
    for n in ("kfs", "kob", "ist", "jst", "kst", "c1", "c2", "d1", "ts"):
        assert np.array_equal(inp_d.arrays[n], a[n]), (case, n)
    for n in ("vmin", "vmax", "zeta", "d2"):
        assert np.float32(inp_d.scalars[n]) == np.float32(inp.scalars[n]), (case, n)
    if case == "lhm_ocean":
        kfs, kob = a["kfs"][3:-3, 3:-3], a["kob"][3:-3, 3:-3]
        assert np.all(kob > kfs) and np.all(kfs > 0), "the ocean bottom lies below the free surface"
    if case == "clipped":
        # the perturbation takes every clip branch: vp above the stability
        # velocity, vs below the cut-off, rho below 1
        import ctypes as C
        from openswpc_amd.host import _perturb_clip, _perturbation, host_lib
        prof = np.zeros((5, spec.nz + 6), np.float32)
        rc = C.c_int()
        lay = np.ascontiguousarray(np.asarray(spec.layers, np.float32).reshape(-1))
        host_lib().swpch_vmodel_profile(1, spec.nz, spec.zbeg, spec.dz, spec.vcut, 0, len(spec.layers), lay, prof,
                                        C.byref(rc))
        r, la, m = prof[0], prof[1], prof[2]
        solid = m > 1e-6
        vp = np.sqrt((la + 2 * m) / r)[None, None, :]
        vs = np.sqrt(m / r)[None, None, :]
        vmin_c, vmax_c, _ = _perturb_clip(spec)
        e = _perturbation(spec)
        sel = solid[None, None, :] & np.ones(e.shape, bool)
        assert np.count_nonzero(sel & (vp * (1 + e) > vmax_c)) > 0
        assert np.count_nonzero(sel & (vs * (1 + e) < vmin_c)) > 0
        assert np.count_nonzero(sel & (r[None, None, :] * (1 + 0.8 * e) < 1.0)) > 0
    # the two simulations step to the same bits
    gh = make_simulation(inp)
    gh.step(1, 20, use_graph=True, block=10)
    gd.step(1, 20, use_graph=True, block=10)
    for n in ("Vx", "Vy", "Vz", "Sxx", "Syz"):
        assert np.array_equal(_bits(gd.get_field(n)), _bits(gh.get_field(n))), (case, n)
    assert np.array_equal(_bits(gd.seismograms("vel")), _bits(gh.seismograms("vel"))), case
    assert np.any(gh.get_field("Vz")), case
    gd.close()
    gh.close()
