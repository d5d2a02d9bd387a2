"""Seeded synthetic workloads of the 3D time loop, shared by the benchmark
and the tests: the ModelSpec inputs of the plain 128^3 run, the
1024x1024x512 production grid, the 64-member Green's-function batch and the
two adjoint grids.  Every seed is numpy.random.default_rng(seed)."""
from __future__ import annotations

import math
import sys

import numpy as np


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def production_spec(nx=1024, ny=1024, nz=512, seed_medium=3, seed_st=4, seed_src=5, with_xi=True):
    from .host import ModelSpec
    dx = 0.25
    vs0, vs1 = 2.0, 4.5
    layers = [[0.0, 2.3, vs0 * math.sqrt(3), vs0, 100.0 * vs0, 50.0 * vs0],
              [60.0, 3.2, vs1 * math.sqrt(3), vs1, 100.0 * vs1, 50.0 * vs1]]
    # dt = 0.85 x the 4th-order stability limit for the fastest perturbed Vp
    vmax = vs1 * math.sqrt(3) * 1.15
    hh = 1.0 / math.sqrt(3.0 / dx**2)
    dt = float(np.float32(math.floor(0.85 * (6.0 / 7.0) * hh / vmax * 1e4) / 1e4))
    rng = np.random.default_rng(seed_src)
    strike, dip, rake = rng.uniform(0, 360), rng.uniform(0, 90), rng.uniform(-180, 180)
    src = [dict(x=0.0, y=0.0, z=10.0, t0=0.0, tr=1.0, mo=1e15, strike=float(strike), dip=float(dip), rake=float(rake))]
    rs = np.random.default_rng(seed_st)
    half = nx * dx / 2 - 30 * dx
    st = [(float(x), float(y), 0.0, "fsb") for x, y in rs.uniform(-half, half, size=(256, 2))]
    # von Karman medium on the (k,i,j) halo grid, generated on the device when the medium is built
    rmedia = dict(ptype="vonkarman", a=(5.0, 5.0, 5.0), epsil=0.05, kappa=0.5, seed=seed_medium) if with_xi else None
    return ModelSpec(nx=nx, ny=ny, nz=nz, dx=dx, dy=dx, dz=dx, dt=dt, nt=3000, na=20, nm=3, zbeg=-1.0,
                     vmodel="lgm", layers=layers, rmedia=rmedia, sources=src, stations=st, ntdec_w=1, wav_what=1,
                     fq_min=0.05, fq_max=5.0, fq_ref=1.0)


def plain_spec(seed_src=1, seed_st=2):
    from .host import ModelSpec
    rng = np.random.default_rng(seed_src)
    strike, dip, rake = rng.uniform(0, 360), rng.uniform(0, 90), rng.uniform(-180, 180)
    src = [dict(x=0.0, y=0.0, z=10.0, t0=0.0, tr=1.0, mo=1e15, strike=float(strike), dip=float(dip), rake=float(rake))]
    rs = np.random.default_rng(seed_st)
    st = [(float(x), float(y), 0.0, "fsb") for x, y in rs.uniform(-21, 21, size=(16, 2))]
    layers = [[0.0, 2.7, 5.5, 3.2, 320.0, 160.0], [5.0, 2.9, 6.5, 3.7, 370.0, 185.0]]
    return ModelSpec(nx=128, ny=128, nz=128, dx=0.5, dy=0.5, dz=0.5, dt=0.02, nt=1000, na=20, nm=3, xbeg=-32.0,
                     ybeg=-32.0, zbeg=-2.0, vmodel="lhm", layers=layers, sources=src, stations=st,
                     fq_min=0.05, fq_max=5.0, fq_ref=1.0)


def batch_specs(n=64, seed_sizes=6, seed_prof=10, seed_src=11, seed_st=12, nt=1500):
    """Batched Green's functions: n independent runs of uneven size, NM=2,
    10-cell PML, one unit force along a seeded axis at a seeded depth, 8
    surface receivers, each on one of four seeded 1D layered profiles."""
    from .host import ModelSpec
    rs = np.random.default_rng(seed_sizes)
    sizes = np.stack([rs.integers(96, 257, n), rs.integers(96, 257, n), rs.integers(96, 193, n)], axis=1)
    rp = np.random.default_rng(seed_prof)
    profiles = []
    for _ in range(4):
        tops = np.concatenate([[0.0], np.sort(rp.uniform(2.0, 30.0, 2))])
        vs = np.sort(rp.uniform(2.0, 4.5, 3))
        profiles.append([[float(t), float(2.2 + 0.2 * v), float(v * math.sqrt(3)), float(v), 100.0 * v, 50.0 * v]
                         for t, v in zip(tops, vs)])
    rsrc = np.random.default_rng(seed_src)
    rst = np.random.default_rng(seed_st)
    dx, dt = 0.5, 0.025
    specs = []
    for m in range(n):
        nx, ny, nz = (int(v) for v in sizes[m])
        axis = int(rsrc.integers(0, 3))
        f = [0.0, 0.0, 0.0]
        f[axis] = 1.0
        depth = float(rsrc.uniform(2.0, min(20.0, (nz - 15) * dx)))
        src = [dict(x=0.0, y=0.0, z=depth, t0=0.0, tr=1.0, fx=f[0], fy=f[1], fz=f[2])]
        hx, hy = (nx / 2 - 14) * dx, (ny / 2 - 14) * dx
        st = [(float(rst.uniform(-hx, hx)), float(rst.uniform(-hy, hy)), 0.0, "fsb") for _ in range(8)]
        specs.append(ModelSpec(nx=nx, ny=ny, nz=nz, dx=dx, dy=dx, dz=dx, dt=dt, nt=nt, na=10, nm=2, zbeg=-1.0,
                               vmodel="lhm", layers=profiles[m % 4], sources=src, bf_mode=True, stations=st,
                               fq_min=0.05, fq_max=5.0, fq_ref=1.0))
    return specs

def gauss_xi(nx, ny, nz, dx, dy, dz, zbeg, centre, amp, radius):
    """relative perturbation amp * exp(-r^2 / radius^2) about centre (x, y, z
    in km) on the (k,i,j) halo layout of an nx x ny x nz grid"""
    cx, cy, cz = centre
    k = np.arange(nz + 6) - 2
    i = np.arange(nx + 6) - 2
    j = np.arange(ny + 6) - 2
    z = zbeg + (k - 0.5) * dz
    x = -nx / 2 * dx + (i - 0.5) * dx
    y = -ny / 2 * dy + (j - 0.5) * dy
    r2 = (y[:, None, None] - cy) ** 2 + (x[None, :, None] - cx) ** 2 + (z[None, None, :] - cz) ** 2
    return (amp * np.exp(-r2 / radius ** 2)).astype(np.float32)


def adjoint_specs(seed_anom=7, seed_src=13, seed_st=14, nt=2000):
    """Adjoint gradient workload: 256 x 256 x 192, dx = 0.2 km, elastic,
    20-cell PML.  True model = layered + a seeded Gaussian velocity anomaly
    (+/- 8 %, radius 3 km); initial model = the layers alone; one seeded
    moment tensor, 64 seeded surface receivers."""
    from .host import ModelSpec
    nx, ny, nz, dx = 256, 256, 192, 0.2
    layers = [[0.0, 2.4, 4.5, 2.6, 1e6, 1e6], [4.0, 2.6, 5.6, 3.2, 1e6, 1e6], [15.0, 2.9, 6.6, 3.8, 1e6, 1e6]]
    ra = np.random.default_rng(seed_anom)
    sign = 1.0 if ra.uniform() < 0.5 else -1.0
    cx, cy, cz = ra.uniform(-8, 8), ra.uniform(-8, 8), ra.uniform(8, 20)
    xi = gauss_xi(nx, ny, nz, dx, dx, dx, -1.0, (cx, cy, cz), sign * 0.08, 3.0)
    rs = np.random.default_rng(seed_src)
    src = [dict(x=float(rs.uniform(-5, 5)), y=float(rs.uniform(-5, 5)), z=float(rs.uniform(10, 20)), t0=0.0, tr=1.0,
                mo=1e15, strike=float(rs.uniform(0, 360)), dip=float(rs.uniform(0, 90)),
                rake=float(rs.uniform(-180, 180)))]
    rr = np.random.default_rng(seed_st)
    half = nx * dx / 2 - 25 * dx
    st = [(float(a), float(b), 0.0, "fsb") for a, b in rr.uniform(-half, half, size=(64, 2))]
    common = dict(nx=nx, ny=ny, nz=nz, dx=dx, dy=dx, dz=dx, dt=0.012, nt=nt, na=20, nm=0, zbeg=-1.0, vmodel="lhm",
                  layers=layers, sources=src, stations=st, wav_what=1, fq_min=0.05, fq_max=5.0, fq_ref=1.0)
    return ModelSpec(xi=xi, **common), ModelSpec(**common), dict(sign=sign, centre=(cx, cy, cz))


# layers of the multi-source adjoint: depth rho vp vs qp qs, Qs = 50 Vs, Qp = 2 Qs
MULTI_LAYERS = [[0.0, 2.4, 4.5, 2.6, 260.0, 130.0], [4.0, 2.6, 5.6, 3.2, 320.0, 160.0],
                [15.0, 2.9, 6.6, 3.8, 380.0, 190.0]]


def adjoint_multi_geometry(nsrc=16, nrec=128, extent=(76.8, 76.8, 51.2), seed_anom=21, seed_src=9, seed_st=23):
    """The seeded part of the multi-source adjoint, in km and independent of
    the grid: the Gaussian anomaly (+/- 8 %, radius 5 km) of the true model,
    nsrc moment tensors (Kupper, 1 s) at 8-25 km depth, and nrec surface
    receivers per source."""
    lx, ly, _ = extent
    ra = np.random.default_rng(seed_anom)
    sign = 1.0 if ra.uniform() < 0.5 else -1.0
    centre = (ra.uniform(-15, 15), ra.uniform(-15, 15), ra.uniform(8, 25))
    rs = np.random.default_rng(seed_src)
    sx, sy = lx / 2 - 13.0, ly / 2 - 13.0
    srcs = [dict(x=float(rs.uniform(-sx, sx)), y=float(rs.uniform(-sy, sy)), z=float(rs.uniform(8, 25)), t0=0.0,
                 tr=1.0, mo=1e15, strike=float(rs.uniform(0, 360)), dip=float(rs.uniform(0, 90)),
                 rake=float(rs.uniform(-180, 180))) for _ in range(nsrc)]
    rr = np.random.default_rng(seed_st)
    hx, hy = lx / 2 - 5.0, ly / 2 - 5.0
    stations = [[(float(a), float(b), 0.0, "fsb") for a, b in
                 zip(rr.uniform(-hx, hx, nrec), rr.uniform(-hy, hy, nrec))] for _ in range(nsrc)]
    return dict(sign=sign, centre=centre, sources=srcs, stations=stations)


def adjoint_multi_spec(geo, n=(384, 384, 256), d=(0.2, 0.2, 0.2), dt=0.012, nt=2500, na=20, true=False):
    """ModelSpec of the multi-source adjoint grid with the first source and
    its receivers (the others are switched in with set_acquisition)"""
    from .host import ModelSpec
    nx, ny, nz = n
    xi = gauss_xi(nx, ny, nz, *d, -1.0, geo["centre"], geo["sign"] * 0.08, 5.0) if true else None
    return ModelSpec(nx=nx, ny=ny, nz=nz, dx=d[0], dy=d[1], dz=d[2], dt=dt, nt=nt, na=na, nm=3, zbeg=-1.0,
                     vmodel="lhm", layers=MULTI_LAYERS, xi=xi, sources=[geo["sources"][0]],
                     stations=geo["stations"][0], wav_what=1, fq_min=0.05, fq_max=5.0, fq_ref=1.0)
