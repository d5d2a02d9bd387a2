"""Shared helpers for the GPU tests: feed oracle-built kernel inputs to the
HIP library through its C ABI and compare results."""
import numpy as np

from oracle import oracle as O


def gpu_from_model(m, kchunk=0, by=0, variant=-1):
    from openswpc_amd.gpu import SwpcGPU
    cfg = m.cfg
    g = SwpcGPU(cfg.nx, cfg.ny, cfg.nz, cfg.na, cfg.nm, cfg.abc, cfg.dx, cfg.dy, cfg.dz, cfg.dt)
    if kchunk or by or variant >= 0:
        g.set_tuning(kchunk, by, variant)
    g.set_medium(m.rho, m.lam, m.mu, m.taup, m.taus, m.kfs_top, m.kfs_bot, m.kob_top, m.kob_bot, m.kob)
    g.set_visco(m.c1, m.c2, m.d1, m.d2)
    if cfg.abc == "pml":
        g.set_pml(m.gxc, m.gxe, m.gyc, m.gye, m.gzc, m.gze)
    else:
        g.set_cerjan(m.cgx_c, m.cgx_b, m.cgy_c, m.cgy_b, m.cgz_c, m.cgz_b)
    if m.nsrc:
        g.set_sources(m.isrc, m.jsrc, m.ksrc, m.mo, m.mij, m.M0, m.stf, bf_mode=cfg.bf_mode)
    if m.nst:
        g.set_receivers(m.ist, m.jst, m.kst, cfg.wav_flags, cfg.ntdec_w, m.ntw, m.M0, m.UC)
    return g


def ulp_diff(a, b):
    """max distance in units in the last place between two float32 arrays"""
    a = np.ascontiguousarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.ascontiguousarray(b, np.float32).view(np.int32).astype(np.int64)
    a = np.where(a < 0, np.int64(-2**31) - a, a)
    b = np.where(b < 0, np.int64(-2**31) - b, b)
    return int(np.max(np.abs(a - b))) if a.size else 0


def compare_fields(g, run, names=O.FIELDS_V + O.FIELDS_S):
    out = {}
    for n in names:
        out[n] = (ulp_diff(g.get_field(n), run.field(n)), O.rel_l2(g.get_field(n), run.field(n)))
    return out


def small_config(**kw):
    base = dict(nx=40, ny=36, nz=32, na=8, nt=40, dx=0.5, dy=0.5, dz=0.5, dt=0.02, zbeg=-2.0,
                sources=[dict(x=0.3, y=-0.6, z=5.2, t0=0.05, tr=0.8, mo=1e15, strike=30.0, dip=60.0, rake=90.0)],
                stations=[(2.1, 1.3, 0.0, 'fsb'), (-3.2, 2.7, 0.0, 'fsb'), (0.7, -2.9, 6.0, 'dep')],
                wav_flags=15)
    base.update(kw)
    return O.Config(**base)


def oracle_config(spec, **kw):
    """the oracle's Config of a host ModelSpec (same grid, medium, sources,
    stations; wav_flags = spec.wav_what)"""
    import dataclasses
    names = {f.name for f in dataclasses.fields(O.Config)}
    base = {k: getattr(spec, k) for k in names if hasattr(spec, k)}
    base["wav_flags"] = spec.wav_what
    base.update(kw)
    return O.Config(**base)
