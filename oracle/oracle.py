"""CPU oracle for OpenSWPC's 3D velocity-stress loop -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module.  It is the checker the HIP library is compared with, never
part of the product path.

Two halves:
  * ``build_model(cfg)`` runs the oracle's restatement of the reference set-up
    (medium__setup -> kernel__setup -> source__setup -> absorb__setup,
    src/swpc_3d/main.f90:70-78) and returns every kernel input as float32 /
    int32 numpy arrays in the reference (k,i,j) halo-3 layout.
  * ``OracleRun`` integrates those inputs with the restated time loop
    (src/swpc_3d/main.f90:119-139) in one of three precisions: 32 (reference
    FP32 build), 16 (reference default mixed build), 64 (same scheme in FP64).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")
HALO = 3   # cells on each side of a halo array (indices -2..n+3)
OFF = 2    # numpy index of Fortran index 0

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")

STF_CODES = {"boxcar": 0, "triangle": 1, "herrmann": 2, "kupper": 3, "cosine": 4, "texp": 5}


def build_oracle() -> None:
    """Compile the three oracle variants (gcc, seconds)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_LIBS: dict = {}


def lib(prec: int = 32, vectorised: bool = False) -> C.CDLL:
    """the oracle library of a precision variant; vectorised: the FP32
    variant built for AVX2 hosts (same arithmetic, for timing)"""
    key = (prec, vectorised)
    if key in _LIBS:
        return _LIBS[key]
    if vectorised and prec != 32:
        raise ValueError("the vectorised build is FP32 only")
    path = os.path.join(BUILD, f"liborc{prec}{'v' if vectorised else ''}.so")
    if not os.path.exists(path):
        build_oracle()
    L = C.CDLL(path)
    f, i, d = C.c_float, C.c_int, C.c_double
    L.orc_x2i.restype = i
    L.orc_x2i.argtypes = [f, f, f]
    L.orc_i2x.restype = f
    L.orc_i2x.argtypes = [i, f, f]
    L.orc_fdm_stable_dt.restype = f
    L.orc_fdm_stable_dt.argtypes = [f, f, f, f]
    L.orc_visco_set_relaxtime.argtypes = [i, f, f, _f32p]
    L.orc_visco_constq_zeta.restype = f
    L.orc_visco_constq_zeta.argtypes = [i, f, f, _f32p]
    L.orc_kernel_coef.argtypes = [i, _f32p, f, _f32p, _f32p, _f32p, C.POINTER(C.c_float)]
    L.orc_momentrate.restype = f
    L.orc_momentrate.argtypes = [f, i, f, f]
    L.orc_stf_table.argtypes = [i, i, i, _f32p, _f32p, f, f, f, _f32p]
    L.orc_sdr2moment.argtypes = [f, f, f, _f32p]
    L.orc_seismic_moment.restype = f
    L.orc_seismic_moment.argtypes = [f]
    for nm in ("orc_vmodel_lhm", "orc_vmodel_lgm"):
        getattr(L, nm).restype = i
        getattr(L, nm).argtypes = [i, i, i, f, f, f, i, i, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p]
    L.orc_seawater_vel.restype = f
    L.orc_seawater_vel.argtypes = [f, i]
    L.orc_homogenize_absorber.argtypes = [i, i, i, i, _f32p, _f32p, _f32p, _f32p, _f32p]
    L.orc_q_to_tau.argtypes = [C.c_int64, i, f, _f32p, _f32p]
    L.orc_relaxed_medium.argtypes = [C.c_int64, i, _f32p, f, _f32p, _f32p, _f32p, _f32p]
    L.orc_surface_detection.argtypes = [i, i, i, _f32p, _f32p, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p]
    L.orc_kbeg_a.argtypes = [i, i, i, i, _i32p]
    L.orc_velocity_minmax.argtypes = [i, i, i, _f32p, _f32p, _f32p, _i32p, C.POINTER(C.c_float),
                                      C.POINTER(C.c_float)]
    L.orc_damping_profile.argtypes = [f, f, f, f, i, f, f, _f32p]
    L.orc_pml_profile_axis.argtypes = [i, i, f, f, f, f, _f32p, _f32p]
    L.orc_cerjan_profile_axis.argtypes = [i, i, f, i, _f32p, _f32p]
    vp = C.c_void_p
    for nm in ("orc_kernel_update_vel", "orc_kernel_update_stress", "orc_absorb_update_stress",
               "orc_absorb_update_vel", "orc_absorb_p_update_vel", "orc_absorb_p_update_stress",
               "orc_absorb_c_update_vel", "orc_absorb_c_update_stress"):
        getattr(L, nm).argtypes = [vp]
    L.orc_source_stressglut.argtypes = [vp, i, _i32p, _i32p, _i32p, _f32p, f, _f32p, _f32p, _f32p, _f32p,
                                        _f32p, _f32p, _f32p]
    L.orc_source_bodyforce.argtypes = [vp, i, _i32p, _i32p, _i32p, _f32p, _f32p, _f32p, f, _f32p]
    L.orc_wav_store.argtypes = [vp, i, i, _i32p, _i32p, _i32p, i, i, i, f, f, vp, vp, vp, vp, vp, vp]
    L.orc_kernel_vmax.argtypes = [vp, _f32p]
    L.orc_snap_write.argtypes = [vp, i, _i32p, _i32p, i, i, i, i, i, i, i, f, f, vp, vp, vp, vp]
    L.orc_green_store.argtypes = [vp, i, i, _i32p, _i32p, _i32p, i, i, i, vp, vp]
    L.orc_green_source.argtypes = [vp, i, i, i, f, f, f, f]
    L.orc_prec.restype = i
    _LIBS[key] = L
    return L


# ----------------------------------------------------------------------------
# Model description and set-up (reference order, src/swpc_3d/main.f90:64-78)
# ----------------------------------------------------------------------------
@dataclass
class Config:
    nx: int
    ny: int
    nz: int
    dx: float = 0.5
    dy: float = 0.5
    dz: float = 0.5
    dt: float = 0.02
    nt: int = 100
    na: int = 20
    nm: int = 3
    abc: str = "pml"                 # 'pml' | 'cerjan'
    xbeg: float | None = None        # default -nx/2*dx (m_global.f90:165)
    ybeg: float | None = None
    zbeg: float = -2.0
    tbeg: float = 0.0
    fq_min: float = 0.05
    fq_max: float = 5.0
    fq_ref: float = 1.0
    vmodel: str = "lhm"              # 'lhm' | 'lgm' | 'uni'
    layers: list = field(default_factory=lambda: [[0.0, 2.7, 5.5, 3.2, 320.0, 160.0],
                                                  [5.0, 2.9, 6.5, 3.7, 370.0, 185.0]])
    vcut: float = 0.0
    munk: bool = False               # seawater: Munk's profile instead of 1.5 km/s (m_seawater.f90)
    # sources: list of dict(x,y,z,t0,tr,mo, mxx..mxy) or body force dict(x,y,z,t0,tr,fx,fy,fz)
    sources: list = field(default_factory=list)
    bf_mode: bool = False
    stftype: str = "kupper"
    # receivers: list of (x, y, z, zsw) with zsw in {'dep','fsb','obb'}
    stations: list = field(default_factory=list)
    ntdec_w: int = 1
    wav_flags: int = 1               # bit0 V, bit1 U, bit2 stress, bit3 strain
    xi: np.ndarray | None = None     # optional relative perturbation field (k,i,j) halo layout

    def __post_init__(self):
        if self.xbeg is None:
            self.xbeg = -self.nx / 2 * float(np.float32(self.dx))
        if self.ybeg is None:
            self.ybeg = -self.ny / 2 * float(np.float32(self.dy))


def n3(cfg: Config) -> tuple[int, int, int]:
    """numpy shape (j, i, k) of a halo-3 (k,i,j) Fortran array."""
    return (cfg.ny + 2 * HALO, cfg.nx + 2 * HALO, cfg.nz + 2 * HALO)


def n2(cfg: Config) -> tuple[int, int]:
    return (cfg.ny + 2 * HALO, cfg.nx + 2 * HALO)


@dataclass
class Model:
    cfg: Config
    rho: np.ndarray
    lam: np.ndarray
    mu: np.ndarray
    taup: np.ndarray
    taus: np.ndarray
    kfs: np.ndarray
    kob: np.ndarray
    kfs_top: np.ndarray
    kfs_bot: np.ndarray
    kob_top: np.ndarray
    kob_bot: np.ndarray
    kbeg_a: np.ndarray
    ts: np.ndarray
    c1: np.ndarray
    c2: np.ndarray
    d1: np.ndarray
    d2: float
    zeta: float
    vmin: float
    vmax: float
    fcut: float
    M0: float
    UC: float
    gxc: np.ndarray
    gxe: np.ndarray
    gyc: np.ndarray
    gye: np.ndarray
    gzc: np.ndarray
    gze: np.ndarray
    cgx_c: np.ndarray
    cgx_b: np.ndarray
    cgy_c: np.ndarray
    cgy_b: np.ndarray
    cgz_c: np.ndarray
    cgz_b: np.ndarray
    # sources (normalised on the device side; raw single-precision values here)
    nsrc: int
    isrc: np.ndarray
    jsrc: np.ndarray
    ksrc: np.ndarray
    mo: np.ndarray
    mij: np.ndarray        # (6, nsrc): xx yy zz yz xz xy  or (3, nsrc) fx fy fz
    stf: np.ndarray        # (nt, nsrc) moment rate at the injection times
    # receivers
    nst: int
    ist: np.ndarray
    jst: np.ndarray
    kst: np.ndarray
    ntw: int


def _f(x):
    return np.float32(x)


def build_model(cfg: Config) -> Model:
    L = lib(32)
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    shp = n3(cfg)
    rho = np.zeros(shp, np.float32)
    lam = np.zeros(shp, np.float32)
    mu = np.zeros(shp, np.float32)
    qp = np.zeros(shp, np.float32)
    qs = np.zeros(shp, np.float32)
    layers = np.ascontiguousarray(np.asarray(cfg.layers, np.float32).reshape(-1))
    nl = len(cfg.layers)
    if cfg.vmodel == "lhm":
        rc = L.orc_vmodel_lhm(nx, ny, nz, cfg.zbeg, cfg.dz, cfg.vcut, int(cfg.munk), nl, layers, rho, lam, mu,
                              qp, qs)
    elif cfg.vmodel == "lgm":
        rc = L.orc_vmodel_lgm(nx, ny, nz, cfg.zbeg, cfg.dz, cfg.vcut, int(cfg.munk), nl, layers, rho, lam, mu,
                              qp, qs)
    else:
        raise ValueError(cfg.vmodel)
    if rc != 0:
        raise ValueError(f"vmodel failed rc={rc}")
    if cfg.xi is not None:
        # per-cell relative perturbation of vp, vs and 0.8 x of rho below the
        # free surface (the intent of m_vmodel_lgm_rmed.f90:195-203)
        xi = np.asarray(cfg.xi, np.float32)
        solid = mu > 0
        vp = np.sqrt((lam + 2 * mu) / rho)
        vs = np.sqrt(mu / rho)
        r1 = np.where(solid, rho * (1 + np.float32(0.8) * xi), rho).astype(np.float32)
        vp1 = np.where(solid, vp * (1 + xi), vp).astype(np.float32)
        vs1 = np.where(solid, vs * (1 + xi), vs).astype(np.float32)
        rho[...] = r1
        mu[...] = np.where(solid, r1 * vs1 * vs1, mu)
        lam[...] = np.where(solid, r1 * (vp1 * vp1 - 2 * vs1 * vs1), lam)
    na = cfg.na
    # absorber homogenisation (m_medium.f90:124-190); taup/taus still hold Q
    L.orc_homogenize_absorber(nx, ny, nz, na, rho, lam, mu, qp, qs)
    nm = cfg.nm
    ts = np.zeros(max(nm, 1), np.float32)
    L.orc_visco_set_relaxtime(nm, cfg.fq_min, cfg.fq_max, ts)
    zeta = L.orc_visco_constq_zeta(nm, cfg.fq_min, cfg.fq_max, ts)
    L.orc_q_to_tau(rho.size, nm, zeta, qp, qs)
    taup, taus = qp, qs
    L.orc_relaxed_medium(rho.size, nm, ts, cfg.fq_ref, lam, mu, taup, taus)
    kfs = np.zeros(n2(cfg), np.int32)
    kob = np.zeros_like(kfs)
    kfs_top = np.zeros_like(kfs)
    kfs_bot = np.zeros_like(kfs)
    kob_top = np.zeros_like(kfs)
    kob_bot = np.zeros_like(kfs)
    L.orc_surface_detection(nx, ny, nz, lam, mu, kfs, kob, kfs_top, kfs_bot, kob_top, kob_bot)
    vmin = C.c_float()
    vmax = C.c_float()
    L.orc_velocity_minmax(nx, ny, nz, rho, lam, mu, kfs, C.byref(vmin), C.byref(vmax))
    kbeg_a = np.zeros_like(kfs)
    L.orc_kbeg_a(nx, ny, nz, na, kbeg_a)
    # kernel__setup (m_kernel.f90:58-67)
    c1 = np.zeros(max(nm, 1), np.float32)
    c2 = np.zeros_like(c1)
    d1 = np.zeros_like(c1)
    d2 = C.c_float()
    L.orc_kernel_coef(nm, ts, cfg.dt, c1, c2, d1, C.byref(d2))

    # source__setup (m_source.f90:184-306)
    srcs = cfg.sources
    nsrc = len(srcs)
    f32 = np.float32
    dx32, dy32, dz32 = f32(cfg.dx), f32(cfg.dy), f32(cfg.dz)
    isrc = np.array([L.orc_x2i(s["x"], cfg.xbeg, dx32) for s in srcs], np.int32)
    jsrc = np.array([L.orc_x2i(s["y"], cfg.ybeg, dy32) for s in srcs], np.int32)
    ksrc = np.array([L.orc_x2i(s["z"], cfg.zbeg, dz32) for s in srcs], np.int32)
    t0 = np.array([s["t0"] for s in srcs], np.float32)
    tr = np.array([s["tr"] for s in srcs], np.float32)
    fcut = f32(0.0)
    for v in tr:
        fcut = max(fcut, f32(1) / v)
    UC = f32(1.0000001095e-15)  # 10.0**(-15) as evaluated by the Fortran runtime (m_global.f90:29)
    if cfg.bf_mode:
        fx = np.array([s["fx"] for s in srcs], f32)
        fy = np.array([s["fy"] for s in srcs], f32)
        fz = np.array([s["fz"] for s in srcs], f32)
        M0 = f32(math.sqrt(float(np.sum(fx * fx + fy * fy + fz * fz, dtype=f32)))) if nsrc else f32(1)
        UC = f32(UC * f32(1000))
        mij = np.stack([fx, fy, fz]) if nsrc else np.zeros((3, 0), f32)
        mo = np.ones(nsrc, f32)
        shift = 0.0
    else:
        mo = np.array([s["mo"] for s in srcs], f32)
        mij = np.zeros((6, nsrc), f32)
        for n, s in enumerate(srcs):
            if "strike" in s:
                m6 = np.zeros(6, f32)
                L.orc_sdr2moment(s["strike"], s["dip"], s["rake"], m6)
                mij[:, n] = m6
            else:
                mij[:, n] = [s["mxx"], s["myy"], s["mzz"], s["myz"], s["mxz"], s["mxy"]]
        M0 = f32(np.sum(mo, dtype=f32)) if nsrc else f32(1)
        shift = 0.5
    stf = np.zeros((cfg.nt, max(nsrc, 1)), f32)
    if nsrc:
        L.orc_stf_table(cfg.nt, nsrc, STF_CODES[cfg.stftype], t0, tr, cfg.tbeg, cfg.dt, shift, stf)
    stf = np.ascontiguousarray(stf[:, :nsrc])

    # absorb__setup (m_absorb_p.f90:60-93 / m_absorb_c.f90:27-101)
    gxc = np.zeros(4 * nx, f32); gxe = np.zeros_like(gxc)
    gyc = np.zeros(4 * ny, f32); gye = np.zeros_like(gyc)
    gzc = np.zeros(4 * nz, f32); gze = np.zeros_like(gzc)
    L.orc_pml_profile_axis(nx, na, cfg.xbeg, dx32, fcut, cfg.dt, gxc, gxe)
    L.orc_pml_profile_axis(ny, na, cfg.ybeg, dy32, fcut, cfg.dt, gyc, gye)
    L.orc_pml_profile_axis(nz, na, cfg.zbeg, dz32, fcut, cfg.dt, gzc, gze)
    cgx_c = np.zeros(nx + 2 * HALO, f32); cgx_b = np.zeros_like(cgx_c)
    cgy_c = np.zeros(ny + 2 * HALO, f32); cgy_b = np.zeros_like(cgy_c)
    cgz_c = np.zeros(nz + 2 * HALO, f32); cgz_b = np.zeros_like(cgz_c)
    L.orc_cerjan_profile_axis(nx, na, dx32, 1, cgx_c, cgx_b)
    L.orc_cerjan_profile_axis(ny, na, dy32, 1, cgy_c, cgy_b)
    L.orc_cerjan_profile_axis(nz, na, dz32, 0, cgz_c, cgz_b)

    # stations (m_wav.f90:141-271), 'xy' format
    ist, jst, kst = [], [], []
    for (x, y, z, zsw) in cfg.stations:
        i = L.orc_x2i(x, cfg.xbeg, dx32)
        j = L.orc_x2i(y, cfg.ybeg, dy32)
        k = L.orc_x2i(z, cfg.zbeg, dz32)
        xi1 = L.orc_i2x(1, cfg.xbeg, dx32); xin = L.orc_i2x(nx, cfg.xbeg, dx32)
        yj1 = L.orc_i2x(1, cfg.ybeg, dy32); yjn = L.orc_i2x(ny, cfg.ybeg, dy32)
        if not (xi1 < f32(x) < xin and yj1 < f32(y) < yjn and 1 < k < nz):
            continue
        if zsw == "fsb":
            k = int(kfs[j + OFF, i + OFF]) + 1
        elif zsw == "obb":
            k = int(kob[j + OFF, i + OFF]) + 1
        if k > nz:
            k = nz - 1
        if k < 1:
            k = 2
        ist.append(i); jst.append(j); kst.append(k)
    ntw = int(math.floor(float(f32(cfg.nt - 1) / f32(cfg.ntdec_w) + f32(1.0))))
    return Model(cfg=cfg, rho=rho, lam=lam, mu=mu, taup=taup, taus=taus, kfs=kfs, kob=kob,
                 kfs_top=kfs_top, kfs_bot=kfs_bot, kob_top=kob_top, kob_bot=kob_bot, kbeg_a=kbeg_a,
                 ts=ts[:nm].copy(), c1=c1, c2=c2, d1=d1, d2=float(d2.value), zeta=float(zeta),
                 vmin=float(vmin.value), vmax=float(vmax.value), fcut=float(fcut), M0=float(M0), UC=float(UC),
                 gxc=gxc, gxe=gxe, gyc=gyc, gye=gye, gzc=gzc, gze=gze,
                 cgx_c=cgx_c, cgx_b=cgx_b, cgy_c=cgy_c, cgy_b=cgy_b, cgz_c=cgz_c, cgz_b=cgz_b,
                 nsrc=nsrc, isrc=isrc, jsrc=jsrc, ksrc=ksrc, mo=mo, mij=np.ascontiguousarray(mij), stf=stf,
                 nst=len(ist), ist=np.array(ist, np.int32), jst=np.array(jst, np.int32),
                 kst=np.array(kst, np.int32), ntw=ntw)


# ----------------------------------------------------------------------------
# Time integration
# ----------------------------------------------------------------------------
def _state_type(sp_is_double: bool):
    sp = C.c_double if sp_is_double else C.c_float
    fields = ([(n, C.c_int) for n in ("nx", "ny", "nz", "na", "nm", "abc", "ibeg_k", "iend_k", "jbeg_k",
                                       "jend_k", "kbeg_k", "kend_k", "kbeg_min")]
              + [("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_float)]
              + [(n, C.c_void_p) for n in ("Vx", "Vy", "Vz", "Sxx", "Syy", "Szz", "Syz", "Sxz", "Sxy",
                                           "Rxx", "Ryy", "Rzz", "Ryz", "Rxz", "Rxy",
                                           "rho", "lam", "mu", "taup", "taus",
                                           "kfs_top", "kfs_bot", "kob_top", "kob_bot", "kbeg_a", "kob",
                                           "c1", "c2", "d1")]
              + [("d2", sp)]
              + [(n, C.c_void_p) for n in ("gxc", "gxe", "gyc", "gye", "gzc", "gze",
                                           "axVx", "ayVx", "azVx", "axVy", "ayVy", "azVy", "axVz", "ayVz",
                                           "azVz", "axSxx", "aySxy", "azSxz", "axSxy", "aySyy", "azSyz",
                                           "axSxz", "aySyz", "azSzz",
                                           "cgx_c", "cgx_b", "cgy_c", "cgy_b", "cgz_c", "cgz_b")])
    return type(f"OrcState_{'d' if sp_is_double else 'f'}", (C.Structure,), {"_fields_": fields})


FIELDS_V = ("Vx", "Vy", "Vz")
FIELDS_S = ("Sxx", "Syy", "Szz", "Syz", "Sxz", "Sxy")
FIELDS_R = ("Rxx", "Ryy", "Rzz", "Ryz", "Rxz", "Rxy")
FIELDS_ADE = ("axVx", "ayVx", "azVx", "axVy", "ayVy", "azVy", "axVz", "ayVz", "azVz",
              "axSxx", "aySxy", "azSxz", "axSxy", "aySyy", "azSyz", "axSxz", "aySyz", "azSzz")


def kernel_box(cfg: Config):
    """Interior (kernel) region of m_global.f90:345-376 for one rank."""
    if cfg.abc == "pml":
        return (cfg.na + 1, cfg.nx - cfg.na, cfg.na + 1, cfg.ny - cfg.na, 1, cfg.nz - cfg.na)
    return (1, cfg.nx, 1, cfg.ny, 1, cfg.nz)


class OracleRun:
    """Owns the state arrays of one oracle integration."""

    def __init__(self, model: Model, prec: int = 32, vectorised: bool = False):
        self.m = model
        self.prec = prec
        self.L = lib(prec, vectorised)
        cfg = model.cfg
        mp = np.float32 if prec == 32 else np.float64
        sp = np.float64 if prec == 64 else np.float32
        self.mp, self.sp = mp, sp
        shp = n3(cfg)
        self.arr = {}
        for n in FIELDS_V + FIELDS_S:
            self.arr[n] = np.zeros(shp, mp)
        ib, ie, jb, je, kb, ke = kernel_box(cfg)
        self.kbox = (ib, ie, jb, je, kb, ke)
        nm = cfg.nm
        rshape = (je - jb + 1, ie - ib + 1, ke - kb + 1, max(nm, 1))
        for n in FIELDS_R:
            self.arr[n] = np.zeros(rshape, sp)
        for n in ("rho", "lam", "mu", "taup", "taus"):
            self.arr[n] = np.ascontiguousarray(getattr(model, n).astype(sp))
        for n in ("kfs_top", "kfs_bot", "kob_top", "kob_bot", "kbeg_a", "kob"):
            self.arr[n] = np.ascontiguousarray(getattr(model, n))
        for n in ("c1", "c2", "d1"):
            self.arr[n] = np.ascontiguousarray(getattr(model, n).astype(sp))
        for n in ("gxc", "gxe", "gyc", "gye", "gzc", "gze", "cgx_c", "cgx_b", "cgy_c", "cgy_b", "cgz_c", "cgz_b"):
            self.arr[n] = np.ascontiguousarray(getattr(model, n).astype(sp))
        kbeg_min = 1
        self.kbeg_min = kbeg_min
        if cfg.abc == "pml":
            for n in FIELDS_ADE:
                self.arr[n] = np.zeros((cfg.ny, cfg.nx, cfg.nz - kbeg_min + 1), sp)
        st_t = _state_type(sp is np.float64)
        st = st_t()
        st.nx, st.ny, st.nz, st.na, st.nm = cfg.nx, cfg.ny, cfg.nz, cfg.na, nm
        st.abc = 0 if cfg.abc == "pml" else 1
        st.ibeg_k, st.iend_k, st.jbeg_k, st.jend_k, st.kbeg_k, st.kend_k = ib, ie, jb, je, kb, ke
        st.kbeg_min = kbeg_min
        st.dx, st.dy, st.dz = self._mpval(cfg.dx), self._mpval(cfg.dy), self._mpval(cfg.dz)
        st.dt = float(np.float32(cfg.dt))
        st.d2 = float(sp(model.d2))
        for n, a in self.arr.items():
            setattr(st, n, a.ctypes.data)
        self.st = st
        self.ptr = C.cast(C.pointer(st), C.c_void_p)
        # receivers
        self.uacc = np.zeros(3 * max(model.nst, 1), sp)
        self.eacc = np.zeros(6 * max(model.nst, 1), sp)
        self.wav = {k: np.zeros((max(model.nst, 1), nc, model.ntw), sp)
                    for k, nc in (("vel", 3), ("disp", 3), ("stress", 6), ("strain", 6))}
        self.it = 0

    def _mpval(self, v):
        # dx is real(MP): rounded to float in the FP32 build
        return float(np.float32(v)) if self.mp is np.float32 else float(v)

    def field(self, name: str) -> np.ndarray:
        return self.arr[name]

    # one step in the reference order (src/swpc_3d/main.f90:119-139)
    def step(self):
        m, L, p = self.m, self.L, self.ptr
        self.it += 1
        it = self.it
        self.green_store(it)
        self.wav_store(it)
        self.snap_write(it)
        L.orc_kernel_update_stress(p)
        L.orc_absorb_update_stress(p)
        if m.nsrc and not m.cfg.bf_mode:
            st = np.ascontiguousarray(m.stf[it - 1])
            L.orc_source_stressglut(p, m.nsrc, m.isrc, m.jsrc, m.ksrc, m.mo, m.M0, *[np.ascontiguousarray(m.mij[c])
                                                                                     for c in range(6)], st)
        L.orc_kernel_update_vel(p)
        L.orc_absorb_update_vel(p)
        if m.nsrc and m.cfg.bf_mode:
            st = np.ascontiguousarray(m.stf[it - 1])
            L.orc_source_bodyforce(p, m.nsrc, m.isrc, m.jsrc, m.ksrc, np.ascontiguousarray(m.mij[0]),
                                   np.ascontiguousarray(m.mij[1]), np.ascontiguousarray(m.mij[2]), m.M0, st)
        self.green_source(it)

    # ---- reciprocity Green's functions (m_green.f90) ----
    GREEN_CMP = {"x": (1.0, 0.0, 0.0), "y": (0.0, 1.0, 0.0), "z": (0.0, 0.0, 1.0)}

    def green_setup(self, ig, jg, kg, isrc, jsrc, ksrc, cmp, bforce, ntdec_w, ntw, stf):
        """GF points (ig, jg, kg); virtual source cell (isrc, jsrc, ksrc) and
        its unit force direction cmp in 'xyz' (m_green.f90:136-143); stf[it-1]
        is the moment rate at t = tbeg + it*dt (m_green.f90:632)."""
        self.green = dict(ig=np.ascontiguousarray(ig, np.int32), jg=np.ascontiguousarray(jg, np.int32),
                          kg=np.ascontiguousarray(kg, np.int32), src=(int(isrc), int(jsrc), int(ksrc)),
                          f=self.GREEN_CMP[cmp], bforce=int(bool(bforce)), ntdec_w=int(ntdec_w), ntw=int(ntw),
                          stf=np.ascontiguousarray(stf, np.float32))
        ng = self.green["ig"].size
        ncmp = 9 if bforce else 6
        self.green_acc = np.zeros(12 * max(ng, 1), self.sp)
        self.gf = np.zeros((max(ng, 1) * ncmp, int(ntw)), self.sp)    # gf(ntw, ncmp*ng), Fortran order

    def green_store(self, it):
        if not hasattr(self, "green"):
            return
        g = self.green
        self.L.orc_green_store(self.ptr, it, g["ig"].size, g["ig"], g["jg"], g["kg"], g["bforce"], g["ntdec_w"],
                               g["ntw"], self.green_acc.ctypes.data, self.gf.ctypes.data)

    def green_source(self, it):
        if not hasattr(self, "green"):
            return
        g = self.green
        if it > g["stf"].size:
            return
        self.L.orc_green_source(self.ptr, *g["src"], *g["f"], float(g["stf"][it - 1]))

    def wav_store(self, it):
        m = self.m
        if m.nst == 0:
            return
        vp = lambda a: a.ctypes.data
        self.L.orc_wav_store(self.ptr, it, m.nst, m.ist, m.jst, m.kst, m.cfg.wav_flags, m.cfg.ntdec_w, m.ntw,
                             m.M0, m.UC, vp(self.uacc), vp(self.eacc), vp(self.wav["vel"]), vp(self.wav["disp"]),
                             vp(self.wav["stress"]), vp(self.wav["strain"]))

    # ---- snapshots / sections (m_snap.f90) ----
    SNAP_PLANES = ("xy", "yz", "xz", "fs", "ob")
    SNAP_QS = ("ps", "v", "u")

    def snap_setup(self, sections, k0_xy=1, i0_yz=1, j0_xz=1, idec=1, jdec=1, kdec=1, ntdec_s=1, nframe=1):
        """sections: names like 'xz_ps', 'fs_v', 'ob_u' (m_snap.f90:105-119)"""
        cfg = self.m.cfg
        nxs, nys, nzs = ((n + d // 2) // d for n, d in ((cfg.nx, idec), (cfg.ny, jdec), (cfg.nz, kdec)))
        dims = {"xy": (nxs, nys), "yz": (nys, nzs), "xz": (nxs, nzs), "fs": (nxs, nys), "ob": (nxs, nys)}
        self.snap_sw = np.zeros(15, np.int32)
        self.snap_frames, self.snap_u, self.snap_vmax, self.snap_umax = {}, {}, {}, {}
        sp = self.sp
        for name in sections:
            pl, q = name.split("_")
            sec = self.SNAP_PLANES.index(pl) * 3 + self.SNAP_QS.index(q)
            self.snap_sw[sec] = 1
            ns1, ns2 = dims[pl]
            self.snap_frames[name] = np.zeros((nframe, 4 if q == "ps" else 3, ns2, ns1), sp)
        for pl in self.SNAP_PLANES:
            ns1, ns2 = dims[pl]
            self.snap_u[pl] = np.zeros((3, ns2, ns1), sp)
            self.snap_vmax[pl] = np.zeros((3, ns2, ns1), sp)
            self.snap_umax[pl] = np.zeros((3, ns2, ns1), sp)
        self.snap_par = (k0_xy, i0_yz, j0_xz, idec, jdec, kdec, ntdec_s)
        dummy = np.zeros(1, sp)
        self._snap_keep = [dummy]
        ptrs = []
        for sec in range(15):
            name = f"{self.SNAP_PLANES[sec // 3]}_{self.SNAP_QS[sec % 3]}"
            ptrs.append(self.snap_frames[name].ctypes.data if name in self.snap_frames else dummy.ctypes.data)
        self._snap_ptrs = (C.c_void_p * 15)(*ptrs)
        self._snap_u = (C.c_void_p * 5)(*[self.snap_u[p].ctypes.data for p in self.SNAP_PLANES])
        self._snap_vm = (C.c_void_p * 5)(*[self.snap_vmax[p].ctypes.data for p in self.SNAP_PLANES])
        self._snap_um = (C.c_void_p * 5)(*[self.snap_umax[p].ctypes.data for p in self.SNAP_PLANES])

    def snap_write(self, it):
        if not hasattr(self, "snap_sw"):
            return
        m = self.m
        k0, i0, j0, idec, jdec, kdec, nts = self.snap_par
        self.L.orc_snap_write(self.ptr, it, m.kfs, self.snap_sw, k0, i0, j0, idec, jdec, kdec, nts, m.M0, m.UC,
                              C.cast(self._snap_ptrs, C.c_void_p), C.cast(self._snap_u, C.c_void_p),
                              C.cast(self._snap_vm, C.c_void_p), C.cast(self._snap_um, C.c_void_p))

    def run(self, nsteps: int):
        for _ in range(nsteps):
            self.step()

    def vmax(self) -> np.ndarray:
        out = np.zeros(3, np.float32)
        self.L.orc_kernel_vmax(self.ptr, out)
        return out


def green_stf(cfg: Config, trise: float, nt: int | None = None) -> np.ndarray:
    """Moment rate of the virtual source at t = tbeg + it*dt, it = 1..nt,
    start time 0 and rise time trise (m_green.f90:25,632)."""
    nt = cfg.nt if nt is None else nt
    out = np.zeros(nt, np.float32)
    t0 = np.zeros(1, np.float32)
    tr = np.array([trise], np.float32)
    lib(32).orc_stf_table(nt, 1, STF_CODES[cfg.stftype], t0, tr, cfg.tbeg, cfg.dt, 0.0, out)
    return out


def interior(a: np.ndarray) -> np.ndarray:
    """Strip the 3-cell halo of a (j,i,k)-shaped numpy view of a (k,i,j) array."""
    return a[HALO:-HALO, HALO:-HALO, HALO:-HALO]


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b.ravel())
    if den == 0:
        return float(np.linalg.norm(a.ravel()))
    return float(np.linalg.norm((a - b).ravel()) / den)


# ----------------------------------------------------------------------------
# Discrete adjoint (elastic, PML): forward with every step kept, misfit,
# backward pass with gradients (adjoint.c)
# ----------------------------------------------------------------------------
FIELDS_AV = FIELDS_ADE[:9]
FIELDS_AS = FIELDS_ADE[9:]


def _adj_type(mp_double: bool, sp_double: bool):
    mp = C.c_void_p
    fields = ([(n, mp) for n in FIELDS_V + FIELDS_S]
              + [("aV", C.c_void_p * 9), ("aS", C.c_void_p * 9), ("D", C.c_void_p * 9)]
              + [("g_lam", C.c_void_p), ("g_mu", C.c_void_p), ("g_rho", C.c_void_p), ("aR", C.c_void_p * 6)])
    return type("OrcAdj", (C.Structure,), {"_fields_": fields})


class OracleAdjoint:
    """Misfit J = 1/2 sum (d - d_obs)^2 over the velocity traces of every step
    (the wav__store formula, m_wav.f90:517-531) and its gradient with respect
    to lam, mu and rho, by the discrete adjoint of the PML step with its
    memory variables."""

    def __init__(self, model: Model, prec: int = 64):
        cfg = model.cfg
        if cfg.abc != "pml":
            raise ValueError("the adjoint covers the PML scheme")
        self.run = OracleRun(model, prec)
        self.m = model
        L = self.run.L
        L.orc_adj_vel.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_adj_stress.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_adj_record.argtypes = [C.c_void_p, C.c_void_p, C.c_int, _i32p, _i32p, _i32p, C.c_float, C.c_float,
                                     np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")]
        self.L = L

    def traces(self):
        """velocity traces (nst, 3) of the current state"""
        m, a = self.m, self.run.arr
        sc = np.float64(np.float32(m.M0)) * np.float64(np.float32(m.UC)) * 1e9
        out = np.zeros((m.nst, 3))
        for n in range(m.nst):
            i, j, k = int(m.ist[n]) + OFF, int(m.jst[n]) + OFF, int(m.kst[n]) + OFF
            out[n, 0] = (a["Vx"][j, i, k] + a["Vx"][j, i - 1, k]) / 2 * sc
            out[n, 1] = (a["Vy"][j, i, k] + a["Vy"][j - 1, i, k]) / 2 * sc
            out[n, 2] = -(a["Vz"][j, i, k] + a["Vz"][j, i, k - 1]) / 2 * sc
        return out

    def forward(self, nt, keep=True):
        """nt steps from rest; returns traces (nt, nst, 3) of V_0 .. V_{nt-1}"""
        run, L, p, m = self.run, self.L, self.run.ptr, self.m
        self.A, self.B = [], []
        d = np.zeros((nt, m.nst, 3))
        for n in range(nt):
            it = n + 1
            d[n] = self.traces()
            if keep:
                self.A.append({k: run.arr[k].copy() for k in FIELDS_V + FIELDS_AV})
            L.orc_kernel_update_stress(p)
            L.orc_absorb_update_stress(p)
            if m.nsrc and not m.cfg.bf_mode:
                st = np.ascontiguousarray(m.stf[it - 1])
                L.orc_source_stressglut(p, m.nsrc, m.isrc, m.jsrc, m.ksrc, m.mo, m.M0,
                                        *[np.ascontiguousarray(m.mij[c]) for c in range(6)], st)
            if keep:
                self.B.append({k: run.arr[k].copy() for k in FIELDS_S + FIELDS_AS})
            L.orc_kernel_update_vel(p)
            L.orc_absorb_update_vel(p)
            if m.nsrc and m.cfg.bf_mode:
                st = np.ascontiguousarray(m.stf[it - 1])
                L.orc_source_bodyforce(p, m.nsrc, m.isrc, m.jsrc, m.ksrc, np.ascontiguousarray(m.mij[0]),
                                       np.ascontiguousarray(m.mij[1]), np.ascontiguousarray(m.mij[2]), m.M0, st)
        return d

    def new_adjoint(self):
        run = self.run
        shp = n3(self.m.cfg)
        self.adj = {k: np.zeros(shp, run.mp) for k in FIELDS_V + FIELDS_S}
        for k in FIELDS_ADE:
            self.adj["a" + k] = np.zeros_like(run.arr[k])
        self.dens = [np.zeros(shp, run.mp) for _ in range(9)]
        for k in FIELDS_R:
            self.adj["a" + k] = np.zeros_like(run.arr[k])
        self.grad = {k: np.zeros(shp) for k in ("lam", "mu", "rho")}
        st = _adj_type(True, True)()
        for k in FIELDS_V + FIELDS_S:
            setattr(st, k, self.adj[k].ctypes.data)
        for q, k in enumerate(FIELDS_AV):
            st.aV[q] = self.adj["a" + k].ctypes.data
        for q, k in enumerate(FIELDS_AS):
            st.aS[q] = self.adj["a" + k].ctypes.data
        for q in range(9):
            st.D[q] = self.dens[q].ctypes.data
        st.g_lam, st.g_mu, st.g_rho = (self.grad[k].ctypes.data for k in ("lam", "mu", "rho"))
        for q, k in enumerate(FIELDS_R):
            st.aR[q] = self.adj["a" + k].ctypes.data
        self.adj_st = st
        self.adj_ptr = C.cast(C.pointer(st), C.c_void_p)

    def adj_vel(self):
        self.L.orc_adj_vel(self.run.ptr, self.adj_ptr)

    def adj_stress(self):
        self.L.orc_adj_stress(self.run.ptr, self.adj_ptr)

    def backward(self, res):
        """res (nt, nst, 3) = d - d_obs; returns gradients (lam, mu, rho)"""
        nt = res.shape[0]
        run, m = self.run, self.m
        self.new_adjoint()
        for n in range(nt - 1, -1, -1):
            for k, v in self.B[n].items():
                run.arr[k][...] = v
            self.adj_vel()
            for k, v in self.A[n].items():
                run.arr[k][...] = v
            self.adj_stress()
            self.L.orc_adj_record(run.ptr, self.adj_ptr, m.nst, m.ist, m.jst, m.kst, m.M0, m.UC,
                                  np.ascontiguousarray(res[n], np.float64).reshape(-1))
        return self.grad["lam"], self.grad["mu"], self.grad["rho"]


def misfit_and_gradient(model: Model, nt: int, d_obs: np.ndarray, prec: int = 64):
    adj = OracleAdjoint(model, prec)
    d = adj.forward(nt)
    res = d - d_obs
    g = adj.backward(res)
    return 0.5 * float(np.sum(res * res)), g, d
