"""Host-side set-up for libswpc_gpu, backed by the Fortran library
libswpc_host.so (openswpc_amd/host/swpc_host.f90).

``ModelSpec`` describes a run the way an OpenSWPC input.inf does (grid,
layered medium, sources, stations); ``build_inputs`` runs the reference's
set-up sequence (medium__setup -> kernel__setup -> source__setup ->
absorb__setup, src/swpc_3d/main.f90:70-78) through the Fortran routines and
returns the kernel inputs; ``make_simulation`` loads them into a SwpcGPU
context.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HOST_LIB = os.path.join(HERE, "lib", "libswpc_host.so")
HALO = 3
OFF = 2

STF_CODES = {"boxcar": 0, "triangle": 1, "herrmann": 2, "kupper": 3, "cosine": 4, "texp": 5}
ZSW = {"dep": 0, "fsb": 1, "obb": 2}
UC_SP = np.float32(1.0000001095e-15)   # 10.0**(-15) in single precision (m_global.f90:29)

_h = None


def host_lib() -> C.CDLL:
    global _h
    if _h is not None:
        return _h
    if not os.path.exists(HOST_LIB):
        raise RuntimeError(f"{HOST_LIB} is missing: build it with `make -C openswpc_amd/host`")
    L = C.CDLL(HOST_LIB)
    f, i = C.c_float, C.c_int
    fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    L.swpch_i2x.restype = f
    L.swpch_i2x.argtypes = [i, f, f]
    L.swpch_x2i.restype = i
    L.swpch_x2i.argtypes = [f, f, f]
    L.swpch_stable_dt.restype = f
    L.swpch_stable_dt.argtypes = [f, f, f, f]
    L.swpch_vmodel_layers.argtypes = [i, i, i, i, f, f, f, i, i, fp, fp, fp, fp, fp, fp, C.POINTER(C.c_int)]
    L.swpch_vmodel_profile.argtypes = [i, i, f, f, f, i, i, fp, fp, C.POINTER(C.c_int)]
    L.swpch_visco_setup.argtypes = [i, f, f, f, fp, C.POINTER(C.c_float), fp]
    L.swpch_seawater_vel.restype = f
    L.swpch_seawater_vel.argtypes = [f, i]
    L.swpch_perturb.argtypes = [i, i, i, fp, f, f, f, fp, fp, fp]
    L.swpch_medium_setup.argtypes = [i, i, i, i, i, f, f, f, fp, fp, fp, fp, fp, fp, C.POINTER(C.c_float),
                                     ip, ip, ip, ip, ip, ip, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.swpch_kernel_coef.argtypes = [i, fp, f, fp, fp, fp, C.POINTER(C.c_float)]
    L.swpch_pml_profile.argtypes = [i, i, f, f, f, f, fp, fp]
    L.swpch_cerjan_profile.argtypes = [i, i, f, i, fp, fp]
    L.swpch_sdr2moment.argtypes = [f, f, f, fp]
    L.swpch_momentrate.restype = f
    L.swpch_momentrate.argtypes = [f, i, f, f]
    L.swpch_stf_table.argtypes = [i, i, i, fp, fp, f, f, f, fp]
    L.swpch_station_index.argtypes = [i, fp, fp, fp, ip, i, i, i, f, f, f, f, f, f, ip, ip, ip, ip, ip, ip]
    L.swpch_green_points.argtypes = [i, fp, fp, fp, f, f, f, i, i, i, f, f, f, f, f, f, ip, ip, ip, ip, ip]
    _h = L
    return L


@dataclass
class GreenSpec:
    """The green_* parameters of input.inf (src/swpc_3d/m_green.f90:117-135).

    station: index into ModelSpec.stations of the virtual source (green_stnm);
    points: (ng, 3) x, y, z of the GF points (fn_glst with green_fmt = 'xyz');
    gid: their ids."""
    station: int
    points: np.ndarray
    gid: np.ndarray | None = None
    cmp: str = "z"              # green_cmp
    trise: float = 1.0          # green_trise
    bforce: bool = False        # green_bforce
    maxdist: float = 1e30       # green_maxdist


def read_glst(path: str):
    """fn_glst in green_fmt = 'xyz': 'x y z gid' per line, '#' comments and
    blank lines skipped (m_green.f90:196-215).  Returns (points, gid)."""
    pts, gid = [], []
    with open(path) as fh:
        for line in fh:
            t = line.strip()
            if not t or t.startswith("#"):
                continue
            w = t.split()
            if len(w) < 4:
                raise ValueError(f"{path}: expected 'x y z gid', got {t!r}")
            g = int(w[3])
            if not 0 <= g <= 99999999:
                raise ValueError(f"{path}: green id {g} out of range")
            pts.append([float(w[0]), float(w[1]), float(w[2])])
            gid.append(g)
    return np.array(pts, np.float32).reshape(-1, 3), np.array(gid, np.int64)


def green_export(gf: np.ndarray, cmp: str) -> np.ndarray:
    """Green's functions as green__export writes them: positive upward for
    a vertical virtual source (m_green.f90:563-566)."""
    return -gf if cmp == "z" else gf


@dataclass
class ModelSpec:
    nx: int
    ny: int
    nz: int
    dx: float = 0.5
    dy: float = 0.5
    dz: float = 0.5
    dt: float = 0.02
    nt: int = 1000
    na: int = 20
    nm: int = 3
    abc: str = "pml"
    xbeg: float | None = None
    ybeg: float | None = None
    zbeg: float = -2.0
    tbeg: float = 0.0
    fq_min: float = 0.05
    fq_max: float = 5.0
    fq_ref: float = 1.0
    vmodel: str = "lhm"          # 'lhm' (homogeneous layers) | 'lgm' (linear gradients)
    layers: list = field(default_factory=list)   # rows: depth rho vp vs qp qs
    vcut: float = 0.0
    munk: bool = False            # seawater sound speed: Munk's profile instead of 1.5 km/s
    xi: np.ndarray | None = None  # optional relative perturbation, (k,i,j) halo layout
    # or a random medium generated on the device (m_rmedia.f90 method, openswpc_amd.gpu.rmedia_3d):
    # dict(ptype='gauss'|'exp'|'vonkarman', a=(ax, ay, az) km, epsil, kappa, seed)
    rmedia: dict | None = None
    sources: list = field(default_factory=list)
    bf_mode: bool = False
    stftype: str = "kupper"
    stations: list = field(default_factory=list)   # (x, y, z, zsw)
    ntdec_w: int = 1
    wav_what: int = 1
    green: GreenSpec | None = None   # reciprocity Green's-function mode (green_mode)

    def __post_init__(self):
        if self.xbeg is None:
            self.xbeg = -self.nx / 2 * float(np.float32(self.dx))
        if self.ybeg is None:
            self.ybeg = -self.ny / 2 * float(np.float32(self.dy))


@dataclass
class Inputs:
    spec: ModelSpec
    arrays: dict
    scalars: dict


def _source_inputs(spec: ModelSpec, srcs: list):
    """grid cells, moment tensors or forces, M0, UC and the STF table of the
    sources (m_source.f90:84-303), and the PML cut-off frequency they set"""
    L = host_lib()
    f32 = np.float32
    nsrc = len(srcs)
    dx32, dy32, dz32 = f32(spec.dx), f32(spec.dy), f32(spec.dz)
    isrc = np.array([L.swpch_x2i(s["x"], spec.xbeg, dx32) for s in srcs], np.int32)
    jsrc = np.array([L.swpch_x2i(s["y"], spec.ybeg, dy32) for s in srcs], np.int32)
    ksrc = np.array([L.swpch_x2i(s["z"], spec.zbeg, dz32) for s in srcs], np.int32)
    t0 = np.array([s["t0"] for s in srcs], np.float32)
    tr = np.array([s["tr"] for s in srcs], np.float32)
    fcut = f32(0)
    for v in tr:
        fcut = max(fcut, f32(1) / v)
    UC = UC_SP
    if spec.bf_mode:
        comp = np.array([[s["fx"] for s in srcs], [s["fy"] for s in srcs], [s["fz"] for s in srcs]], np.float32)
        M0 = f32(math.sqrt(float(np.sum(comp * comp, dtype=f32)))) if nsrc else f32(1)
        UC = f32(UC * f32(1000))
        mo = np.ones(nsrc, np.float32)
        shift = 0.0
    else:
        mo = np.array([s["mo"] for s in srcs], np.float32)
        comp = np.zeros((6, nsrc), np.float32)
        for n, s in enumerate(srcs):
            if "strike" in s:
                m6 = np.zeros(6, np.float32)
                L.swpch_sdr2moment(s["strike"], s["dip"], s["rake"], m6)
                comp[:, n] = m6
            else:
                comp[:, n] = [s["mxx"], s["myy"], s["mzz"], s["myz"], s["mxz"], s["mxy"]]
        M0 = f32(np.sum(mo, dtype=f32)) if nsrc else f32(1)
        shift = 0.5
    stf = np.zeros((spec.nt, max(nsrc, 1)), np.float32)
    if nsrc:
        L.swpch_stf_table(spec.nt, nsrc, STF_CODES[spec.stftype], t0, tr, spec.tbeg, spec.dt, shift, stf)
    stf = np.ascontiguousarray(stf[:, :nsrc])
    arrays = dict(isrc=isrc, jsrc=jsrc, ksrc=ksrc, mo=mo, comp=np.ascontiguousarray(comp), stf=stf)
    return arrays, dict(fcut=fcut, M0=M0, UC=UC, nsrc=nsrc)


def _station_inputs(spec: ModelSpec, kfs: np.ndarray, kob: np.ndarray) -> dict:
    """grid cells of the stations inside the domain (src/swpc_3d/m_wav.f90:188-240,
    through swpch_station_index)"""
    L = host_lib()
    f32 = np.float32
    dx32, dy32, dz32 = f32(spec.dx), f32(spec.dy), f32(spec.dz)
    nx, ny, nz = spec.nx, spec.ny, spec.nz
    nst = len(spec.stations)
    if nst:
        xs = np.array([s[0] for s in spec.stations], np.float32)
        ys = np.array([s[1] for s in spec.stations], np.float32)
        zs = np.array([s[2] for s in spec.stations], np.float32)
        zsw = np.array([ZSW[s[3]] for s in spec.stations], np.int32)
        ist, jst, kst, inside = (np.zeros(nst, np.int32) for _ in range(4))
        L.swpch_station_index(nst, xs, ys, zs, zsw, nx, ny, nz, spec.xbeg, spec.ybeg, spec.zbeg, dx32, dy32, dz32,
                              kfs, kob, ist, jst, kst, inside)
        keep = inside.astype(bool)
        return dict(ist=ist[keep].copy(), jst=jst[keep].copy(), kst=kst[keep].copy())
    return dict(ist=np.zeros(0, np.int32), jst=np.zeros(0, np.int32), kst=np.zeros(0, np.int32))


def _rmedia_args(spec: ModelSpec):
    """arguments of gpu.rmedia_3d for spec.rmedia: the (k,i,j) halo grid, z fastest"""
    r = spec.rmedia
    ax, ay, az = r["a"]
    return (r["ptype"], (spec.nz + 6, spec.nx + 6, spec.ny + 6), (spec.dz, spec.dx, spec.dy), (az, ax, ay),
            r["epsil"]), dict(kappa=r.get("kappa", 0.5), seed=r["seed"])


def _perturbation(spec: ModelSpec):
    """the relative perturbation of spec as a host array, or None"""
    if spec.xi is not None and spec.rmedia is not None:
        raise ValueError("ModelSpec: give xi or rmedia, not both")
    if spec.rmedia is not None:
        from .gpu import rmedia_3d
        a, kw = _rmedia_args(spec)
        return rmedia_3d(*a, **kw)
    return spec.xi


def _perturb_clip(spec: ModelSpec):
    """vmin, vmax and rhomin of the random-medium perturbation: the velocity
    cut-off, the velocity the 4th-order stability condition allows at dt, 1"""
    f32 = np.float32
    dx32, dy32, dz32 = f32(spec.dx), f32(spec.dy), f32(spec.dz)
    vmax_cfl = (f32(6) / f32(7)) * f32(1) / np.sqrt(1 / dx32**2 + 1 / dy32**2 + 1 / dz32**2) / f32(spec.dt)
    return f32(spec.vcut), f32(vmax_cfl), f32(1.0)


def build_inputs(spec: ModelSpec) -> Inputs:
    L = host_lib()
    nx, ny, nz, na, nm = spec.nx, spec.ny, spec.nz, spec.na, spec.nm
    shp = (ny + 6, nx + 6, nz + 6)
    rho, lam, mu, taup, taus = (np.zeros(shp, np.float32) for _ in range(5))
    layers = np.ascontiguousarray(np.asarray(spec.layers, np.float32).reshape(-1))
    rc = C.c_int()
    kind = {"lhm": 0, "lgm": 1}[spec.vmodel]
    L.swpch_vmodel_layers(kind, nx, ny, nz, spec.zbeg, spec.dz, spec.vcut, int(spec.munk), len(spec.layers), layers,
                          rho, lam, mu, taup, taus, C.byref(rc))
    if rc.value != 0:
        raise ValueError(f"velocity model rejected (rc={rc.value})")
    xi = _perturbation(spec)
    if xi is not None:
        L.swpch_perturb(nx, ny, nz, np.ascontiguousarray(xi, np.float32), *_perturb_clip(spec), rho, lam, mu)
        del xi
    ts = np.zeros(max(nm, 1), np.float32)
    zeta = C.c_float()
    kfs, kob, kfs_top, kfs_bot, kob_top, kob_bot = (np.zeros((ny + 6, nx + 6), np.int32) for _ in range(6))
    vmin, vmax = C.c_float(), C.c_float()
    L.swpch_medium_setup(nx, ny, nz, na, nm, spec.fq_min, spec.fq_max, spec.fq_ref, rho, lam, mu, taup, taus, ts,
                         C.byref(zeta), kfs, kob, kfs_top, kfs_bot, kob_top, kob_bot, C.byref(vmin), C.byref(vmax))
    medium = dict(rho=rho, lam=lam, mu=mu, taup=taup, taus=taus, kfs_top=kfs_top, kfs_bot=kfs_bot,
                  kob_top=kob_top, kob_bot=kob_bot)
    return _complete_inputs(spec, medium, kfs, kob, ts, float(zeta.value), float(vmin.value), float(vmax.value))


def _complete_inputs(spec: ModelSpec, medium: dict, kfs, kob, ts, zeta: float, vmin: float, vmax: float) -> Inputs:
    """the inputs after the medium: kernel coefficients (kernel__setup), sources,
    absorber profiles, stations and Green's-function points"""
    L = host_lib()
    nx, ny, nz, na, nm = spec.nx, spec.ny, spec.nz, spec.na, spec.nm
    f32 = np.float32
    c1, c2, d1 = (np.zeros(max(nm, 1), np.float32) for _ in range(3))
    d2 = C.c_float()
    L.swpch_kernel_coef(nm, ts, spec.dt, c1, c2, d1, C.byref(d2))

    srcs = [] if spec.green is not None else spec.sources   # green_mode has none (m_source.f90:84-90)
    sa, ssc = _source_inputs(spec, srcs)
    fcut, M0, UC, nsrc = ssc["fcut"], ssc["M0"], ssc["UC"], ssc["nsrc"]
    dx32, dy32, dz32 = f32(spec.dx), f32(spec.dy), f32(spec.dz)

    # absorbers
    arrays = dict(**medium, kfs=kfs, kob=kob, ts=ts[:nm].copy(), c1=c1, c2=c2, d1=d1, **sa)
    if spec.abc == "pml":
        for ax, n, beg, h in (("x", nx, spec.xbeg, dx32), ("y", ny, spec.ybeg, dy32), ("z", nz, spec.zbeg, dz32)):
            gc = np.zeros(4 * n, np.float32)
            ge = np.zeros(4 * n, np.float32)
            L.swpch_pml_profile(n, na, beg, h, fcut, spec.dt, gc, ge)
            arrays[f"g{ax}c"] = gc
            arrays[f"g{ax}e"] = ge
    else:
        for ax, n, h, top in (("x", nx, dx32, 1), ("y", ny, dy32, 1), ("z", nz, dz32, 0)):
            g_c = np.zeros(n + 6, np.float32)
            g_b = np.zeros(n + 6, np.float32)
            L.swpch_cerjan_profile(n, na, h, top, g_c, g_b)
            arrays[f"cg{ax}_c"] = g_c
            arrays[f"cg{ax}_b"] = g_b

    arrays.update(_station_inputs(spec, kfs, kob))
    ntw = int(math.floor(float(f32(spec.nt - 1) / f32(spec.ntdec_w) + f32(1.0))))
    scalars = dict(zeta=zeta, d2=float(d2.value), vmin=vmin, vmax=vmax,
                   fcut=float(fcut), M0=float(M0), UC=float(UC), ntw=ntw, nsrc=nsrc)
    if spec.green is not None:
        _green_inputs(spec, arrays, scalars, kob)
    return Inputs(spec=spec, arrays=arrays, scalars=scalars)


def _green_inputs(spec: ModelSpec, arrays: dict, scalars: dict, kob: np.ndarray) -> None:
    """green__setup (m_green.f90:66-404): M0 = 1, no regular source (the PML
    cut-off frequency stays 0), no station output, the virtual source at the
    station's cell, GF cells filtered by distance and solid part."""
    L = host_lib()
    gs = spec.green
    f32 = np.float32
    if gs.cmp not in ("x", "y", "z"):
        raise ValueError("no matching green_cmp")
    st = spec.stations[gs.station]
    xs = np.array([st[0]], f32); ys = np.array([st[1]], f32); zs = np.array([st[2]], f32)
    zsw = np.array([ZSW[st[3]]], np.int32)
    ist, jst, kst, inside = (np.zeros(1, np.int32) for _ in range(4))
    dx32, dy32, dz32 = f32(spec.dx), f32(spec.dy), f32(spec.dz)
    L.swpch_station_index(1, xs, ys, zs, zsw, spec.nx, spec.ny, spec.nz, spec.xbeg, spec.ybeg, spec.zbeg,
                          dx32, dy32, dz32, arrays["kfs"], kob, ist, jst, kst, inside)
    if not inside[0]:
        raise ValueError("green_stnm: the virtual source station is outside the domain")
    pts = np.ascontiguousarray(np.asarray(gs.points, f32).reshape(-1, 3))
    ng0 = pts.shape[0]
    ig, jg, kg, keep = (np.zeros(max(ng0, 1), np.int32) for _ in range(4))
    if ng0:
        L.swpch_green_points(ng0, np.ascontiguousarray(pts[:, 0]), np.ascontiguousarray(pts[:, 1]),
                             np.ascontiguousarray(pts[:, 2]), xs[0], ys[0], f32(gs.maxdist), spec.nx, spec.ny,
                             spec.nz, spec.xbeg, spec.ybeg, spec.zbeg, dx32, dy32, dz32, kob, ig, jg, kg, keep)
    sel = keep[:ng0].astype(bool)
    gid = np.arange(ng0) if gs.gid is None else np.asarray(gs.gid)
    stf = np.zeros(spec.nt, f32)
    L.swpch_stf_table(spec.nt, 1, STF_CODES[spec.stftype], np.zeros(1, f32), np.array([gs.trise], f32),
                      spec.tbeg, spec.dt, 0.0, stf)
    arrays.update(green_ig=ig[:ng0][sel].copy(), green_jg=jg[:ng0][sel].copy(), green_kg=kg[:ng0][sel].copy(),
                  green_gid=gid[sel].copy(), green_stf=stf)
    scalars.update(M0=1.0, green_src=(int(ist[0]), int(jst[0]), int(kst[0])))


def make_simulation(inp: Inputs, device: int = 0):
    """Create a SwpcGPU context loaded with the inputs of ``build_inputs``."""
    from .gpu import SwpcGPU
    s, a = inp.spec, inp.arrays
    g = SwpcGPU(s.nx, s.ny, s.nz, s.na, s.nm, s.abc, s.dx, s.dy, s.dz, s.dt, device=device)
    g.set_medium(a["rho"], a["lam"], a["mu"], a["taup"], a["taus"], a["kfs_top"], a["kfs_bot"], a["kob_top"],
                 a["kob_bot"], a["kob"])
    return _load_rest(g, inp)


def make_simulation_device(spec: ModelSpec, device: int = 0):
    """build_inputs and make_simulation with the medium built on the device
    (swpc_medium_layered): the host library gives the per-plane profile of
    the layered model and the scalars of the tau method, and the 3D medium
    arrays never exist on the host.  A spec.rmedia field is generated into
    the context's staging buffer on the device; a spec.xi array is uploaded
    through it.  Returns (simulation, inputs); inputs.arrays holds everything
    build_inputs gives but the 3D medium and the band maps."""
    from .gpu import SwpcGPU
    L = host_lib()
    nx, ny, nz, nm = spec.nx, spec.ny, spec.nz, spec.nm
    prof = np.zeros((5, nz + 6), np.float32)
    layers = np.ascontiguousarray(np.asarray(spec.layers, np.float32).reshape(-1))
    rc = C.c_int()
    kind = {"lhm": 0, "lgm": 1}[spec.vmodel]
    L.swpch_vmodel_profile(kind, nz, spec.zbeg, spec.dz, spec.vcut, int(spec.munk), len(spec.layers), layers, prof,
                           C.byref(rc))
    if rc.value != 0:
        raise ValueError(f"velocity model rejected (rc={rc.value})")
    ts = np.zeros(max(nm, 1), np.float32)
    zeta = C.c_float()
    ccv = np.zeros(2, np.float32)
    L.swpch_visco_setup(nm, spec.fq_min, spec.fq_max, spec.fq_ref, ts, C.byref(zeta), ccv)
    if spec.xi is not None and spec.rmedia is not None:
        raise ValueError("ModelSpec: give xi or rmedia, not both")
    g = SwpcGPU(nx, ny, nz, spec.na, nm, spec.abc, spec.dx, spec.dy, spec.dz, spec.dt, device=device)
    try:
        xi = spec.xi
        if spec.rmedia is not None:
            from .gpu import rmedia_3d
            ptr, n = g.staging_buffer()
            a, kw = _rmedia_args(spec)
            rmedia_3d(*a, **kw, device=device, out_ptr=ptr)
            xi = ptr
        kfs, kob, vr = g.medium_layered(prof, xi, _perturb_clip(spec), (zeta.value, ccv[0], ccv[1]))
        inp = _complete_inputs(spec, {}, kfs, kob, ts, float(zeta.value), float(vr[0]), float(vr[1]))
        _load_rest(g, inp)
    except Exception:
        g.close()
        raise
    return g, inp


def _load_rest(g, inp: Inputs):
    """everything but the medium into a context"""
    s, a, sc = inp.spec, inp.arrays, inp.scalars
    g.set_visco(a["c1"], a["c2"], a["d1"], sc["d2"])
    if s.abc == "pml":
        g.set_pml(a["gxc"], a["gxe"], a["gyc"], a["gye"], a["gzc"], a["gze"])
    else:
        g.set_cerjan(a["cgx_c"], a["cgx_b"], a["cgy_c"], a["cgy_b"], a["cgz_c"], a["cgz_b"])
    if sc["nsrc"]:
        g.set_sources(a["isrc"], a["jsrc"], a["ksrc"], a["mo"], a["comp"], sc["M0"], a["stf"], bf_mode=s.bf_mode)
    if a["ist"].size and s.green is None:
        g.set_receivers(a["ist"], a["jst"], a["kst"], s.wav_what, s.ntdec_w, sc["ntw"], sc["M0"], sc["UC"])
    if s.green is not None:
        g.set_green(a["green_ig"], a["green_jg"], a["green_kg"], *sc["green_src"], s.green.cmp, s.green.bforce,
                    s.ntdec_w, sc["ntw"], a["green_stf"])
    return g


def set_acquisition(sim, inp: Inputs, spec: ModelSpec) -> dict:
    """Switch a simulation made by make_simulation(inp) to the sources and
    stations of spec -- same grid, medium and time step; the sources must
    give the same PML cut-off frequency, which shaped the absorber profiles.
    Returns the arrays (ist, jst, kst, ...) and scalars (M0, UC, ntw) used."""
    s0 = inp.spec
    for k in ("nx", "ny", "nz", "dx", "dy", "dz", "dt", "na", "nm", "xbeg", "ybeg", "zbeg", "abc", "bf_mode"):
        if getattr(spec, k) != getattr(s0, k):
            raise ValueError(f"set_acquisition: {k} differs from the simulation's")
    sa, ssc = _source_inputs(spec, spec.sources)
    if np.float32(ssc["fcut"]) != np.float32(inp.scalars["fcut"]):
        raise ValueError("set_acquisition: the sources change the PML cut-off frequency")
    st = _station_inputs(spec, inp.arrays["kfs"], inp.arrays["kob"])
    ntw = int(math.floor(float(np.float32(spec.nt - 1) / np.float32(spec.ntdec_w) + np.float32(1.0))))
    sim.set_sources(sa["isrc"], sa["jsrc"], sa["ksrc"], sa["mo"], sa["comp"], ssc["M0"], sa["stf"],
                    bf_mode=spec.bf_mode)
    sim.set_receivers(st["ist"], st["jst"], st["kst"], spec.wav_what, spec.ntdec_w, ntw, ssc["M0"], ssc["UC"])
    return dict(arrays={**sa, **st}, scalars=dict(M0=float(ssc["M0"]), UC=float(ssc["UC"]), ntw=ntw))
