"""ctypes binding of libswpc_gpu (include/swpc_gpu.h).

``SwpcGPU`` mirrors the reference's hot-path routines one for one
(src/swpc_3d/main.f90:119-139):

    kernel__update_stress  -> SwpcGPU.kernel_update_stress()
    absorb__update_stress  -> SwpcGPU.absorb_update_stress()
    source__stressglut(it) -> SwpcGPU.source_stressglut(it)
    kernel__update_vel     -> SwpcGPU.kernel_update_vel()
    absorb__update_vel     -> SwpcGPU.absorb_update_vel()
    source__bodyforce(it)  -> SwpcGPU.source_bodyforce(it)
    wav__store(it)         -> SwpcGPU.wav_store(it)
    kernel__vmax           -> SwpcGPU.kernel_vmax()

Arrays cross the boundary in the reference layout: 3D (k,i,j) with k
fastest and a 3-cell halo, i.e. numpy shape (ny+6, nx+6, nz+6); (i,j) maps
as (ny+6, nx+6).  There is no CPU fallback: if the HIP library is missing
or no GPU is present, construction raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libswpc_gpu.so")

FIELDS = {"Vx": 0, "Vy": 1, "Vz": 2, "Sxx": 3, "Syy": 4, "Szz": 5, "Syz": 6, "Sxz": 7, "Sxy": 8}
WAV = {"vel": 1, "disp": 2, "stress": 4, "strain": 8}
NCMP = {"vel": 3, "disp": 3, "stress": 6, "strain": 6}
ABC = {"pml": 0, "cerjan": 1}
TIMERS = ("wav", "snap", "stress_side", "stress", "stress_abc", "glut", "vel_side", "vel", "vel_abc", "bodyforce")
SNAP_PLANES = ("xy", "yz", "xz", "fs", "ob")
SNAP_QS = ("ps", "v", "u")


def snap_section(name: str) -> int:
    """section id of 'xz_ps', 'fs_v', ... (plane * 3 + quantity)"""
    pl, q = name.split("_")
    return SNAP_PLANES.index(pl) * 3 + SNAP_QS.index(q)


class _SnapDesc(C.Structure):
    _fields_ = [("sw", C.c_int * 15), ("k0_xy", C.c_int), ("i0_yz", C.c_int), ("j0_xz", C.c_int),
                ("idec", C.c_int), ("jdec", C.c_int), ("kdec", C.c_int), ("ntdec_s", C.c_int),
                ("nframe", C.c_int), ("M0", C.c_float), ("UC", C.c_float)]


class SwpcError(RuntimeError):
    pass


class _Desc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("na", C.c_int), ("nm", C.c_int),
                ("abc", C.c_int), ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("dt", C.c_float), ("device", C.c_int)]


_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load libswpc_gpu.so; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise SwpcError(f"{path} is missing: build it with openswpc_amd/build.py (no CPU fallback exists)")
    L = C.CDLL(path)
    vp, i, f = C.c_void_p, C.c_int, C.c_float
    fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    L.swpc_create.argtypes = [C.POINTER(_Desc), C.POINTER(vp)]
    L.swpc_destroy.argtypes = [vp]
    L.swpc_last_error.argtypes = [vp]
    L.swpc_last_error.restype = C.c_char_p
    L.swpc_set_medium.argtypes = [vp, fp, fp, fp, fp, fp, ip, ip, ip, ip, ip]
    L.swpc_set_visco.argtypes = [vp, fp, fp, fp, f]
    L.swpc_set_pml.argtypes = [vp, fp, fp, fp, fp, fp, fp]
    L.swpc_set_cerjan.argtypes = [vp, fp, fp, fp, fp, fp, fp]
    L.swpc_set_sources.argtypes = [vp, i, i, ip, ip, ip, fp, fp, f, i, fp]
    L.swpc_set_receivers.argtypes = [vp, i, ip, ip, ip, i, i, i, f, f]
    for nm in ("swpc_kernel_update_stress", "swpc_absorb_update_stress", "swpc_kernel_update_vel",
               "swpc_absorb_update_vel", "swpc_update_stress", "swpc_update_vel", "swpc_synchronize"):
        getattr(L, nm).argtypes = [vp]
    L.swpc_set_green.argtypes = [vp, i, ip, ip, ip, i, i, i, i, i, i, i, i, fp]
    L.swpc_get_green.argtypes = [vp, fp]
    for nm in ("swpc_source_stressglut", "swpc_source_bodyforce", "swpc_wav_store", "swpc_green_store",
               "swpc_green_source"):
        getattr(L, nm).argtypes = [vp, i]
    L.swpc_step.argtypes = [vp, i, i, i, i]
    L.swpc_step_timed.argtypes = [vp, i, i, fp]
    L.swpc_set_snapshots.argtypes = [vp, C.POINTER(_SnapDesc), ip]
    L.swpc_snap_write.argtypes = [vp, i]
    L.swpc_get_snapshot.argtypes = [vp, i, i, fp]
    L.swpc_get_snap_max.argtypes = [vp, i, fp]
    L.swpc_snap_dims.argtypes = [vp, i, ip]
    L.swpc_kernel_vmax.argtypes = [vp, fp]
    L.swpc_kernel_vmax_noalloc.argtypes = [vp, vp]
    L.swpc_medium_layered.argtypes = [vp, fp, vp, fp, fp, ip, ip, fp]
    L.swpc_get_medium.argtypes = [vp, i, vp]
    L.swpc_staging_buffer.argtypes = [vp, C.POINTER(C.c_int64)]
    L.swpc_staging_buffer.restype = vp
    L.swpc_get_field.argtypes = [vp, i, fp]
    L.swpc_set_field.argtypes = [vp, i, fp]
    L.swpc_get_seismograms.argtypes = [vp, i, fp]
    L.swpc_device_bytes.argtypes = [vp]
    L.swpc_device_bytes.restype = C.c_int64
    L.swpc_set_tuning.argtypes = [vp, i, i, i]
    L.swpc_stream.argtypes = [vp]
    L.swpc_stream.restype = vp
    L.swpc_adj_gradient.argtypes = [vp, i, fp, i, C.POINTER(C.c_double), vp, vp, vp]
    L.swpc_adj_plan.argtypes = [vp, i, C.c_int64, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int64),
                                C.POINTER(C.c_int64)]
    L.swpc_adj_reset.argtypes = [vp]
    L.swpc_adj_accumulate.argtypes = [vp, i, fp, i, i, C.POINTER(C.c_double)]
    L.swpc_adj_get_gradient.argtypes = [vp, vp, vp, vp]
    L.swpc_adj_release.argtypes = [vp]
    L.swpc_rmedia_3d.argtypes = [i, i, i, i, i, f, f, f, f, f, f, f, f, C.c_uint64, C.POINTER(C.c_float)]
    L.swpc_rmedia_last_error.argtypes = []
    L.swpc_rmedia_last_error.restype = C.c_char_p
    L.swpc_batch_create.argtypes = [C.POINTER(vp), i, C.POINTER(vp)]
    L.swpc_batch_destroy.argtypes = [vp]
    L.swpc_batch_last_error.argtypes = [vp]
    L.swpc_batch_last_error.restype = C.c_char_p
    L.swpc_batch_step.argtypes = [vp, i, i, i, i]
    L.swpc_batch_synchronize.argtypes = [vp]
    L.swpc_batch_tune.argtypes = [vp, i, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                  C.POINTER(C.c_float)]
    L.swpc_batch_stream.argtypes = [vp]
    L.swpc_batch_set_packing.argtypes = [vp, i]
    L.swpc_batch_get_packing.argtypes = [vp]
    L.swpc_batch_stream.restype = vp
    L.swpc_stripe_reserve.argtypes = [i, C.c_int64, i]
    L.swpc_stripe_release.argtypes = [i]
    L.swpc_stripe_free_bytes.argtypes = [i, i]
    L.swpc_stripe_free_bytes.restype = C.c_int64
    _lib = L
    return L


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class SwpcGPU:
    """One device-resident 3D simulation (one context of libswpc_gpu)."""

    def __init__(self, nx, ny, nz, na, nm, abc="pml", dx=0.5, dy=0.5, dz=0.5, dt=0.01, device=0):
        self.L = load_library()
        d = _Desc(nx=nx, ny=ny, nz=nz, na=na, nm=nm, abc=ABC[abc], dx=dx, dy=dy, dz=dz, dt=dt, device=device)
        h = C.c_void_p()
        rc = self.L.swpc_create(C.byref(d), C.byref(h))
        if rc != 0:
            raise SwpcError(f"swpc_create failed ({rc})")
        self.h = h
        self.nx, self.ny, self.nz, self.na, self.nm = nx, ny, nz, na, nm
        self.abc = abc
        self.nst = 0
        self.ntw = 0

    def _chk(self, rc):
        if rc != 0:
            msg = self.L.swpc_last_error(self.h)
            raise SwpcError(f"libswpc_gpu error {rc}: {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "h", None):
            self.L.swpc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- set-up (the arrays the reference keeps in its modules) ----
    def set_medium(self, rho, lam, mu, taup, taus, kfs_top, kfs_bot, kob_top, kob_bot, kob):
        self._chk(self.L.swpc_set_medium(self.h, _f32(rho), _f32(lam), _f32(mu), _f32(taup), _f32(taus),
                                         _i32(kfs_top), _i32(kfs_bot), _i32(kob_top), _i32(kob_bot), _i32(kob)))

    def set_visco(self, c1, c2, d1, d2):
        self._chk(self.L.swpc_set_visco(self.h, _f32(c1), _f32(c2), _f32(d1), float(d2)))

    def set_pml(self, gxc, gxe, gyc, gye, gzc, gze):
        self._chk(self.L.swpc_set_pml(self.h, _f32(gxc), _f32(gxe), _f32(gyc), _f32(gye), _f32(gzc), _f32(gze)))

    def set_cerjan(self, gx_c, gx_b, gy_c, gy_b, gz_c, gz_b):
        self._chk(self.L.swpc_set_cerjan(self.h, _f32(gx_c), _f32(gx_b), _f32(gy_c), _f32(gy_b), _f32(gz_c),
                                         _f32(gz_b)))

    def set_sources(self, isrc, jsrc, ksrc, mo, comp, M0, stf, bf_mode=False):
        isrc, jsrc, ksrc = _i32(isrc), _i32(jsrc), _i32(ksrc)
        n = int(isrc.size)
        stf = _f32(stf).reshape(-1, max(n, 1)) if n else np.zeros((1, 1), np.float32)
        comp = _f32(comp) if n else np.zeros(6, np.float32)
        mo = _f32(mo) if (n and not bf_mode) else np.ones(max(n, 1), np.float32)
        self._chk(self.L.swpc_set_sources(self.h, n, 1 if bf_mode else 0, isrc if n else np.zeros(1, np.int32),
                                          jsrc if n else np.zeros(1, np.int32), ksrc if n else np.zeros(1, np.int32),
                                          mo, comp, float(M0), int(stf.shape[0]) if n else 0, stf))

    def set_receivers(self, ist, jst, kst, what, ntdec_w, ntw, M0, UC):
        ist, jst, kst = _i32(ist), _i32(jst), _i32(kst)
        self._chk(self.L.swpc_set_receivers(self.h, int(ist.size), ist, jst, kst, int(what), int(ntdec_w), int(ntw),
                                            float(M0), float(UC)))
        self.nst, self.ntw = int(ist.size), int(ntw)

    def set_green(self, ig, jg, kg, isrc, jsrc, ksrc, cmp, bforce, ntdec_w, ntw, stf):
        """Reciprocity Green's-function mode (m_green.f90): GF cells
        (ig, jg, kg), virtual source cell, force direction cmp in 'xyz' (None
        switches the mode off), body-force components, decimation, samples
        and the moment rate stf[it-1] at t = tbeg + it*dt."""
        if cmp is None:
            z = np.zeros(1, np.int32)
            self._chk(self.L.swpc_set_green(self.h, 0, z, z, z, 1, 1, 1, -1, 0, 1, 1, 0, np.zeros(1, np.float32)))
            self.green_shape = None
            return
        ig, jg, kg = _i32(ig), _i32(jg), _i32(kg)
        stf = _f32(stf).reshape(-1)
        ng = int(ig.size)
        pad = np.zeros(1, np.int32)
        self._chk(self.L.swpc_set_green(self.h, ng, ig if ng else pad, jg if ng else pad, kg if ng else pad,
                                        int(isrc), int(jsrc), int(ksrc), "xyz".index(cmp), 1 if bforce else 0,
                                        int(ntdec_w), int(ntw), int(stf.size),
                                        stf if stf.size else np.zeros(1, np.float32)))
        self.green_shape = (ng * (9 if bforce else 6), int(ntw))

    def green(self):
        """gf(ntw, ncmp*ng) as a (ncmp*ng, ntw) array: row n*ncmp + c is
        component c (mxx myy mzz myz mxz mxy [fx fy fz]) of GF cell n, before
        the export's sign flip for cmp = 'z'"""
        out = np.empty(self.green_shape, np.float32)
        if out.size:
            self._chk(self.L.swpc_get_green(self.h, out))
        return out

    def green_store(self, it):
        self._chk(self.L.swpc_green_store(self.h, int(it)))

    def green_source(self, it):
        self._chk(self.L.swpc_green_source(self.h, int(it)))

    # ---- time loop, routine for routine ----
    def kernel_update_stress(self):
        self._chk(self.L.swpc_kernel_update_stress(self.h))

    def absorb_update_stress(self):
        self._chk(self.L.swpc_absorb_update_stress(self.h))

    def update_stress(self):
        """kernel__update_stress + absorb__update_stress in one fused launch"""
        self._chk(self.L.swpc_update_stress(self.h))

    def update_vel(self):
        """kernel__update_vel + absorb__update_vel in one fused launch"""
        self._chk(self.L.swpc_update_vel(self.h))

    def source_stressglut(self, it):
        self._chk(self.L.swpc_source_stressglut(self.h, int(it)))

    def kernel_update_vel(self):
        self._chk(self.L.swpc_kernel_update_vel(self.h))

    def absorb_update_vel(self):
        self._chk(self.L.swpc_absorb_update_vel(self.h))

    def source_bodyforce(self, it):
        self._chk(self.L.swpc_source_bodyforce(self.h, int(it)))

    def wav_store(self, it):
        self._chk(self.L.swpc_wav_store(self.h, int(it)))

    def step(self, it0, nsteps, use_graph=False, block=100):
        self._chk(self.L.swpc_step(self.h, int(it0), int(nsteps), 1 if use_graph else 0, int(block)))

    def step_timed(self, it0, nsteps):
        """direct-launch steps with HIP events between launches; returns the
        milliseconds of each launch slot (TIMERS) summed over the steps"""
        ms = np.zeros(len(TIMERS), np.float32)
        self._chk(self.L.swpc_step_timed(self.h, int(it0), int(nsteps), ms))
        return dict(zip(TIMERS, (float(x) for x in ms)))

    # ---- sections (snap__setup / snap__write) ----
    def set_snapshots(self, sections, kfs, k0_xy=1, i0_yz=1, j0_xz=1, idec=1, jdec=1, kdec=1, ntdec_s=1, nframe=1,
                      M0=1.0, UC=1.0):
        """sections: names like 'xz_ps', 'fs_v', 'ob_u'; kfs: (ny+6, nx+6) free-surface map"""
        d = _SnapDesc()
        for name in sections:
            d.sw[snap_section(name)] = 1
        d.k0_xy, d.i0_yz, d.j0_xz = k0_xy, i0_yz, j0_xz
        d.idec, d.jdec, d.kdec, d.ntdec_s, d.nframe = idec, jdec, kdec, ntdec_s, nframe
        d.M0, d.UC = M0, UC
        self._chk(self.L.swpc_set_snapshots(self.h, C.byref(d), _i32(kfs)))
        self.snap_nframe = nframe

    def snap_write(self, it):
        self._chk(self.L.swpc_snap_write(self.h, int(it)))

    def snap_dims(self, name):
        dims = np.zeros(3, np.int32)
        self._chk(self.L.swpc_snap_dims(self.h, snap_section(name), dims))
        return tuple(int(x) for x in dims)

    def snapshot(self, name, frame):
        """frame of a section as numpy (ncomp, ns2, ns1) = Fortran (ns1, ns2, ncomp)"""
        ns1, ns2, nc = self.snap_dims(name)
        out = np.empty((nc, ns2, ns1), np.float32)
        self._chk(self.L.swpc_get_snapshot(self.h, snap_section(name), int(frame), out))
        return out

    def snap_max(self, name):
        ns1, ns2, _ = self.snap_dims(name)
        out = np.empty((3, ns2, ns1), np.float32)
        self._chk(self.L.swpc_get_snap_max(self.h, snap_section(name), out))
        return out

    def adjoint_gradient(self, nt, obs, ckp=45):
        """misfit 1/2 sum (d - obs)^2 of the velocity traces after nt steps from
        rest and its gradients (lam, mu, rho), reference layout"""
        shp = (self.ny + 6, self.nx + 6, self.nz + 6)
        gl, gm, gr = (np.zeros(shp, np.float32) for _ in range(3))
        J = C.c_double()
        self._chk(self.L.swpc_adj_gradient(self.h, int(nt), _f32(obs), int(ckp), C.byref(J), gl.ctypes.data,
                                           gm.ctypes.data, gr.ctypes.data))
        return J.value, gl, gm, gr

    def adjoint_plan(self, nt, budget=0):
        """(ckp_outer, ckp_inner, device bytes, forward steps) of the cheapest
        checkpoint scheme within budget bytes (0: the free device memory)"""
        c1, c2 = C.c_int(), C.c_int()
        nb, nf = C.c_int64(), C.c_int64()
        self._chk(self.L.swpc_adj_plan(self.h, int(nt), int(budget), C.byref(c1), C.byref(c2), C.byref(nb),
                                       C.byref(nf)))
        return c1.value, c2.value, nb.value, nf.value

    def adjoint_reset(self):
        self._chk(self.L.swpc_adj_reset(self.h))

    def adjoint_accumulate(self, nt, obs, ckp_outer, ckp_inner=0):
        """misfit of the current sources and receivers; the gradient is added
        to the device sums"""
        J = C.c_double()
        self._chk(self.L.swpc_adj_accumulate(self.h, int(nt), _f32(obs), int(ckp_outer), int(ckp_inner),
                                             C.byref(J)))
        return J.value

    def adjoint_sums(self):
        """the device gradient sums (lam, mu, rho), reference layout"""
        shp = (self.ny + 6, self.nx + 6, self.nz + 6)
        gl, gm, gr = (np.zeros(shp, np.float32) for _ in range(3))
        self._chk(self.L.swpc_adj_get_gradient(self.h, gl.ctypes.data, gm.ctypes.data, gr.ctypes.data))
        return gl, gm, gr

    def adjoint_release(self):
        self._chk(self.L.swpc_adj_release(self.h))

    def synchronize(self):
        self._chk(self.L.swpc_synchronize(self.h))

    def kernel_vmax(self):
        out = np.zeros(3, np.float32)
        self._chk(self.L.swpc_kernel_vmax(self.h, out))
        return out

    MEDIUM = ("rho", "lam", "mu", "taup", "taus", "kfs_top", "kfs_bot", "kob_top", "kob_bot", "kob")

    def get_medium(self, name):
        """the device's rho/lam/mu/taup/taus ((k,i,j) halo layout) or band
        and ocean-bottom maps ((i,j) halo layout)"""
        w = self.MEDIUM.index(name)
        if w < 5:
            out = np.empty((self.ny + 6, self.nx + 6, self.nz + 6), np.float32)
        else:
            out = np.empty((self.ny + 6, self.nx + 6), np.int32)
        self._chk(self.L.swpc_get_medium(self.h, w, out.ctypes.data_as(C.c_void_p)))
        return out

    def staging_buffer(self):
        """(device address, floats) of the context's staging buffer, one
        (k,i,j) halo array long"""
        n = C.c_int64()
        ptr = self.L.swpc_staging_buffer(self.h, C.byref(n))
        return int(ptr), int(n.value)

    def medium_layered(self, prof, xi, clip, tau):
        """medium__setup of a layered model on the device: prof (5, nz+6) =
        rho, lam, mu, qp, qs per plane; xi None, a (k,i,j) halo numpy array or
        the device address of one (an int, e.g. the staging buffer's); clip =
        (vmin, vmax, rhomin) of the perturbation; tau =
        (zeta, Re cc, Im cc).  Returns the kfs and kob maps ((ny+6, nx+6),
        the reference's halo layout) and (vmin, vmax)."""
        ny6, nx6, nz6 = self.ny + 6, self.nx + 6, self.nz + 6
        prof = np.ascontiguousarray(prof, np.float32)
        if prof.shape != (5, nz6):
            raise ValueError(f"prof must be (5, {nz6})")
        if xi is None:
            xp = None
        elif isinstance(xi, int):
            xp = C.c_void_p(xi)          # a device address of one (k,i,j) halo array
        else:
            xi = np.ascontiguousarray(xi, np.float32)
            if xi.shape != (ny6, nx6, nz6):
                raise ValueError("xi: the (k,i,j) halo shape")
            xp = xi.ctypes.data_as(C.c_void_p)
        kfs = np.zeros((ny6, nx6), np.int32)
        kob = np.zeros((ny6, nx6), np.int32)
        vr = np.zeros(2, np.float32)
        self._chk(self.L.swpc_medium_layered(self.h, prof, xp, np.asarray(clip, np.float32),
                                             np.asarray(tau, np.float32), kfs, kob, vr))
        return kfs, kob, vr

    def kernel_vmax_noalloc(self, d_out):
        """the three maxima written to d_out (a device address of 3 floats on
        this context's GPU), stream-ordered; nothing is allocated"""
        self._chk(self.L.swpc_kernel_vmax_noalloc(self.h, C.c_void_p(int(d_out))))

    # ---- data out / in ----
    def get_field(self, name):
        out = np.empty((self.ny + 6, self.nx + 6, self.nz + 6), np.float32)
        self._chk(self.L.swpc_get_field(self.h, FIELDS[name], out))
        return out

    def set_field(self, name, arr):
        a = _f32(arr)
        assert a.shape == (self.ny + 6, self.nx + 6, self.nz + 6)
        self._chk(self.L.swpc_set_field(self.h, FIELDS[name], a))

    def seismograms(self, kind="vel"):
        out = np.empty((self.nst, NCMP[kind], self.ntw), np.float32)
        self._chk(self.L.swpc_get_seismograms(self.h, WAV[kind], out))
        return out

    def device_bytes(self):
        return int(self.L.swpc_device_bytes(self.h))

    def set_tuning(self, kchunk=0, by=0, variant=-1):
        self._chk(self.L.swpc_set_tuning(self.h, int(kchunk), int(by), int(variant)))

    def stream(self):
        return self.L.swpc_stream(self.h)


class SwpcBatch:
    """Member simulations stepped together, one launch per phase (swpc_batch_*)."""

    def __init__(self, members):
        self.L = load_library()
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m.h.value for m in self.members])
        h = C.c_void_p()
        rc = self.L.swpc_batch_create(arr, len(self.members), C.byref(h))
        if rc != 0:
            raise SwpcError(f"swpc_batch_create failed ({rc})")
        self.h = h

    def _chk(self, rc):
        if rc != 0:
            msg = self.L.swpc_batch_last_error(self.h)
            raise SwpcError(f"libswpc_gpu batch error {rc}: {msg.decode() if msg else ''}")

    def step(self, it0, nsteps, use_graph=False, block=100):
        self._chk(self.L.swpc_batch_step(self.h, int(it0), int(nsteps), 1 if use_graph else 0, int(block)))

    def synchronize(self):
        self._chk(self.L.swpc_batch_synchronize(self.h))

    def tune(self, trial_steps=3):
        """(k chunk, tile rows, interior tile width, ms per trial step) of the
        fastest shape; the packing chosen is self.packing()"""
        kc, by, w, ms = C.c_int(), C.c_int(), C.c_int(), C.c_float()
        self._chk(self.L.swpc_batch_tune(self.h, int(trial_steps), C.byref(kc), C.byref(by), C.byref(w),
                                         C.byref(ms)))
        return kc.value, by.value, w.value, ms.value

    def set_packing(self, packing):
        self._chk(self.L.swpc_batch_set_packing(self.h, int(packing)))

    def packing(self):
        return int(self.L.swpc_batch_get_packing(self.h))

    def stream(self):
        return self.L.swpc_batch_stream(self.h)

    def close(self):
        if getattr(self, "h", None):
            self.L.swpc_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


RMEDIA_TYPES = {"gauss": 1, "exp": 2, "vonkarman": 3}


def rmedia_3d(ptype, shape, d, a, epsil, kappa=0.5, seed=1, device=0, out_ptr=None):
    """Random medium on the device (m_rmedia.f90 rmedia__3dgen_ani).
    shape = (n1, n2, n3) with axis 1 fastest, d and a per axis; returns a
    float32 numpy array of shape (n3, n2, n1), or, given out_ptr (the device
    address of n1*n2*n3 floats on `device`), writes the field there and
    returns None.  ptype: 1/2/3 or 'gauss'/'exp'/'vonkarman'."""
    L = load_library()
    pt = RMEDIA_TYPES.get(ptype, ptype)
    n1, n2, n3 = (int(v) for v in shape)
    if out_ptr is not None:
        out = None
        ptr = C.cast(C.c_void_p(int(out_ptr)), C.POINTER(C.c_float))
    else:
        out = np.empty((n3, n2, n1), np.float32)
        ptr = out.ctypes.data_as(C.POINTER(C.c_float))
    rc = L.swpc_rmedia_3d(int(device), int(pt), n1, n2, n3, *(float(v) for v in d), *(float(v) for v in a),
                          float(epsil), float(kappa), int(seed) & (2**64 - 1), ptr)
    if rc != 0:
        raise SwpcError(f"swpc_rmedia_3d failed ({rc}): {L.swpc_rmedia_last_error().decode()}")
    return out


# ---- device memory striped over the HBM regions (arena.hip) ----
HBM_REGIONS = 3   # MI355X: 288 GiB in three 96 GiB regions


def stripe_reserve(bytes_per_region, device=0, nregions=HBM_REGIONS):
    """Reserve bytes_per_region of device memory in each HBM region; contexts
    created on the device afterwards take their per-cell arrays from it, each
    array in a fixed region (swpc_stripe_reserve)."""
    L = load_library()
    rc = L.swpc_stripe_reserve(int(device), int(bytes_per_region), int(nregions))
    if rc != 0:
        raise SwpcError(f"swpc_stripe_reserve failed ({rc})")


def stripe_release(device=0):
    """Return the arena's memory; fails while a context still uses it."""
    L = load_library()
    rc = L.swpc_stripe_release(int(device))
    if rc != 0:
        raise SwpcError(f"swpc_stripe_release failed ({rc}): arena memory still in use")


def stripe_free_bytes(device=0, region=0):
    """Bytes left in a region of the device's arena (0 without an arena)."""
    return int(load_library().swpc_stripe_free_bytes(int(device), int(region)))


def stripe_bytes_per_region(nx, ny, nz, nm, na=0, nregions=HBM_REGIONS, margin=1.05, field_sets=1):
    """Arena bytes per region that hold one context's per-cell arrays:
    9 fields and 5 medium arrays of (nx+6+32-aligned) x (ny+6) x (nz+6)
    floats, 6*nm memory variables per interior cell, 18 auxiliaries per PML
    cell; the largest region share, with 256 MiB granules and a 5 % margin
    (add a granule or two of slack for a single context).  field_sets=2
    counts the second set of 9 fields that the fused steps (variant bit 7)
    hold."""
    pitch = -(-(32 + nx + 3) // 32) * 32
    n = pitch * (ny + 6) * (nz + 6) * 4
    rn = -(-nx // 32) * 32 * max(ny - 2 * na, 0) * max(nz - na, 0) * 4
    npml = nx * ny * nz - max(nx - 2 * na, 0) * max(ny - 2 * na, 0) * max(nz - na, 0) if na > 0 else 0
    gran = 256 << 20
    # arrays of a granule or more take whole granules; smaller ones share slabs
    up = lambda b: -(-int(b) // gran) * gran if b >= gran else -(-int(b) // 65536) * 65536  # noqa: E731
    per = [0] * nregions
    # fields and medium: 14 arrays; regions as the library assigns them
    for r in (0, 1, 2, 0, 1, 2, 0, 1, 2, 0, 0, 1, 2, 2) + (0, 1, 2) * 3 * (field_sets - 1):
        per[r % nregions] += up(n)
    npack = (6 * nm + 3) // 4
    if nm > 0:
        for s in range(npack):
            per[s % nregions] += up(rn * (4 if 4 * (s + 1) <= 6 * nm else 2))
    if npml:
        for c in range(18):
            per[c % nregions] += npml * 4
    return int(max(per) * margin)

