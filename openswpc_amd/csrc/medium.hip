// medium.hip -- medium__setup of a layered model on the device
// (swpc_medium_layered).
//
// The reference builds the medium on the host: the layered velocity model
// assigns whole planes (vmodel_lhm / vmodel_lgm, src/swpc_3d/m_vmodel_lhm.f90:
// 20-158, m_vmodel_lgm.f90:20-174), a random medium perturbs every solid cell
// (m_vmodel_lgm_rmed.f90:195-203, vcheck of src/shared/m_fdtool.f90:15-48),
// and medium__setup (src/swpc_3d/m_medium.f90:124-394) copies the nearest
// interior cell into the absorber, turns Q into tau-method strengths, relaxes
// the moduli at the reference frequency and detects the free surface and the
// ocean bottom.  Here the per-plane profile (5 x (nz+6) values) and the
// scalars of the tau method (zeta and the complex mean cc) come from the host;
// one thread per (i,j) column marches k and writes rho, lam, mu, taup, taus
// in the device layout, and column kernels find the surfaces, their
// 2nd-order bands and the velocity extremes.  Every operation keeps the host
// library's single-precision order; the complex root and quotient of the
// relaxed moduli are the forms the host runtime evaluates
// (sqrt(z) = (r, b / (2r)) with r = sqrt((|z| + a) / 2) for a > 0, and
// 1 / w = conj(w) / |w|^2), so the device medium equals the host library's
// bit for bit.
#include <cstdio>
#include <cstring>

#include "swpc_internal.h"

namespace {

#define EPSL_SP 1.1920928955078125e-07f  // epsilon(1.0)

// correctly rounded single-precision root (the host's sqrt of a real(SP))
__device__ static inline float sqrt_rn(float x) { return (float)sqrt((double)x); }

struct MedArgs {
    Grid g;
    const float *prof;   // (5, nz+6): rho, lam, mu, qp, qs at k = -2..nz+3
    const float *xi;     // (k,i,j) halo array or null
    float vclip_min, vclip_max, rhomin;
    int nm;
    float zeta, ccr, cci;
    float *rho, *lam, *mu, *taup, *taus;
};

// relaxed modulus factor chi = 1 / Re(1 / sqrt(1 - cc tau)) in complex
// single precision (m_medium.f90:231-271)
__device__ static inline float chi_of(float ccr, float cci, float tau)
{
    const float zr = 1.0f - ccr * tau;
    const float zi = 0.0f - cci * tau;
    const float d = (float)sqrt((double)zr * (double)zr + (double)zi * (double)zi);   // |z|, correctly rounded
    float wr, wi;
    if (zr > 0.0f) {
        wr = sqrt_rn(0.5f * (d + zr));
        wi = 0.5f * (zi / wr);
    } else {
        const float s = sqrt_rn(0.5f * (d - zr));
        wr = fabsf(0.5f * (zi / s));
        wi = copysignf(s, zi);
    }
    const float den = wr * wr + wi * wi;
    const float qr = wr / den;
    return 1.0f / qr;
}

// One thread per (i,j) column of the halo grid, marching k = -2..nz+3.
// Absorber cells take the values of the nearest interior cell (x, then y,
// then the bottom band), so a cell's values are those of the clamped cell.
__global__ void __launch_bounds__(256) k_medium_cells(MedArgs a)
{
    const Grid &g = a.g;
    const int i = -2 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int j = -2 + (int)blockIdx.y;
    if (i > g.nx + 3 || j > g.ny + 3) return;
    const int na = g.na, nx = g.nx, ny = g.ny, nz = g.nz;
    const int ci = i <= na ? na + 1 : (i > nx - na ? nx - na : i);
    const int cj = j <= na ? na + 1 : (j > ny - na ? ny - na : j);
    const int nzr = nz + 6;
    const long long xcol = ((long long)(cj + 2) * (nx + 6) + (ci + 2)) * nzr;
    const float fnm = (float)a.nm;
    for (int k = -2; k <= nz + 3; ++k) {
        const int ck = k > nz - na ? nz - na : k;
        float r = a.prof[ck + 2];
        float la = a.prof[nzr + ck + 2];
        float m = a.prof[2 * nzr + ck + 2];
        const float qp = a.prof[3 * nzr + ck + 2];
        const float qs = a.prof[4 * nzr + ck + 2];
        // random-medium perturbation of a solid cell (swpch_perturb)
        if (a.xi && m > EPSL_SP) {
            float vp = sqrt_rn((la + 2.0f * m) / r);
            float vs = sqrt_rn(m / r);
            const float e = a.xi[xcol + ck + 2];
            r = r * (1.0f + 0.8f * e);
            vp = vp * (1.0f + e);
            vs = vs * (1.0f + e);
            const float gamma = vp / vs;
            if (vp > a.vclip_max || vs > a.vclip_max) {
                const float xi2 = (1.0f + e) * a.vclip_max / vp - 1.0f;
                vp = a.vclip_max;
                vs = a.vclip_max / gamma;
                r = r * (1.0f + 0.8f * xi2) / (1.0f + 0.8f * e);
            }
            if (vp < a.vclip_min || vs < a.vclip_min) {
                vs = a.vclip_min;
                vp = a.vclip_min * gamma;
            }
            if (r < a.rhomin) r = a.rhomin;
            m = r * vs * vs;
            la = r * (vp * vp - 2.0f * vs * vs);
        }
        // tau method: taup = nm zeta / Qp, taus = nm zeta / Qs
        const float tp = fnm * a.zeta / qp;
        const float ts = fnm * a.zeta / qs;
        if (a.nm > 0) {
            const float rb2 = m;
            const float ra2 = la + 2.0f * m;
            const float chi_mu = chi_of(a.ccr, a.cci, ts);
            const float chi_lam = chi_of(a.ccr, a.cci, tp);
            m = rb2 / (chi_mu * chi_mu);
            la = ra2 / (chi_lam * chi_lam) - 2.0f * m;
        }
        const long long p = gidx(g, i, j, k);
        a.rho[p] = r;
        a.lam[p] = la;
        a.mu[p] = m;
        a.taup[p] = tp;
        a.taus[p] = ts;
    }
}

// free surface and ocean bottom of a column (m_medium.f90:339-394): the
// last k in 1..nz-1 with |lam| (|mu|) below epsilon over one above it, for
// columns 0..n+2; 0 elsewhere
__global__ void __launch_bounds__(256) k_medium_surface(Grid g, const float *lam, const float *mu, int *kfs, int *kob)
{
    const int i = -2 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int j = -2 + (int)blockIdx.y;
    if (i > g.nx + 3 || j > g.ny + 3) return;
    int fs = 0, ob = 0;
    if (i >= 0 && i <= g.nx + 2 && j >= 0 && j <= g.ny + 2) {
        float l0 = lam[gidx(g, i, j, 1)], m0 = mu[gidx(g, i, j, 1)];
        for (int k = 1; k <= g.nz - 1; ++k) {
            const float l1 = lam[gidx(g, i, j, k + 1)], m1 = mu[gidx(g, i, j, k + 1)];
            if (fabsf(m0) < EPSL_SP && fabsf(m1) > EPSL_SP) ob = k;
            if (fabsf(l0) < EPSL_SP && fabsf(l1) > EPSL_SP) fs = k;
            l0 = l1;
            m0 = m1;
        }
    }
    kfs[g2idx(g, i, j)] = fs;
    kob[g2idx(g, i, j)] = ob;
}

// 2nd-order band limits over the 6 x 6 window (i-2..i+3, j-2..j+3) for
// cells 1..n, 0 elsewhere
__global__ void __launch_bounds__(256) k_medium_bands(Grid g, const int *kfs, const int *kob, int *kfs_top,
                                                      int *kfs_bot, int *kob_top, int *kob_bot)
{
    const int i = -2 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int j = -2 + (int)blockIdx.y;
    if (i > g.nx + 3 || j > g.ny + 3) return;
    int ft = 0, fb = 0, ot = 0, ob = 0;
    if (i >= 1 && i <= g.nx && j >= 1 && j <= g.ny) {
        int fmin = 1 << 30, fmax = -(1 << 30), omin = 1 << 30, omax = -(1 << 30);
        for (int jj = j - 2; jj <= j + 3; ++jj)
            for (int ii = i - 2; ii <= i + 3; ++ii) {
                const int f = kfs[g2idx(g, ii, jj)], o = kob[g2idx(g, ii, jj)];
                fmin = min(fmin, f); fmax = max(fmax, f);
                omin = min(omin, o); omax = max(omax, o);
            }
        ft = max(fmin - 2, 1);
        fb = min(fmax + 2, g.nz);
        ot = max(omin - 2, 1);
        ob = min(omax + 2, g.nz);
    }
    const long long q = g2idx(g, i, j);
    kfs_top[q] = ft;
    kfs_bot[q] = fb;
    kob_top[q] = ot;
    kob_bot[q] = ob;
}

// velocity extremes below the free surface (m_medium.f90:396-427): max vp
// and min vs over vs >= epsilon, k = kfs+1..nz of columns 1..n.  The values
// are non-negative, so their bit patterns order like the values and integer
// atomics are exact.  ext[0] = max vp, ext[1] = min vs, ext[2] = 1 if any cell.
__global__ void __launch_bounds__(256) k_medium_vrange(Grid g, const float *rho, const float *lam, const float *mu,
                                                       const int *kfs, unsigned int *ext)
{
    const int i = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int j = 1 + (int)blockIdx.y;
    float vmax = 0.0f, vmin = 1e30f;
    unsigned any = 0;
    if (i <= g.nx && j <= g.ny) {
        for (int k = kfs[g2idx(g, i, j)] + 1; k <= g.nz; ++k) {
            const long long p = gidx(g, i, j, k);
            const float r = rho[p], l = lam[p], m = mu[p];
            const float vp = sqrt_rn((l + 2.0f * m) / r);
            const float vs = sqrt_rn(m / r);
            vmax = fmaxf(vmax, vp);
            any = 1;
            if (vs < EPSL_SP) continue;
            vmin = fminf(vmin, vs);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        vmax = fmaxf(vmax, __shfl_xor(vmax, o));
        vmin = fminf(vmin, __shfl_xor(vmin, o));
        any |= (unsigned)__shfl_xor((int)any, o);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&ext[0], __float_as_uint(vmax));
        atomicMin(&ext[1], __float_as_uint(vmin));
        atomicOr(&ext[2], any);
    }
}

// the two surface maps in the host layout, i fastest over -2..n+3, rows j:
// one contiguous block the host copies back in a single transfer
__global__ void __launch_bounds__(256) k_medium_maps_out(Grid g, const int *kfs, const int *kob, int *out)
{
    const int i = -2 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int j = -2 + (int)blockIdx.y;
    if (i > g.nx + 3 || j > g.ny + 3) return;
    const long long w = g.nx + 6, q = (long long)(j + 2) * w + (i + 2);
    out[q] = kfs[g2idx(g, i, j)];
    out[(long long)(g.ny + 6) * w + q] = kob[g2idx(g, i, j)];
}

__global__ void k_medium_ext_init(unsigned int *ext)
{
    if (threadIdx.x == 0) {
        ext[0] = 0u;
        ext[1] = __float_as_uint(1e30f);
        ext[2] = 0u;
    }
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

}  // namespace

extern "C" int swpc_medium_layered(swpc_ctx *c, const float *prof, const float *xi, const float clip[3],
                                   const float tau[3], int32_t *kfs_out, int32_t *kob_out, float vrange[2])
{
    if (!c) return SWPC_EINVAL;
    if (!prof || !clip || !tau || !kfs_out || !kob_out || !vrange)
        return ctx_fail(c, SWPC_EINVAL, "swpc_medium_layered: null argument");
    const Grid &g = c->g;
    if (g.na < 0 || g.nx - 2 * g.na < 1 || g.ny - 2 * g.na < 1 || g.nz - g.na < 1)
        return ctx_fail(c, SWPC_EINVAL, "swpc_medium_layered: the absorber leaves no interior cell");
    if (hipSetDevice(c->device) != hipSuccess) return ctx_fail(c, SWPC_EHIP, "swpc_medium_layered: hipSetDevice");
    hipStream_t s = c->stream;
    const int nzr = g.nz + 6;
    DevBuf dprof, dmaps, dext, dout;
    const size_t nmap = (size_t)(g.nx + 6) * (g.ny + 6);
    hipError_t e = hipMalloc(&dprof.p, (size_t)5 * nzr * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&dmaps.p, (size_t)2 * g.plane * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&dout.p, 2 * nmap * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&dext.p, 4 * sizeof(unsigned int));
    if (e != hipSuccess) return ctx_fail(c, SWPC_ENOMEM, "swpc_medium_layered: device memory for the set-up");
    e = hipMemcpyAsync(dprof.p, prof, (size_t)5 * nzr * sizeof(float), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemsetAsync(dmaps.p, 0, (size_t)2 * g.plane * sizeof(int), s);
    // the perturbation: device memory is read in place, host memory goes through the staging array
    const float *dxi = nullptr;
    if (e == hipSuccess && xi) {
        hipPointerAttribute_t at;
        const bool on_dev = hipPointerGetAttributes(&at, xi) == hipSuccess && at.type == hipMemoryTypeDevice &&
                            at.device == c->device;
        (void)hipGetLastError();
        if (on_dev) {
            dxi = xi;
        } else {
            e = hipMemcpyAsync(c->staging, xi, (size_t)c->staging_n * sizeof(float), hipMemcpyHostToDevice, s);
            dxi = c->staging;
        }
    }
    if (e != hipSuccess) return ctx_fail(c, SWPC_EHIP, "swpc_medium_layered: upload");
    MedArgs a;
    a.g = g;
    a.prof = (const float *)dprof.p;
    a.xi = dxi;
    a.vclip_min = clip[0];
    a.vclip_max = clip[1];
    a.rhomin = clip[2];
    a.nm = c->desc.nm;
    a.zeta = tau[0];
    a.ccr = tau[1];
    a.cci = tau[2];
    a.rho = c->rho; a.lam = c->lam; a.mu = c->mu; a.taup = c->taup; a.taus = c->taus;
    const dim3 cols((g.nx + 6 + 63) / 64, g.ny + 6);
    hipLaunchKernelGGL(k_medium_cells, cols, dim3(64), 0, s, a);
    int *kfs = (int *)dmaps.p, *kob_tmp = kfs + g.plane;
    hipLaunchKernelGGL(k_medium_surface, cols, dim3(64), 0, s, g, c->lam, c->mu, kfs, kob_tmp);
    hipLaunchKernelGGL(k_medium_bands, cols, dim3(64), 0, s, g, kfs, kob_tmp, c->kfs_top, c->kfs_bot, c->kob_top,
                       c->kob_bot);
    hipLaunchKernelGGL(k_medium_ext_init, dim3(1), dim3(64), 0, s, (unsigned int *)dext.p);
    hipLaunchKernelGGL(k_medium_vrange, dim3((g.nx + 63) / 64, g.ny), dim3(64), 0, s, g, c->rho, c->lam, c->mu, kfs,
                       (unsigned int *)dext.p);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(c->kob, kob_tmp, (size_t)g.plane * sizeof(int), hipMemcpyDeviceToDevice, s);
    // the maps back in the host layout (i fastest over -2..n+3, rows j)
    hipLaunchKernelGGL(k_medium_maps_out, cols, dim3(64), 0, s, g, kfs, kob_tmp, (int *)dout.p);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(kfs_out, dout.p, nmap * sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(kob_out, (int *)dout.p + nmap, nmap * sizeof(int), hipMemcpyDeviceToHost, s);
    unsigned int ext[4] = {0, 0, 0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(ext, dext.p, 3 * sizeof(unsigned int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        char msg[160];
        snprintf(msg, sizeof msg, "swpc_medium_layered: %s", hipGetErrorString(e));
        return ctx_fail(c, SWPC_EHIP, msg);
    }
    float vmax, vmin;
    memcpy(&vmax, &ext[0], 4);
    memcpy(&vmin, &ext[1], 4);
    vrange[0] = vmin;
    vrange[1] = ext[2] ? vmax : -1.0f;
    c->has_medium = true;
    ctx_reset_graph(c);
    return SWPC_OK;
}
