// aux.hip -- source injection, receiver sampling, vmax and layout
// transposes of libswpc_gpu.
//
// Source injection is deterministic: the host groups every (field, cell)
// update of every source into a target list; one thread owns one target and
// applies its contributions in source order, exactly the sequence of
// subtractions the reference performs (m_source.f90:798-841), so no atomics
// and no order-dependent rounding.
#include <cstdio>
#include <cstring>
#include <mutex>

#include "swpc_internal.h"

// --------------------------------------------------------------------------
// source__stressglut, m_source.f90:776-848
// --------------------------------------------------------------------------
struct SrcArgs {
    float *fld[SWPC_NFIELD];
    const SrcTarget *tg;
    const SrcContrib *ct;
    const float *comp;   // (ncomp, nsrc)
    const float *mo;     // nsrc
    const float *stf;    // (nt, nsrc)
    const float *rho;
    int ntarget, nsrc, nt;
    float dt_dxyz;
    int it;
    const int *d_it;
};

__device__ static inline int step_of(const SrcArgs &a) { return a.d_it ? *a.d_it : a.it; }

// target t of one context's stress glut at step it
__device__ static inline void stressglut_body(const SrcArgs &a, int t, int it)
{
    if (t >= a.ntarget) return;
    if (it < 1 || it > a.nt) return;
    const SrcTarget tg = a.tg[t];
    float *f = a.fld[tg.field];
    float v = f[tg.idx];
    for (int q = tg.start; q < tg.start + tg.count; ++q) {
        const SrcContrib c = a.ct[q];
        const int n = c.src;
        const float stime = a.stf[(long long)(it - 1) * a.nsrc + n];
        const float sdrop = a.mo[n] * stime * a.dt_dxyz;
        const float m = a.comp[(long long)c.comp * a.nsrc + n];
        if (c.div4) v = v - m * sdrop / 4.0f;
        else v = v - m * sdrop;
    }
    f[tg.idx] = v;
}

__global__ void k_stressglut(SrcArgs a) { stressglut_body(a, blockIdx.x * blockDim.x + threadIdx.x, step_of(a)); }

// source__bodyforce, m_source.f90:850-898
__device__ static inline void bodyforce_body(const SrcArgs &a, int t, int it)
{
    if (t >= a.ntarget) return;
    if (it < 1 || it > a.nt) return;
    const SrcTarget tg = a.tg[t];
    float *f = a.fld[tg.field];
    float v = f[tg.idx];
    for (int q = tg.start; q < tg.start + tg.count; ++q) {
        const SrcContrib c = a.ct[q];
        const int n = c.src;
        const float stime = a.stf[(long long)(it - 1) * a.nsrc + n];
        const float b = 2.0f / (a.rho[c.rho_a] + a.rho[c.rho_b]);
        const float fc = a.comp[(long long)c.comp * a.nsrc + n];
        v = v + b * fc * stime * a.dt_dxyz / 2.0f;
    }
    f[tg.idx] = v;
}

__global__ void k_bodyforce(SrcArgs a) { bodyforce_body(a, blockIdx.x * blockDim.x + threadIdx.x, step_of(a)); }

// batched launches: context m = blockIdx.y, its arguments in HBM
__global__ void k_stressglut_b(const SrcArgs *__restrict__ args, int it, const int *d_it)
{
    stressglut_body(args[blockIdx.y], blockIdx.x * blockDim.x + threadIdx.x, d_it ? *d_it : it);
}
__global__ void k_bodyforce_b(const SrcArgs *__restrict__ args, int it, const int *d_it)
{
    bodyforce_body(args[blockIdx.y], blockIdx.x * blockDim.x + threadIdx.x, d_it ? *d_it : it);
}

static void srcargs_fill(swpc_ctx *c, SrcArgs &a, int it, const int *d_it)
{
    for (int q = 0; q < SWPC_NFIELD; ++q) a.fld[q] = c->fields[q];
    a.tg = c->d_targets;
    a.ct = c->d_contrib;
    a.comp = c->d_comp;
    a.mo = c->d_mo;
    a.stf = c->d_stf;
    a.rho = c->rho;
    a.ntarget = c->ntarget;
    a.nsrc = c->nsrc;
    a.nt = c->src_nt;
    a.dt_dxyz = c->dt_dxyz;
    a.it = it;
    a.d_it = d_it;
}

hipError_t launch_stressglut(swpc_ctx *c, int it, const int *d_it, hipStream_t s)
{
    if (c->nsrc <= 0 || c->bf_mode || c->ntarget <= 0) return hipSuccess;
    SrcArgs a;
    srcargs_fill(c, a, it, d_it);
    hipLaunchKernelGGL(k_stressglut, dim3((c->ntarget + 63) / 64), dim3(64), 0, (s ? s : c->stream), a);
    return hipGetLastError();
}

hipError_t launch_bodyforce(swpc_ctx *c, int it, const int *d_it, hipStream_t s)
{
    if (c->nsrc <= 0 || !c->bf_mode || c->ntarget <= 0) return hipSuccess;
    SrcArgs a;
    srcargs_fill(c, a, it, d_it);
    hipLaunchKernelGGL(k_bodyforce, dim3((c->ntarget + 63) / 64), dim3(64), 0, (s ? s : c->stream), a);
    return hipGetLastError();
}

// --------------------------------------------------------------------------
// wav__store, m_wav.f90:397-625 (one thread per station)
// --------------------------------------------------------------------------
struct WavArgs {
    Grid g;
    const float *Vx, *Vy, *Vz, *Sxx, *Syy, *Szz, *Syz, *Sxz, *Sxy;
    const int *ist, *jst, *kst;
    float *uacc, *eacc;
    float *wv, *wu, *ws, *we;
    int nst, what, ntdec, ntw;
    float M0, UC, dt;
    float r40x, r40y, r40z, r41x, r41y, r41z;
    int it;
    const int *d_it;
};

// station n of one context at step it
__device__ static inline void wav_store_body(const WavArgs &a, int n, int it)
{
    if (n >= a.nst) return;
    const Grid &g = a.g;
    const int i = a.ist[n], j = a.jst[n], k = a.kst[n];
    const long long P = g.plane;
    const int Y = g.pitch;
    const long long p = gidx(g, i, j, k);
    const float dt = a.dt;
    if (a.what & SWPC_WAV_U) {
        a.uacc[3 * n + 0] = a.uacc[3 * n + 0] + (a.Vx[p] + a.Vx[p - 1]) * 0.5f * dt;
        a.uacc[3 * n + 1] = a.uacc[3 * n + 1] + (a.Vy[p] + a.Vy[p - Y]) * 0.5f * dt;
        a.uacc[3 * n + 2] = a.uacc[3 * n + 2] - (a.Vz[p] + a.Vz[p - P]) * 0.5f * dt;
    }
    if (a.what & SWPC_WAV_STRAIN) {
        const float *Vx = a.Vx, *Vy = a.Vy, *Vz = a.Vz;
        const float r40x = a.r40x, r40y = a.r40y, r40z = a.r40z, r41x = a.r41x, r41y = a.r41y, r41z = a.r41z;
#define AT(F, di, dj, dk) F[p + (di) + (long long)(dj) * Y + (long long)(dk) * P]
        const float dxVx = (AT(Vx, 0, 0, 0) - AT(Vx, -1, 0, 0)) * r40x - (AT(Vx, 1, 0, 0) - AT(Vx, -2, 0, 0)) * r41x;
        const float dyVy = (AT(Vy, 0, 0, 0) - AT(Vy, 0, -1, 0)) * r40y - (AT(Vy, 0, 1, 0) - AT(Vy, 0, -2, 0)) * r41y;
        const float dzVz = (AT(Vz, 0, 0, 0) - AT(Vz, 0, 0, -1)) * r40z - (AT(Vz, 0, 0, 1) - AT(Vz, 0, 0, -2)) * r41z;
        const float dxVy = ((AT(Vy, 1, 0, 0) - AT(Vy, 0, 0, 0)) * r40x - (AT(Vy, 2, 0, 0) - AT(Vy, -1, 0, 0)) * r41x
                          + (AT(Vy, 1, -1, 0) - AT(Vy, 0, -1, 0)) * r40x - (AT(Vy, 2, -1, 0) - AT(Vy, -1, -1, 0)) * r41x
                          + (AT(Vy, 0, 0, 0) - AT(Vy, -1, 0, 0)) * r40x - (AT(Vy, 1, 0, 0) - AT(Vy, -2, 0, 0)) * r41x
                          + (AT(Vy, 0, -1, 0) - AT(Vy, -1, -1, 0)) * r40x - (AT(Vy, 1, -1, 0) - AT(Vy, -2, -1, 0)) * r41x) / 4.0f;
        const float dxVz = ((AT(Vz, 1, 0, 0) - AT(Vz, 0, 0, 0)) * r40x - (AT(Vz, 2, 0, 0) - AT(Vz, -1, 0, 0)) * r41x
                          + (AT(Vz, 1, 0, -1) - AT(Vz, 0, 0, -1)) * r40x - (AT(Vz, 2, 0, -1) - AT(Vz, -1, 0, -1)) * r41x
                          + (AT(Vz, 0, 0, 0) - AT(Vz, -1, 0, 0)) * r40x - (AT(Vz, 1, 0, 0) - AT(Vz, -2, 0, 0)) * r41x
                          + (AT(Vz, 0, 0, -1) - AT(Vz, -1, 0, -1)) * r40x - (AT(Vz, 1, 0, -1) - AT(Vz, -2, 0, -1)) * r41x) / 4.0f;
        const float dyVx = ((AT(Vx, 0, 1, 0) - AT(Vx, 0, 0, 0)) * r40y - (AT(Vx, 0, 2, 0) - AT(Vx, 0, -1, 0)) * r41y
                          + (AT(Vx, -1, 1, 0) - AT(Vx, -1, 0, 0)) * r40y - (AT(Vx, -1, 2, 0) - AT(Vx, -1, -1, 0)) * r41y
                          + (AT(Vx, 0, 0, 0) - AT(Vx, 0, -1, 0)) * r40y - (AT(Vx, 0, 1, 0) - AT(Vx, 0, -2, 0)) * r41y
                          + (AT(Vx, -1, 0, 0) - AT(Vx, -1, -1, 0)) * r40y - (AT(Vx, -1, 1, 0) - AT(Vx, -1, -2, 0)) * r41y) / 4.0f;
        const float dyVz = ((AT(Vz, 0, 1, 0) - AT(Vz, 0, 0, 0)) * r40y - (AT(Vz, 0, 2, 0) - AT(Vz, 0, -1, 0)) * r41y
                          + (AT(Vz, 0, 1, -1) - AT(Vz, 0, 0, -1)) * r40y - (AT(Vz, 0, 2, -1) - AT(Vz, 0, -1, -1)) * r41y
                          + (AT(Vz, 0, 0, 0) - AT(Vz, 0, -1, 0)) * r40y - (AT(Vz, 0, 1, 0) - AT(Vz, 0, -2, 0)) * r41y
                          + (AT(Vz, 0, 0, -1) - AT(Vz, 0, -1, -1)) * r40y - (AT(Vz, 0, 1, -1) - AT(Vz, 0, -2, -1)) * r41y) / 4.0f;
        const float dzVx = ((AT(Vx, 0, 0, 1) - AT(Vx, 0, 0, 0)) * r40z - (AT(Vx, 0, 0, 2) - AT(Vx, 0, 0, -1)) * r41z
                          + (AT(Vx, -1, 0, 1) - AT(Vx, -1, 0, 0)) * r40z - (AT(Vx, -1, 0, 2) - AT(Vx, -1, 0, -1)) * r41z
                          + (AT(Vx, 0, 0, 0) - AT(Vx, 0, 0, -1)) * r40z - (AT(Vx, 0, 0, 1) - AT(Vx, 0, 0, -2)) * r41z
                          + (AT(Vx, -1, 0, 0) - AT(Vx, -1, 0, -1)) * r40z - (AT(Vx, -1, 0, 1) - AT(Vx, -1, 0, -2)) * r41z) / 4.0f;
        const float dzVy = ((AT(Vy, 0, 0, 1) - AT(Vy, 0, 0, 0)) * r40z - (AT(Vy, 0, 0, 2) - AT(Vy, 0, 0, -1)) * r41z
                          + (AT(Vy, 0, -1, 1) - AT(Vy, 0, -1, 0)) * r40z - (AT(Vy, 0, -1, 2) - AT(Vy, 0, -1, -1)) * r41z
                          + (AT(Vy, 0, 0, 0) - AT(Vy, 0, 0, -1)) * r40z - (AT(Vy, 0, 0, 1) - AT(Vy, 0, 0, -2)) * r41z
                          + (AT(Vy, 0, -1, 0) - AT(Vy, 0, -1, -1)) * r40z - (AT(Vy, 0, -1, 1) - AT(Vy, 0, -1, -2)) * r41z) / 4.0f;
#undef AT
        float *e = a.eacc + 6 * n;
        e[0] = e[0] + dxVx * dt;
        e[1] = e[1] + dyVy * dt;
        e[2] = e[2] + dzVz * dt;
        e[3] = e[3] + (dyVz + dzVy) / 2.0f * dt;
        e[4] = e[4] + (dxVz + dzVx) / 2.0f * dt;
        e[5] = e[5] + (dxVy + dyVx) / 2.0f * dt;
    }
    if ((it - 1) % a.ntdec != 0) return;
    const int itw = (it - 1) / a.ntdec;
    if (itw < 0 || itw >= a.ntw) return;
    const float M0 = a.M0, UC = a.UC;
    const long long ntw = a.ntw;
    if (a.what & SWPC_WAV_V) {
        a.wv[(n * 3LL + 0) * ntw + itw] = (a.Vx[p] + a.Vx[p - 1]) / 2.0f * M0 * UC * 1e9f;
        a.wv[(n * 3LL + 1) * ntw + itw] = (a.Vy[p] + a.Vy[p - Y]) / 2.0f * M0 * UC * 1e9f;
        a.wv[(n * 3LL + 2) * ntw + itw] = -((a.Vz[p] + a.Vz[p - P]) / 2.0f * M0 * UC * 1e9f);
    }
    if (a.what & SWPC_WAV_U) {
        for (int q = 0; q < 3; ++q) a.wu[(n * 3LL + q) * ntw + itw] = a.uacc[3 * n + q] * M0 * UC * 1e9f;
    }
    if (a.what & SWPC_WAV_STRESS) {
        a.ws[(n * 6LL + 0) * ntw + itw] = a.Sxx[p] * M0 * UC * 1e6f;
        a.ws[(n * 6LL + 1) * ntw + itw] = a.Syy[p] * M0 * UC * 1e6f;
        a.ws[(n * 6LL + 2) * ntw + itw] = a.Szz[p] * M0 * UC * 1e6f;
        a.ws[(n * 6LL + 3) * ntw + itw] = (a.Syz[p] + a.Syz[p - Y] + a.Syz[p - P] + a.Syz[p - Y - P]) / 4.0f * M0 * UC * 1e6f;
        a.ws[(n * 6LL + 4) * ntw + itw] = (a.Sxz[p] + a.Sxz[p - 1] + a.Sxz[p - P] + a.Sxz[p - 1 - P]) / 4.0f * M0 * UC * 1e6f;
        a.ws[(n * 6LL + 5) * ntw + itw] = (a.Sxy[p] + a.Sxy[p - Y] + a.Sxy[p - 1] + a.Sxy[p - 1 - Y]) / 4.0f * M0 * UC * 1e6f;
    }
    if (a.what & SWPC_WAV_STRAIN) {
        for (int q = 0; q < 6; ++q) a.we[(n * 6LL + q) * ntw + itw] = a.eacc[6 * n + q] * M0 * UC * 1e-3f;
    }
}

__global__ void k_wav_store(WavArgs a)
{
    wav_store_body(a, blockIdx.x * blockDim.x + threadIdx.x, a.d_it ? *a.d_it : a.it);
}

__global__ void k_wav_store_b(const WavArgs *__restrict__ args, int it, const int *d_it)
{
    wav_store_body(args[blockIdx.y], blockIdx.x * blockDim.x + threadIdx.x, d_it ? *d_it : it);
}

static void wavargs_fill(swpc_ctx *c, WavArgs &a, int it, const int *d_it)
{
    a.g = c->g;
    a.Vx = c->fields[SWPC_VX]; a.Vy = c->fields[SWPC_VY]; a.Vz = c->fields[SWPC_VZ];
    a.Sxx = c->fields[SWPC_SXX]; a.Syy = c->fields[SWPC_SYY]; a.Szz = c->fields[SWPC_SZZ];
    a.Syz = c->fields[SWPC_SYZ]; a.Sxz = c->fields[SWPC_SXZ]; a.Sxy = c->fields[SWPC_SXY];
    a.ist = c->d_ist; a.jst = c->d_jst; a.kst = c->d_kst;
    a.uacc = c->d_uacc; a.eacc = c->d_eacc;
    a.wv = c->d_wav[0]; a.wu = c->d_wav[1]; a.ws = c->d_wav[2]; a.we = c->d_wav[3];
    a.nst = c->nst; a.what = c->wav_what; a.ntdec = c->ntdec_w; a.ntw = c->ntw;
    a.M0 = c->wav_M0; a.UC = c->wav_UC; a.dt = c->desc.dt;
    const float dx = (float)c->desc.dx, dy = (float)c->desc.dy, dz = (float)c->desc.dz;
    a.r40x = 9.0f / 8.0f / dx; a.r40y = 9.0f / 8.0f / dy; a.r40z = 9.0f / 8.0f / dz;
    a.r41x = 1.0f / 24.0f / dx; a.r41y = 1.0f / 24.0f / dy; a.r41z = 1.0f / 24.0f / dz;
    a.it = it;
    a.d_it = d_it;
}

hipError_t launch_wav_store(swpc_ctx *c, int it, const int *d_it, hipStream_t s)
{
    if (c->nst <= 0) return hipSuccess;
    WavArgs a;
    wavargs_fill(c, a, it, d_it);
    hipLaunchKernelGGL(k_wav_store, dim3((c->nst + 63) / 64), dim3(64), 0, (s ? s : c->stream), a);
    return hipGetLastError();
}

// --------------------------------------------------------------------------
// batched receivers and sources: one launch for every member context
// --------------------------------------------------------------------------
size_t batch_aux_size(int kind)
{
    if (kind == BATCH_AUX_GREEN_STORE || kind == BATCH_AUX_GREEN_SOURCE) return batch_green_size(kind);
    return kind == BATCH_AUX_WAV ? sizeof(WavArgs) : sizeof(SrcArgs);
}

int batch_aux_fill(swpc_ctx *c, int kind, void *out)
{
    if (kind == BATCH_AUX_GREEN_STORE || kind == BATCH_AUX_GREEN_SOURCE) return batch_green_fill(c, kind, out);
    if (kind == BATCH_AUX_WAV) {
        if (c->nst <= 0) return 0;
        WavArgs a;
        wavargs_fill(c, a, 0, nullptr);
        memcpy(out, &a, sizeof(a));
        return c->nst;
    }
    const bool bf = kind == BATCH_AUX_BODYFORCE;
    if (c->nsrc <= 0 || c->ntarget <= 0 || (c->bf_mode != 0) != bf) return 0;
    SrcArgs a;
    srcargs_fill(c, a, 0, nullptr);
    memcpy(out, &a, sizeof(a));
    return c->ntarget;
}

hipError_t launch_batch_aux(int kind, const void *d_args, int nmem, int nmax, int it, const int *d_it, hipStream_t s)
{
    if (nmem <= 0 || nmax <= 0) return hipSuccess;
    if (kind == BATCH_AUX_GREEN_STORE || kind == BATCH_AUX_GREEN_SOURCE)
        return launch_batch_green(kind, d_args, nmem, nmax, it, d_it, s);
    const dim3 grid((nmax + 63) / 64, nmem), block(64);
    if (kind == BATCH_AUX_WAV) hipLaunchKernelGGL(k_wav_store_b, grid, block, 0, s, (const WavArgs *)d_args, it, d_it);
    else if (kind == BATCH_AUX_BODYFORCE)
        hipLaunchKernelGGL(k_bodyforce_b, grid, block, 0, s, (const SrcArgs *)d_args, it, d_it);
    else hipLaunchKernelGGL(k_stressglut_b, grid, block, 0, s, (const SrcArgs *)d_args, it, d_it);
    return hipGetLastError();
}

// --------------------------------------------------------------------------
// device step counter used by graph-captured step blocks
// --------------------------------------------------------------------------
__global__ void k_step_advance(int *d_it)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *d_it = *d_it + 1;
}

hipError_t launch_step_advance(swpc_ctx *c)
{
    hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, c->stream, c->d_it);
    return hipGetLastError();
}

hipError_t launch_counter_advance(int *d, hipStream_t s)
{
    hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, s, d);
    return hipGetLastError();
}

__global__ void k_set_counter(int *p, int v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

hipError_t launch_set_counter(int *d, int v, hipStream_t s)
{
    hipLaunchKernelGGL(k_set_counter, dim3(1), dim3(64), 0, s, d, v);
    return hipGetLastError();
}

// --------------------------------------------------------------------------
// kernel__vmax, m_kernel.f90:350-374.  |V| >= 0, so the float bit patterns
// order like the values and an integer atomicMax is exact.
// --------------------------------------------------------------------------
struct VmaxArgs {
    Grid g;
    const float *Vx, *Vy, *Vz;
    const int *kob;
    int i0, i1, j0, j1;
    unsigned int *out;
};

__global__ void k_vmax_clear(unsigned int *out)
{
    if (threadIdx.x < 3) out[threadIdx.x] = 0u;
}

__global__ void k_vmax(VmaxArgs a)
{
    const int i = a.i0 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = a.j0 + blockIdx.y;
    float x = 0.0f, y = 0.0f, z = 0.0f;
    if (i <= a.i1 && j <= a.j1) {
        const int k = a.kob[g2idx(a.g, i, j)] + 1;
        if (k >= 1 && k <= a.g.nz) {  // kob comes from the caller: never index outside the grid
            const long long p = gidx(a.g, i, j, k);
            x = fabsf(a.Vx[p]);
            y = fabsf(a.Vy[p]);
            z = fabsf(a.Vz[p]);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        x = fmaxf(x, __shfl_xor(x, o));
        y = fmaxf(y, __shfl_xor(y, o));
        z = fmaxf(z, __shfl_xor(z, o));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&a.out[0], __float_as_uint(x));
        atomicMax(&a.out[1], __float_as_uint(y));
        atomicMax(&a.out[2], __float_as_uint(z));
    }
}

hipError_t launch_vmax(swpc_ctx *c)
{
    const Grid &g = c->g;
    const int margin = 5;
    VmaxArgs a;
    a.g = g;
    a.Vx = c->fields[SWPC_VX]; a.Vy = c->fields[SWPC_VY]; a.Vz = c->fields[SWPC_VZ];
    a.kob = c->kob;
    a.j0 = max(g.na + margin + 1, g.jb);
    a.j1 = min(g.ny - g.na - margin, g.je);
    a.i0 = max(g.na + margin + 1, g.ib);
    a.i1 = min(g.nx - g.na - margin, g.ie);
    a.out = (unsigned int *)c->d_vmax;
    hipLaunchKernelGGL(k_vmax_clear, dim3(1), dim3(64), 0, c->stream, a.out);
    if (a.i1 >= a.i0 && a.j1 >= a.j0) {
        dim3 grid((a.i1 - a.i0 + 1 + 255) / 256, a.j1 - a.j0 + 1);
        hipLaunchKernelGGL(k_vmax, grid, dim3(256), 0, c->stream, a);
    }
    return hipGetLastError();
}

// --------------------------------------------------------------------------
// kernel__vmax without allocation (swpc_kernel_vmax_noalloc), for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops).  The sampling blocks leave their partial maxima in
// `scratch`, a file-scope device array, and one block reduces them into the
// caller's device output.  Every partial is written before it is read, so
// nothing needs clearing.  Each GPU holds its own copy of the array (192 KiB)
// until the process ends; calls that overlap on one GPU share it and race.
// --------------------------------------------------------------------------
#define VMAX_SCRATCH_BLOCKS 16384
static __device__ __align__(256) float scratch[3 * VMAX_SCRATCH_BLOCKS];

// the sampling of k_vmax; partials (x, y, z) of block b at part[3b .. 3b+2]
__global__ void __launch_bounds__(256) k_vmax_part(VmaxArgs a, float *part)
{
    __shared__ float w[4][3];
    const int i = a.i0 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = a.j0 + blockIdx.y;
    float x = 0.0f, y = 0.0f, z = 0.0f;
    if (i <= a.i1 && j <= a.j1) {
        const int k = a.kob[g2idx(a.g, i, j)] + 1;
        if (k >= 1 && k <= a.g.nz) {
            const long long p = gidx(a.g, i, j, k);
            x = fabsf(a.Vx[p]);
            y = fabsf(a.Vy[p]);
            z = fabsf(a.Vz[p]);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        x = fmaxf(x, __shfl_xor(x, o));
        y = fmaxf(y, __shfl_xor(y, o));
        z = fmaxf(z, __shfl_xor(z, o));
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        w[wave][0] = x;
        w[wave][1] = y;
        w[wave][2] = z;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const float m = fmaxf(fmaxf(w[0][threadIdx.x], w[1][threadIdx.x]), fmaxf(w[2][threadIdx.x], w[3][threadIdx.x]));
        part[3LL * (blockIdx.y * (long long)gridDim.x + blockIdx.x) + threadIdx.x] = m;
    }
}

// one block: the maxima of nb partials into out[0..2] (0 when nb = 0)
__global__ void __launch_bounds__(256) k_vmax_reduce(const float *part, int nb, float *out)
{
    __shared__ float w[4][3];
    float x = 0.0f, y = 0.0f, z = 0.0f;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        x = fmaxf(x, part[3 * b]);
        y = fmaxf(y, part[3 * b + 1]);
        z = fmaxf(z, part[3 * b + 2]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        x = fmaxf(x, __shfl_xor(x, o));
        y = fmaxf(y, __shfl_xor(y, o));
        z = fmaxf(z, __shfl_xor(z, o));
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        w[wave][0] = x;
        w[wave][1] = y;
        w[wave][2] = z;
    }
    __syncthreads();
    if (threadIdx.x < 3)
        out[threadIdx.x] = fmaxf(fmaxf(w[0][threadIdx.x], w[1][threadIdx.x]), fmaxf(w[2][threadIdx.x], w[3][threadIdx.x]));
}

// Device address of `scratch` on each device, taken once per device with
// that device current (the array's host-side name is not a device address).
#define VMAX_SCRATCH_MAXDEV 64
static float *g_scratch_ptr[VMAX_SCRATCH_MAXDEV];
static std::mutex g_scratch_mu;

static hipError_t scratch_ptr(int dev, float **out)
{
    if (dev < 0 || dev >= VMAX_SCRATCH_MAXDEV) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = (float *)p;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

extern "C" int swpc_kernel_vmax_noalloc(swpc_ctx *c, float *d_out)
{
    if (!c || !d_out) return SWPC_EINVAL;
    const Grid &g = c->g;
    const int margin = 5;
    VmaxArgs a;
    a.g = g;
    a.j0 = max(g.na + margin + 1, g.jb);
    a.j1 = min(g.ny - g.na - margin, g.je);
    a.i0 = max(g.na + margin + 1, g.ib);
    a.i1 = min(g.nx - g.na - margin, g.ie);
    const bool any = a.i1 >= a.i0 && a.j1 >= a.j0;
    const long long nbx = any ? (a.i1 - a.i0 + 1 + 255) / 256 : 0, nby = any ? a.j1 - a.j0 + 1 : 0;
    if (nbx * nby > VMAX_SCRATCH_BLOCKS) {
        char msg[200];
        snprintf(msg, sizeof msg, "swpc_kernel_vmax_noalloc: the sampled surface needs %lld blocks of 256 cells, "
                 "beyond the %d its scratch holds", nbx * nby, VMAX_SCRATCH_BLOCKS);
        return ctx_fail(c, SWPC_EINVAL, msg);
    }
    if (!c->has_medium) return ctx_fail(c, SWPC_ESTATE, "swpc_kernel_vmax_noalloc: medium not set");
    hipError_t e = hipSetDevice(c->device);
    // the output must be device memory of the context's GPU
    hipPointerAttribute_t at;
    if (e == hipSuccess && (hipPointerGetAttributes(&at, d_out) != hipSuccess || at.type != hipMemoryTypeDevice ||
                            at.device != c->device)) {
        (void)hipGetLastError();
        return ctx_fail(c, SWPC_EINVAL, "swpc_kernel_vmax_noalloc: d_out is not device memory of the context's GPU");
    }
    float *part = nullptr;
    if (e == hipSuccess) e = scratch_ptr(c->device, &part);
    if (e != hipSuccess) return ctx_fail(c, SWPC_EHIP, hipGetErrorString(e));
    a.Vx = c->fields[SWPC_VX]; a.Vy = c->fields[SWPC_VY]; a.Vz = c->fields[SWPC_VZ];
    a.kob = c->kob;
    a.out = nullptr;
    if (any) hipLaunchKernelGGL(k_vmax_part, dim3((unsigned)nbx, (unsigned)nby), dim3(256), 0, c->stream, a, part);
    hipLaunchKernelGGL(k_vmax_reduce, dim3(1), dim3(256), 0, c->stream, part, (int)(nbx * nby), d_out);
    e = hipGetLastError();
    if (e != hipSuccess) return ctx_fail(c, SWPC_EHIP, hipGetErrorString(e));
    return SWPC_OK;
}

// --------------------------------------------------------------------------
// layout transposes: reference (k,i,j) z-fastest halo-3 <-> device x-fastest.
// Tiled 32 x 32 over (k, i) for each j, through LDS.
// --------------------------------------------------------------------------
struct TrArgs {
    Grid g;
    const float *src;
    float *dst;
    int nzr, nxr;   // reference extents nz+6, nx+6
};

__global__ void __launch_bounds__(256) k_ref_to_dev(TrArgs a)
{
    __shared__ float tile[32][33];
    const int jr = blockIdx.z;                       // 0..ny+5
    const int k0 = blockIdx.y * 32, i0 = blockIdx.x * 32;
    const long long jbase = (long long)jr * a.nxr * a.nzr;
    for (int r = threadIdx.y; r < 32; r += 8) {     // read: k fastest
        const int i = i0 + r, k = k0 + threadIdx.x;
        if (i < a.nxr && k < a.nzr) tile[r][threadIdx.x] = a.src[jbase + (long long)i * a.nzr + k];
    }
    __syncthreads();
    for (int r = threadIdx.y; r < 32; r += 8) {     // write: i fastest
        const int k = k0 + r, i = i0 + threadIdx.x;
        if (i < a.nxr && k < a.nzr)
            a.dst[(long long)k * a.g.plane + (long long)jr * a.g.pitch + (i - 3 + a.g.ox)] = tile[threadIdx.x][r];
    }
}

__global__ void __launch_bounds__(256) k_dev_to_ref(TrArgs a)
{
    __shared__ float tile[32][33];
    const int jr = blockIdx.z;
    const int k0 = blockIdx.y * 32, i0 = blockIdx.x * 32;
    const long long jbase = (long long)jr * a.nxr * a.nzr;
    for (int r = threadIdx.y; r < 32; r += 8) {     // read: i fastest
        const int k = k0 + r, i = i0 + threadIdx.x;
        if (i < a.nxr && k < a.nzr)
            tile[r][threadIdx.x] = a.src[(long long)k * a.g.plane + (long long)jr * a.g.pitch + (i - 3 + a.g.ox)];
    }
    __syncthreads();
    for (int r = threadIdx.y; r < 32; r += 8) {     // write: k fastest
        const int i = i0 + r, k = k0 + threadIdx.x;
        if (i < a.nxr && k < a.nzr) a.dst[jbase + (long long)i * a.nzr + k] = tile[threadIdx.x][r];
    }
}

hipError_t launch_ref_to_dev(swpc_ctx *c, const float *src_ref, float *dst, hipStream_t s)
{
    TrArgs a;
    a.g = c->g;
    a.src = src_ref;
    a.dst = dst;
    a.nzr = c->g.nz + 6;
    a.nxr = c->g.nx + 6;
    dim3 grid((a.nxr + 31) / 32, (a.nzr + 31) / 32, c->g.ny + 6);
    hipLaunchKernelGGL(k_ref_to_dev, grid, dim3(32, 8), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_dev_to_ref(swpc_ctx *c, const float *src, float *dst_ref, hipStream_t s)
{
    TrArgs a;
    a.g = c->g;
    a.src = src;
    a.dst = dst_ref;
    a.nzr = c->g.nz + 6;
    a.nxr = c->g.nx + 6;
    dim3 grid((a.nxr + 31) / 32, (a.nzr + 31) / 32, c->g.ny + 6);
    hipLaunchKernelGGL(k_dev_to_ref, grid, dim3(32, 8), 0, s, a);
    return hipGetLastError();
}
