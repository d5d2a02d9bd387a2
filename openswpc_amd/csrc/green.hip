// green.hip -- reciprocity Green's-function mode of libswpc_gpu
// (src/swpc_3d/m_green.f90): the strain Green's functions sampled at the GF
// cells (green__store) and the unit force at the virtual source, a station
// (green__source).
//
// green__store is a gather of 4th-order staggered derivatives at a few
// scattered cells: one thread per GF cell, accumulators component-major so
// that each of the twelve read-modify-writes is coalesced.  Every expression
// keeps the reference's FP32 evaluation order (the FP32 build, MP = SP).
#include <cstring>

#include "swpc_internal.h"

// UC_DERIV = 10.0**(-15) and UC_BF = 10.0**(-12) as the Fortran compiler
// folds them to single precision (m_green.f90:30-31)
#define GREEN_UC_DERIV 0x1.203afcp-50f
#define GREEN_UC_BF 0x1.197998p-40f

struct GreenArgs {
    const float *Vx, *Vy, *Vz;
    const int *ig, *jg, *kg;
    float *acc;   // (12, ng): dxUx dxUy dxUz dyUx dyUy dyUz dzUx dzUy dzUz Ux Uy Uz
    float *gf;    // (ntw, ncmp*ng), Fortran order
    Grid g;
    int ng, ncmp, ntdec, ntw;
    float dt;
    float r40x, r40y, r40z, r41x, r41y, r41z;
    int it;
    const int *d_it;
};

// (F(c) - F(c-s)) * r40 - (F(c+s) - F(c-2s)) * r41: the 4th-order staggered
// derivative whose upper inner node is c, stride s along the axis
__device__ static inline float dstag(const float *F, long long c, long long s, float r40, float r41)
{
    return (F[c] - F[c - s]) * r40 - (F[c + s] - F[c - 2 * s]) * r41;
}

// green__store, m_green.f90:406-545
// GF point n of one context at step it
__device__ static inline void green_store_body(const GreenArgs &a, int n, int it)
{
    if (n >= a.ng) return;
    const Grid &g = a.g;
    const long long p = gidx(g, a.ig[n], a.jg[n], a.kg[n]);
    const long long X = 1, Y = g.pitch, Z = g.plane;
    const float *Vx = a.Vx, *Vy = a.Vy, *Vz = a.Vz;
    const float r40x = a.r40x, r40y = a.r40y, r40z = a.r40z, r41x = a.r41x, r41y = a.r41y, r41z = a.r41z;

    // m_green.f90:458-460
    const float dxVx = dstag(Vx, p, X, r40x, r41x);
    const float dyVy = dstag(Vy, p, Y, r40y, r41y);
    const float dzVz = dstag(Vz, p, Z, r40z, r41z);
    // the four staggered derivatives around the cell, m_green.f90:462-483
    const float dxVy1 = dstag(Vy, p + X, X, r40x, r41x), dxVy2 = dstag(Vy, p + X - Y, X, r40x, r41x);
    const float dxVy3 = dstag(Vy, p, X, r40x, r41x), dxVy4 = dstag(Vy, p - Y, X, r40x, r41x);
    const float dxVz1 = dstag(Vz, p + X, X, r40x, r41x), dxVz2 = dstag(Vz, p + X - Z, X, r40x, r41x);
    const float dxVz3 = dstag(Vz, p, X, r40x, r41x), dxVz4 = dstag(Vz, p - Z, X, r40x, r41x);
    const float dyVx1 = dstag(Vx, p + Y, Y, r40y, r41y), dyVx2 = dstag(Vx, p - X + Y, Y, r40y, r41y);
    const float dyVx3 = dstag(Vx, p, Y, r40y, r41y), dyVx4 = dstag(Vx, p - X, Y, r40y, r41y);
    const float dyVz1 = dstag(Vz, p + Y, Y, r40y, r41y), dyVz2 = dstag(Vz, p - Z + Y, Y, r40y, r41y);
    const float dyVz3 = dstag(Vz, p, Y, r40y, r41y), dyVz4 = dstag(Vz, p - Z, Y, r40y, r41y);
    const float dzVx1 = dstag(Vx, p + Z, Z, r40z, r41z), dzVx2 = dstag(Vx, p - X + Z, Z, r40z, r41z);
    const float dzVx3 = dstag(Vx, p, Z, r40z, r41z), dzVx4 = dstag(Vx, p - X, Z, r40z, r41z);
    const float dzVy1 = dstag(Vy, p + Z, Z, r40z, r41z), dzVy2 = dstag(Vy, p - Y + Z, Z, r40z, r41z);
    const float dzVy3 = dstag(Vy, p, Z, r40z, r41z), dzVy4 = dstag(Vy, p - Y, Z, r40z, r41z);
    // their means, m_green.f90:485-490
    const float dxVy = (dxVy1 + dxVy2 + dxVy3 + dxVy4) * 0.25f;
    const float dxVz = (dxVz1 + dxVz2 + dxVz3 + dxVz4) * 0.25f;
    const float dyVx = (dyVx1 + dyVx2 + dyVx3 + dyVx4) * 0.25f;
    const float dyVz = (dyVz1 + dyVz2 + dyVz3 + dyVz4) * 0.25f;
    const float dzVx = (dzVx1 + dzVx2 + dzVx3 + dzVx4) * 0.25f;
    const float dzVy = (dzVy1 + dzVy2 + dzVy3 + dzVy4) * 0.25f;

    // time integration, m_green.f90:492-506
    const long long ng = a.ng;
    const float dt = a.dt;
    float *acc = a.acc + n;
    const float d[9] = {dxVx, dxVy, dxVz, dyVx, dyVy, dyVz, dzVx, dzVy, dzVz};
#pragma unroll
    for (int q = 0; q < 9; ++q) acc[q * ng] = acc[q * ng] + d[q] * dt;
    const bool bforce = a.ncmp == 9;
    if (bforce) {
        acc[9 * ng] = acc[9 * ng] + 0.5f * (Vx[p] + Vx[p - X]) * dt;
        acc[10 * ng] = acc[10 * ng] + 0.5f * (Vy[p] + Vy[p - Y]) * dt;
        acc[11 * ng] = acc[11 * ng] + 0.5f * (Vz[p] + Vz[p - Z]) * dt;
    }

    // output every ntdec_w steps, m_green.f90:510-541
    if ((it - 1) % a.ntdec != 0) return;
    const int itw = (it - 1) / a.ntdec;
    if (itw < 0 || itw >= a.ntw) return;
    const float ucd = GREEN_UC_DERIV, ucb = GREEN_UC_BF;
    const long long ntw = a.ntw;
    float *gf = a.gf + (long long)n * a.ncmp * ntw + itw;
    gf[0 * ntw] = acc[0] * ucd * 1e9f;
    gf[1 * ntw] = acc[4 * ng] * ucd * 1e9f;
    gf[2 * ntw] = acc[8 * ng] * ucd * 1e9f;
    gf[3 * ntw] = (acc[5 * ng] + acc[7 * ng]) * ucd * 1e9f;
    gf[4 * ntw] = (acc[2 * ng] + acc[6 * ng]) * ucd * 1e9f;
    gf[5 * ntw] = (acc[1 * ng] + acc[3 * ng]) * ucd * 1e9f;
    if (bforce) {
        gf[6 * ntw] = acc[9 * ng] * ucb * 1e9f;
        gf[7 * ntw] = acc[10 * ng] * ucb * 1e9f;
        gf[8 * ntw] = acc[11 * ng] * ucb * 1e9f;
    }
}

__global__ void __launch_bounds__(64) k_green_store(GreenArgs a)
{
    green_store_body(a, blockIdx.x * blockDim.x + threadIdx.x, a.d_it ? *a.d_it : a.it);
}

// batched: context m = blockIdx.y, its arguments in HBM
__global__ void __launch_bounds__(64) k_green_store_b(const GreenArgs *__restrict__ args, int it, const int *d_it)
{
    green_store_body(args[blockIdx.y], blockIdx.x * blockDim.x + threadIdx.x, d_it ? *d_it : it);
}

static void green_store_fill(swpc_ctx *c, GreenArgs &a, int it, const int *d_it)
{
    a.Vx = c->fields[SWPC_VX]; a.Vy = c->fields[SWPC_VY]; a.Vz = c->fields[SWPC_VZ];
    a.ig = c->d_green_ijk;
    a.jg = c->d_green_ijk + c->green_ng;
    a.kg = c->d_green_ijk + 2 * (long long)c->green_ng;
    a.acc = c->d_green_acc;
    a.gf = c->d_gf;
    a.g = c->g;
    a.ng = c->green_ng; a.ncmp = c->green_ncmp; a.ntdec = c->green_ntdec; a.ntw = c->green_ntw;
    a.dt = c->desc.dt;
    // r40 = C40/d, r41 = C41/d with C40 = 9/8, C41 = 1/24 (m_green.f90:43-45,110-115)
    const float dx = (float)c->desc.dx, dy = (float)c->desc.dy, dz = (float)c->desc.dz;
    const float C40 = 9.0f / 8.0f, C41 = 1.0f / 24.0f;
    a.r40x = C40 / dx; a.r40y = C40 / dy; a.r40z = C40 / dz;
    a.r41x = C41 / dx; a.r41y = C41 / dy; a.r41z = C41 / dz;
    a.it = it;
    a.d_it = d_it;
}

hipError_t launch_green_store(swpc_ctx *c, int it, const int *d_it, hipStream_t s)
{
    if (!c->green_on || c->green_ng <= 0) return hipSuccess;
    GreenArgs a;
    green_store_fill(c, a, it, d_it);
    hipLaunchKernelGGL(k_green_store, dim3((c->green_ng + 63) / 64), dim3(64), 0, (s ? s : c->stream), a);
    return hipGetLastError();
}

struct GreenSrcArgs {
    float *Vx, *Vy, *Vz;
    const float *rho;
    const float *stf;
    Grid g;
    int i, j, k, nt;
    float fx1, fy1, fz1, dt_dxyz;
    int it;
    const int *d_it;
};

// green__source, m_green.f90:614-649: the force split onto the two staggered
// velocity nodes of each axis with the averaged buoyancy; all six updates
// are made, the zero ones included
__device__ static inline void green_source_body(const GreenSrcArgs &a, int it)
{
    if (it < 1 || it > a.nt) return;
    const Grid &g = a.g;
    const long long p = gidx(g, a.i, a.j, a.k);
    const long long Y = g.pitch, Z = g.plane;
    const float st = a.stf[it - 1];
    const float fx = a.fx1 * a.dt_dxyz * st;
    const float fy = a.fy1 * a.dt_dxyz * st;
    const float fz = a.fz1 * a.dt_dxyz * st;
    const float *rho = a.rho;
    const float r0 = rho[p];
    a.Vx[p] = a.Vx[p] + (2.0f / (r0 + rho[p + 1])) * fx / 2.0f;
    a.Vx[p - 1] = a.Vx[p - 1] + (2.0f / (r0 + rho[p - 1])) * fx / 2.0f;
    a.Vy[p] = a.Vy[p] + (2.0f / (r0 + rho[p + Y])) * fy / 2.0f;
    a.Vy[p - Y] = a.Vy[p - Y] + (2.0f / (r0 + rho[p - Y])) * fy / 2.0f;
    a.Vz[p] = a.Vz[p] + (2.0f / (r0 + rho[p + Z])) * fz / 2.0f;
    a.Vz[p - Z] = a.Vz[p - Z] + (2.0f / (r0 + rho[p - Z])) * fz / 2.0f;
}

__global__ void k_green_source(GreenSrcArgs a)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    green_source_body(a, a.d_it ? *a.d_it : a.it);
}

// batched: context m = blockIdx.y
__global__ void k_green_source_b(const GreenSrcArgs *__restrict__ args, int it, const int *d_it)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    green_source_body(args[blockIdx.y], d_it ? *d_it : it);
}

static void green_source_fill(swpc_ctx *c, GreenSrcArgs &a, int it, const int *d_it)
{
    a.Vx = c->fields[SWPC_VX]; a.Vy = c->fields[SWPC_VY]; a.Vz = c->fields[SWPC_VZ];
    a.rho = c->rho;
    a.stf = c->d_green_stf;
    a.g = c->g;
    a.i = c->green_src[0]; a.j = c->green_src[1]; a.k = c->green_src[2];
    a.nt = c->green_nt;
    a.fx1 = c->green_f[0]; a.fy1 = c->green_f[1]; a.fz1 = c->green_f[2];
    a.dt_dxyz = c->green_dt_dxyz;
    a.it = it;
    a.d_it = d_it;
}

hipError_t launch_green_source(swpc_ctx *c, int it, const int *d_it, hipStream_t s)
{
    if (!c->green_on || c->green_nt <= 0) return hipSuccess;
    GreenSrcArgs a;
    green_source_fill(c, a, it, d_it);
    hipLaunchKernelGGL(k_green_source, dim3(1), dim3(64), 0, (s ? s : c->stream), a);
    return hipGetLastError();
}

// batched Green's-function store and source (swpc_batch): per-member
// arguments, one launch for every member that has them
size_t batch_green_size(int kind) { return kind == BATCH_AUX_GREEN_STORE ? sizeof(GreenArgs) : sizeof(GreenSrcArgs); }

int batch_green_fill(swpc_ctx *c, int kind, void *out)
{
    if (!c->green_on) return 0;
    if (kind == BATCH_AUX_GREEN_STORE) {
        if (c->green_ng <= 0) return 0;
        GreenArgs a;
        green_store_fill(c, a, 0, nullptr);
        memcpy(out, &a, sizeof(a));
        return c->green_ng;
    }
    if (c->green_nt <= 0) return 0;
    GreenSrcArgs a;
    green_source_fill(c, a, 0, nullptr);
    memcpy(out, &a, sizeof(a));
    return 1;
}

hipError_t launch_batch_green(int kind, const void *d_args, int nmem, int nmax, int it, const int *d_it,
                              hipStream_t s)
{
    if (nmem <= 0 || nmax <= 0) return hipSuccess;
    if (kind == BATCH_AUX_GREEN_STORE)
        hipLaunchKernelGGL(k_green_store_b, dim3((nmax + 63) / 64, nmem), dim3(64), 0, s, (const GreenArgs *)d_args,
                           it, d_it);
    else
        hipLaunchKernelGGL(k_green_source_b, dim3(1, nmem), dim3(64), 0, s, (const GreenSrcArgs *)d_args, it, d_it);
    return hipGetLastError();
}
