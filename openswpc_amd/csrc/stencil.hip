// stencil.hip -- velocity/stress updates of OpenSWPC's 3D loop on gfx950.
//
// Interior (m_kernel) kernels march along z: each thread owns one (i,j)
// column of a 64 x BY tile and walks kchunk planes, keeping the z-neighbours
// of its column in registers (a rotating queue), while x/y neighbours come
// from the plane currently streamed by the tile (L1/L2 resident).  Every
// arithmetic expression keeps the reference's FP32 evaluation order; the
// library is compiled with -ffp-contract=off so no FMA changes the rounding.
//
// PML kernels (m_absorb_p) run one thread per PML cell over a compact
// numbering of the five PML boxes; their 18 auxiliary arrays exist only
// there.  Cerjan tapers (m_absorb_c) are plain per-cell kernels.
#include "swpc_internal.h"

#define BX 64

// Blocks are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8).
// Give each XCD a contiguous run of logical tiles so tiles that share halo
// rows share one L2.
__device__ static inline int xcd_remap(int b, int n)
{
    const int x = b & 7, t = b >> 3;
    const int q = n >> 3, r = n & 7;
    return x < r ? x * (q + 1) + t : r * (q + 1) + (x - r) * q + t;
}

// isign of m_kernel.f90:103-104: true inside the 2nd-order band
__device__ static inline bool in_band(int k, int kft, int kfb, int kot, int kob)
{
    const int a = (k - kft) * (kfb - k);
    const int b = (k - kot) * (kob - k);
    return (a > b ? a : b) >= 0;
}

// harmonic mean of mu over four cells, m_kernel.f90:293-309
__device__ static inline float hmean4(float a, float b, float c, float d, float epsl)
{
    return 4.0f * a * b * c * d / (a * b * c + a * b * d + a * c * d + b * c * d + epsl);
}

#define EPSL 1.1920928955078125e-07f  // epsilon(1.0)

struct KArgs {
    Grid g;
    FieldPtrs f;
    MediumPtrs m;
    float *R;
    ViscoCoef v;
    Coef4 c;
    float dt;
    int kchunk, nbx, nby, nbz, by;
};

// ---------------------------------------------------------------------------
// kernel__update_stress (m_kernel.f90:142-348): normal and shear parts fused
// (they read V and write disjoint arrays).
// ---------------------------------------------------------------------------
template <int NM>
__global__ void __launch_bounds__(256) k_stress_interior(KArgs a)
{
    const Grid &g = a.g;
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int bx = bid % a.nbx;
    const int t = bid / a.nbx;
    const int byy = t % a.nby;
    const int bz = t / a.nby;
    const int i = g.ib + bx * BX + threadIdx.x;
    const int j = g.jb + byy * a.by + threadIdx.y;
    if (i > g.ie || j > g.je) return;
    const int k0 = g.kb + bz * a.kchunk;
    const int k1 = min(g.ke, k0 + a.kchunk - 1);

    const long long c2 = g2idx(g, i, j);
    const int kft = a.m.kfs_top[c2], kfb = a.m.kfs_bot[c2];
    const int kot = a.m.kob_top[c2], kobb = a.m.kob_bot[c2];

    const long long P = g.plane;
    const int Y = g.pitch;
    const float *__restrict__ Vx = a.f.Vx;
    const float *__restrict__ Vy = a.f.Vy;
    const float *__restrict__ Vz = a.f.Vz;
    const float *__restrict__ mu = a.m.mu;
    const float *__restrict__ lam = a.m.lam;
    const float *__restrict__ taup = a.m.taup;
    const float *__restrict__ taus = a.m.taus;
    float *__restrict__ Sxx = a.f.Sxx;
    float *__restrict__ Syy = a.f.Syy;
    float *__restrict__ Szz = a.f.Szz;
    float *__restrict__ Syz = a.f.Syz;
    float *__restrict__ Sxz = a.f.Sxz;
    float *__restrict__ Sxy = a.f.Sxy;
    const float dt = a.dt;
    const float d2 = a.v.d2;

    long long p = gidx(g, i, j, k0);
    float vx_m1 = Vx[p - P], vx_0 = Vx[p], vx_p1 = Vx[p + P];
    float vy_m1 = Vy[p - P], vy_0 = Vy[p], vy_p1 = Vy[p + P];
    float vz_m2 = Vz[p - 2 * P], vz_m1 = Vz[p - P], vz_0 = Vz[p];
    float mu00 = mu[p], mu10 = mu[p + 1], mu01 = mu[p + Y], mu11 = mu[p + 1 + Y];

    long long r = ((long long)(k0 - g.kb) * g.nyk + (j - g.jb)) * g.rpitch + (i - g.ib);
    const long long rn = g.rn;

    for (int k = k0; k <= k1; ++k, p += P, r += g.rplane) {
        const float vx_p2 = Vx[p + 2 * P], vy_p2 = Vy[p + 2 * P], vz_p1 = Vz[p + P];
        const float mu00n = mu[p + P], mu10n = mu[p + P + 1], mu01n = mu[p + P + Y];

        const bool in = in_band(k, kft, kfb, kot, kobb);
        const float re40x = in ? a.c.re40x[1] : a.c.re40x[0];
        const float re41x = in ? a.c.re41x[1] : a.c.re41x[0];
        const float re40y = in ? a.c.re40y[1] : a.c.re40y[0];
        const float re41y = in ? a.c.re41y[1] : a.c.re41y[0];
        const float re40z = in ? a.c.re40z[1] : a.c.re40z[0];
        const float re41z = in ? a.c.re41z[1] : a.c.re41z[0];

        const float vx_im2 = Vx[p - 2], vx_im1 = Vx[p - 1], vx_ip1 = Vx[p + 1];
        const float vx_jm1 = Vx[p - Y], vx_jp1 = Vx[p + Y], vx_jp2 = Vx[p + 2 * Y];
        const float vy_jm2 = Vy[p - 2 * Y], vy_jm1 = Vy[p - Y], vy_jp1 = Vy[p + Y];
        const float vy_im1 = Vy[p - 1], vy_ip1 = Vy[p + 1], vy_ip2 = Vy[p + 2];
        const float vz_im1 = Vz[p - 1], vz_ip1 = Vz[p + 1], vz_ip2 = Vz[p + 2];
        const float vz_jm1 = Vz[p - Y], vz_jp1 = Vz[p + Y], vz_jp2 = Vz[p + 2 * Y];

        // ---- normal components, m_kernel.f90:198-238 ----
        const float dxVx = (vx_0 - vx_im1) * re40x - (vx_ip1 - vx_im2) * re41x;
        const float dyVy = (vy_0 - vy_jm1) * re40y - (vy_jp1 - vy_jm2) * re41y;
        const float dzVz = (vz_0 - vz_m1) * re40z - (vz_p1 - vz_m2) * re41z;

        const float mu2 = 2.0f * mu00;
        const float lam2mu = lam[p] + mu2;
        const float taup1 = taup[p];
        const float taus1 = taus[p];

        const float d3v3 = dxVx + dyVy + dzVz;
        const float dyVy_dzVz = dyVy + dzVz;
        const float dxVx_dzVz = dxVx + dzVz;
        const float dxVx_dyVy = dxVx + dyVy;

        float Rxx_n = 0.0f, Ryy_n = 0.0f, Rzz_n = 0.0f;
#pragma unroll
        for (int m = 0; m < NM; ++m) {
            float *rxx = a.R + (0 * NM + m) * rn + r;
            float *ryy = a.R + (1 * NM + m) * rn + r;
            float *rzz = a.R + (2 * NM + m) * rn + r;
            const float c1 = a.v.c1[m], c2 = a.v.c2[m], d1 = a.v.d1[m];
            const float nxx = c1 * *rxx - c2 * (lam2mu * taup1 * d3v3 - mu2 * taus1 * dyVy_dzVz) * dt;
            const float nyy = c1 * *ryy - c2 * (lam2mu * taup1 * d3v3 - mu2 * taus1 * dxVx_dzVz) * dt;
            const float nzz = c1 * *rzz - c2 * (lam2mu * taup1 * d3v3 - mu2 * taus1 * dxVx_dyVy) * dt;
            *rxx = nxx;
            *ryy = nyy;
            *rzz = nzz;
            Rxx_n = Rxx_n + d1 * nxx;
            Ryy_n = Ryy_n + d1 * nyy;
            Rzz_n = Rzz_n + d1 * nzz;
        }
        const float taup_plus1 = 1.0f + taup1 * (1.0f + d2);
        const float taus_plus1 = 1.0f + taus1 * (1.0f + d2);
        Sxx[p] = Sxx[p] + (lam2mu * taup_plus1 * d3v3 - mu2 * taus_plus1 * dyVy_dzVz + Rxx_n) * dt;
        Syy[p] = Syy[p] + (lam2mu * taup_plus1 * d3v3 - mu2 * taus_plus1 * dxVx_dzVz + Ryy_n) * dt;
        Szz[p] = Szz[p] + (lam2mu * taup_plus1 * d3v3 - mu2 * taus_plus1 * dxVx_dyVy + Rzz_n) * dt;

        // ---- shear components, m_kernel.f90:285-333 ----
        const float dxVy_dyVx = (vy_ip1 - vy_0) * re40x - (vy_ip2 - vy_im1) * re41x
                              + (vx_jp1 - vx_0) * re40y - (vx_jp2 - vx_jm1) * re41y;
        const float dxVz_dzVx = (vz_ip1 - vz_0) * re40x - (vz_ip2 - vz_im1) * re41x
                              + (vx_p1 - vx_0) * re40z - (vx_p2 - vx_m1) * re41z;
        const float dyVz_dzVy = (vz_jp1 - vz_0) * re40y - (vz_jp2 - vz_jm1) * re41y
                              + (vy_p1 - vy_0) * re40z - (vy_p2 - vy_m1) * re41z;

        const float muxz = hmean4(mu00, mu00n, mu10, mu10n, EPSL);
        const float muxy = hmean4(mu00, mu10, mu01, mu11, EPSL);
        const float muyz = hmean4(mu00, mu00n, mu01, mu01n, EPSL);

        float Ryz_n = 0.0f, Rxz_n = 0.0f, Rxy_n = 0.0f;
#pragma unroll
        for (int m = 0; m < NM; ++m) {
            float *ryz = a.R + (3 * NM + m) * rn + r;
            float *rxz = a.R + (4 * NM + m) * rn + r;
            float *rxy = a.R + (5 * NM + m) * rn + r;
            const float c1 = a.v.c1[m], c2 = a.v.c2[m], d1 = a.v.d1[m];
            const float nyz = c1 * *ryz - c2 * muyz * taus1 * dyVz_dzVy * dt;
            const float nxz = c1 * *rxz - c2 * muxz * taus1 * dxVz_dzVx * dt;
            const float nxy = c1 * *rxy - c2 * muxy * taus1 * dxVy_dyVx * dt;
            *ryz = nyz;
            *rxz = nxz;
            *rxy = nxy;
            Ryz_n = Ryz_n + d1 * nyz;
            Rxz_n = Rxz_n + d1 * nxz;
            Rxy_n = Rxy_n + d1 * nxy;
        }
        Syz[p] = Syz[p] + (muyz * taus_plus1 * dyVz_dzVy + Ryz_n) * dt;
        Sxz[p] = Sxz[p] + (muxz * taus_plus1 * dxVz_dzVx + Rxz_n) * dt;
        Sxy[p] = Sxy[p] + (muxy * taus_plus1 * dxVy_dyVx + Rxy_n) * dt;

        // rotate the z queues
        vx_m1 = vx_0; vx_0 = vx_p1; vx_p1 = vx_p2;
        vy_m1 = vy_0; vy_0 = vy_p1; vy_p1 = vy_p2;
        vz_m2 = vz_m1; vz_m1 = vz_0; vz_0 = vz_p1;
        mu00 = mu00n; mu10 = mu10n; mu01 = mu01n;
        mu11 = mu[p + P + 1 + Y];
    }
}

// ---------------------------------------------------------------------------
// kernel__update_vel (m_kernel.f90:75-140)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_vel_interior(KArgs a)
{
    const Grid &g = a.g;
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int bx = bid % a.nbx;
    const int t = bid / a.nbx;
    const int byy = t % a.nby;
    const int bz = t / a.nby;
    const int i = g.ib + bx * BX + threadIdx.x;
    const int j = g.jb + byy * a.by + threadIdx.y;
    if (i > g.ie || j > g.je) return;
    const int k0 = g.kb + bz * a.kchunk;
    const int k1 = min(g.ke, k0 + a.kchunk - 1);

    const long long c2 = g2idx(g, i, j);
    const int kft = a.m.kfs_top[c2], kfb = a.m.kfs_bot[c2];
    const int kot = a.m.kob_top[c2], kobb = a.m.kob_bot[c2];

    const long long P = g.plane;
    const int Y = g.pitch;
    const float *__restrict__ Sxx = a.f.Sxx;
    const float *__restrict__ Syy = a.f.Syy;
    const float *__restrict__ Szz = a.f.Szz;
    const float *__restrict__ Syz = a.f.Syz;
    const float *__restrict__ Sxz = a.f.Sxz;
    const float *__restrict__ Sxy = a.f.Sxy;
    const float *__restrict__ rho = a.m.rho;
    float *__restrict__ Vx = a.f.Vx;
    float *__restrict__ Vy = a.f.Vy;
    float *__restrict__ Vz = a.f.Vz;
    const float dt = a.dt;

    long long p = gidx(g, i, j, k0);
    float sxz_m2 = Sxz[p - 2 * P], sxz_m1 = Sxz[p - P], sxz_0 = Sxz[p];
    float syz_m2 = Syz[p - 2 * P], syz_m1 = Syz[p - P], syz_0 = Syz[p];
    float szz_m1 = Szz[p - P], szz_0 = Szz[p], szz_p1 = Szz[p + P];
    float rho0 = rho[p];

    for (int k = k0; k <= k1; ++k, p += P) {
        const float sxz_p1 = Sxz[p + P], syz_p1 = Syz[p + P], szz_p2 = Szz[p + 2 * P];
        const float rho_kp = rho[p + P];

        const bool in = in_band(k, kft, kfb, kot, kobb);
        const float re40x = in ? a.c.re40x[1] : a.c.re40x[0];
        const float re41x = in ? a.c.re41x[1] : a.c.re41x[0];
        const float re40y = in ? a.c.re40y[1] : a.c.re40y[0];
        const float re41y = in ? a.c.re41y[1] : a.c.re41y[0];
        const float re40z = in ? a.c.re40z[1] : a.c.re40z[0];
        const float re41z = in ? a.c.re41z[1] : a.c.re41z[0];

        const float sxx_im1 = Sxx[p - 1], sxx_0 = Sxx[p], sxx_ip1 = Sxx[p + 1], sxx_ip2 = Sxx[p + 2];
        const float sxy_0 = Sxy[p];
        const float sxy_jm2 = Sxy[p - 2 * Y], sxy_jm1 = Sxy[p - Y], sxy_jp1 = Sxy[p + Y];
        const float sxy_im2 = Sxy[p - 2], sxy_im1 = Sxy[p - 1], sxy_ip1 = Sxy[p + 1];
        const float syy_jm1 = Syy[p - Y], syy_0 = Syy[p], syy_jp1 = Syy[p + Y], syy_jp2 = Syy[p + 2 * Y];
        const float syz_jm2 = Syz[p - 2 * Y], syz_jm1 = Syz[p - Y], syz_jp1 = Syz[p + Y];
        const float sxz_im2 = Sxz[p - 2], sxz_im1 = Sxz[p - 1], sxz_ip1 = Sxz[p + 1];

        const float d3Sx3 = (sxx_ip1 - sxx_0) * re40x - (sxx_ip2 - sxx_im1) * re41x
                          + (sxy_0 - sxy_jm1) * re40y - (sxy_jp1 - sxy_jm2) * re41y
                          + (sxz_0 - sxz_m1) * re40z - (sxz_p1 - sxz_m2) * re41z;
        const float d3Sy3 = (sxy_0 - sxy_im1) * re40x - (sxy_ip1 - sxy_im2) * re41x
                          + (syy_jp1 - syy_0) * re40y - (syy_jp2 - syy_jm1) * re41y
                          + (syz_0 - syz_m1) * re40z - (syz_p1 - syz_m2) * re41z;
        const float d3Sz3 = (sxz_0 - sxz_im1) * re40x - (sxz_ip1 - sxz_im2) * re41x
                          + (syz_0 - syz_jm1) * re40y - (syz_jp1 - syz_jm2) * re41y
                          + (szz_p1 - szz_0) * re40z - (szz_p2 - szz_m1) * re41z;

        Vx[p] = Vx[p] + 2.0f / (rho0 + rho[p + 1]) * d3Sx3 * dt;
        Vy[p] = Vy[p] + 2.0f / (rho0 + rho[p + Y]) * d3Sy3 * dt;
        Vz[p] = Vz[p] + 2.0f / (rho0 + rho_kp) * d3Sz3 * dt;

        sxz_m2 = sxz_m1; sxz_m1 = sxz_0; sxz_0 = sxz_p1;
        syz_m2 = syz_m1; syz_m1 = syz_0; syz_0 = syz_p1;
        szz_m1 = szz_0; szz_0 = szz_p1; szz_p1 = szz_p2;
        rho0 = rho_kp;
    }
}

static void kargs_fill(swpc_ctx *c, KArgs &a)
{
    a.g = c->g;
    a.f.Vx = c->fields[SWPC_VX]; a.f.Vy = c->fields[SWPC_VY]; a.f.Vz = c->fields[SWPC_VZ];
    a.f.Sxx = c->fields[SWPC_SXX]; a.f.Syy = c->fields[SWPC_SYY]; a.f.Szz = c->fields[SWPC_SZZ];
    a.f.Syz = c->fields[SWPC_SYZ]; a.f.Sxz = c->fields[SWPC_SXZ]; a.f.Sxy = c->fields[SWPC_SXY];
    a.m.rho = c->rho; a.m.lam = c->lam; a.m.mu = c->mu; a.m.taup = c->taup; a.m.taus = c->taus;
    a.m.kfs_top = c->kfs_top; a.m.kfs_bot = c->kfs_bot; a.m.kob_top = c->kob_top; a.m.kob_bot = c->kob_bot;
    a.m.kob = c->kob;
    a.R = c->R;
    a.v = c->visco;
    a.c = c->coef;
    a.dt = c->desc.dt;
    a.by = c->by;
    a.kchunk = c->kchunk;
    a.nbx = (c->g.nxk + BX - 1) / BX;
    a.nby = (c->g.nyk + a.by - 1) / a.by;
    a.nbz = (c->g.nzk + a.kchunk - 1) / a.kchunk;
}

hipError_t launch_kernel_stress(swpc_ctx *c)
{
    if (c->g.nxk <= 0 || c->g.nyk <= 0 || c->g.nzk <= 0) return hipSuccess;
    KArgs a;
    kargs_fill(c, a);
    dim3 block(BX, a.by);
    dim3 grid(a.nbx * a.nby * a.nbz);
    switch (c->g.nm) {
    case 0: hipLaunchKernelGGL(k_stress_interior<0>, grid, block, 0, c->stream, a); break;
    case 1: hipLaunchKernelGGL(k_stress_interior<1>, grid, block, 0, c->stream, a); break;
    case 2: hipLaunchKernelGGL(k_stress_interior<2>, grid, block, 0, c->stream, a); break;
    case 3: hipLaunchKernelGGL(k_stress_interior<3>, grid, block, 0, c->stream, a); break;
    case 4: hipLaunchKernelGGL(k_stress_interior<4>, grid, block, 0, c->stream, a); break;
    case 5: hipLaunchKernelGGL(k_stress_interior<5>, grid, block, 0, c->stream, a); break;
    case 6: hipLaunchKernelGGL(k_stress_interior<6>, grid, block, 0, c->stream, a); break;
    case 7: hipLaunchKernelGGL(k_stress_interior<7>, grid, block, 0, c->stream, a); break;
    case 8: hipLaunchKernelGGL(k_stress_interior<8>, grid, block, 0, c->stream, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_kernel_vel(swpc_ctx *c)
{
    if (c->g.nxk <= 0 || c->g.nyk <= 0 || c->g.nzk <= 0) return hipSuccess;
    KArgs a;
    kargs_fill(c, a);
    dim3 block(BX, a.by);
    dim3 grid(a.nbx * a.nby * a.nbz);
    hipLaunchKernelGGL(k_vel_interior, grid, block, 0, c->stream, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// ADE-CFS PML (m_absorb_p.f90)
// ---------------------------------------------------------------------------
__device__ static inline bool pml_cell(const PmlArgs &a, long long t, int &i, int &j, int &k)
{
    if (t >= a.npml) return false;
    PmlBox b = a.box[0];
#pragma unroll
    for (int q = 1; q < 5; ++q)
        if (t >= a.box[q].off) b = a.box[q];
    const long long l = t - b.off;
    const long long plane = (long long)b.nxb * b.nyb;
    const int kk = (int)(l / plane);
    const int rem = (int)(l - (long long)kk * plane);
    const int jj = rem / b.nxb;
    const int ii = rem - jj * b.nxb;
    i = b.i0 + ii;
    j = b.j0 + jj;
    k = b.k0 + kk;
    return true;
}

// absorb_p__update_vel, m_absorb_p.f90:261-306
__global__ void __launch_bounds__(256) k_pml_vel(PmlArgs a)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (!pml_cell(a, t, i, j, k)) return;
    const Grid &g = a.g;
    const long long P = g.plane;
    const int Y = g.pitch;
    const long long p = gidx(g, i, j, k);
    const float *Sxx = a.f.Sxx, *Syy = a.f.Syy, *Szz = a.f.Szz, *Syz = a.f.Syz, *Sxz = a.f.Sxz, *Sxy = a.f.Sxy;
    const float *rho = a.m.rho;
    const float dt = a.dt;
    const long long n = a.npml;

    const float dxSxx = (Sxx[p + 1] - Sxx[p]) * a.r20x;
    const float dySyy = (Syy[p + Y] - Syy[p]) * a.r20y;
    const float dzSzz = (Szz[p + P] - Szz[p]) * a.r20z;
    const float dySyz = (Syz[p] - Syz[p - Y]) * a.r20y;
    const float dzSyz = (Syz[p] - Syz[p - P]) * a.r20z;
    const float dxSxz = (Sxz[p] - Sxz[p - 1]) * a.r20x;
    const float dzSxz = (Sxz[p] - Sxz[p - P]) * a.r20z;
    const float dxSxy = (Sxy[p] - Sxy[p - 1]) * a.r20x;
    const float dySxy = (Sxy[p] - Sxy[p - Y]) * a.r20y;

    const float r0 = rho[p];
    const float bx = 2.0f / (r0 + rho[p + 1]);
    const float by = 2.0f / (r0 + rho[p + Y]);
    const float bz = 2.0f / (r0 + rho[p + P]);

    const float4 gxc = a.gxc[i - 1], gxe = a.gxe[i - 1];
    const float4 gyc = a.gyc[j - 1], gye = a.gye[j - 1];
    const float4 gzc = a.gzc[k - 1], gze = a.gze[k - 1];

    float *ade = a.ade;
    float axSxx = ade[A_axSxx * n + t], aySxy = ade[A_aySxy * n + t], azSxz = ade[A_azSxz * n + t];
    float axSxy = ade[A_axSxy * n + t], aySyy = ade[A_aySyy * n + t], azSyz = ade[A_azSyz * n + t];
    float axSxz = ade[A_axSxz * n + t], aySyz = ade[A_aySyz * n + t], azSzz = ade[A_azSzz * n + t];

    a.f.Vx[p] = a.f.Vx[p] + bx * (gxe.x * dxSxx + gyc.x * dySxy + gzc.x * dzSxz
                                  + gxe.y * axSxx + gyc.y * aySxy + gzc.y * azSxz) * dt;
    a.f.Vy[p] = a.f.Vy[p] + by * (gxc.x * dxSxy + gye.x * dySyy + gzc.x * dzSyz
                                  + gxc.y * axSxy + gye.y * aySyy + gzc.y * azSyz) * dt;
    a.f.Vz[p] = a.f.Vz[p] + bz * (gxc.x * dxSxz + gyc.x * dySyz + gze.x * dzSzz
                                  + gxc.y * axSxz + gyc.y * aySyz + gze.y * azSzz) * dt;

    ade[A_axSxx * n + t] = gxe.z * axSxx + gxe.w * dxSxx * dt;
    ade[A_aySxy * n + t] = gyc.z * aySxy + gyc.w * dySxy * dt;
    ade[A_azSxz * n + t] = gzc.z * azSxz + gzc.w * dzSxz * dt;
    ade[A_axSxy * n + t] = gxc.z * axSxy + gxc.w * dxSxy * dt;
    ade[A_aySyy * n + t] = gye.z * aySyy + gye.w * dySyy * dt;
    ade[A_azSyz * n + t] = gzc.z * azSyz + gzc.w * dzSyz * dt;
    ade[A_axSxz * n + t] = gxc.z * axSxz + gxc.w * dxSxz * dt;
    ade[A_aySyz * n + t] = gyc.z * aySyz + gyc.w * dySyz * dt;
    ade[A_azSzz * n + t] = gze.z * azSzz + gze.w * dzSzz * dt;
}

// absorb_p__update_stress, m_absorb_p.f90:448-519 (both k loops of a column
// touch disjoint arrays, so one thread does both for its cell)
__global__ void __launch_bounds__(256) k_pml_stress(PmlArgs a)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (!pml_cell(a, t, i, j, k)) return;
    const Grid &g = a.g;
    const long long P = g.plane;
    const int Y = g.pitch;
    const long long p = gidx(g, i, j, k);
    const float *Vx = a.f.Vx, *Vy = a.f.Vy, *Vz = a.f.Vz;
    const float *mu = a.m.mu;
    const float dt = a.dt;
    const long long n = a.npml;
    float *ade = a.ade;

    const float4 gxc = a.gxc[i - 1], gxe = a.gxe[i - 1];
    const float4 gyc = a.gyc[j - 1], gye = a.gye[j - 1];
    const float4 gzc = a.gzc[k - 1], gze = a.gze[k - 1];

    // normal components
    {
        const float vx0 = Vx[p], vy0 = Vy[p], vz0 = Vz[p];
        const float dxVx = (vx0 - Vx[p - 1]) * a.r20x;
        const float dyVy = (vy0 - Vy[p - Y]) * a.r20y;
        const float dzVz = (vz0 - Vz[p - P]) * a.r20z;
        const float mu0 = mu[p];
        const float lam2mu_R = (a.m.lam[p] + 2.0f * mu0);
        const float lam_R = lam2mu_R - 2.0f * mu0;
        const float axVx = ade[A_axVx * n + t], ayVy = ade[A_ayVy * n + t], azVz = ade[A_azVz * n + t];
        const float dxVx_ade = gxc.x * dxVx + gxc.y * axVx;
        const float dyVy_ade = gyc.x * dyVy + gyc.y * ayVy;
        const float dzVz_ade = gzc.x * dzVz + gzc.y * azVz;
        a.f.Sxx[p] = a.f.Sxx[p] + (lam2mu_R * dxVx_ade + lam_R * (dyVy_ade + dzVz_ade)) * dt;
        a.f.Syy[p] = a.f.Syy[p] + (lam2mu_R * dyVy_ade + lam_R * (dxVx_ade + dzVz_ade)) * dt;
        a.f.Szz[p] = a.f.Szz[p] + (lam2mu_R * dzVz_ade + lam_R * (dxVx_ade + dyVy_ade)) * dt;
        ade[A_axVx * n + t] = gxc.z * axVx + gxc.w * dxVx * dt;
        ade[A_ayVy * n + t] = gyc.z * ayVy + gyc.w * dyVy * dt;
        ade[A_azVz * n + t] = gzc.z * azVz + gzc.w * dzVz * dt;
    }
    // shear components
    {
        const float vx0 = Vx[p], vy0 = Vy[p], vz0 = Vz[p];
        const float dxVy = (Vy[p + 1] - vy0) * a.r20x;
        const float dxVz = (Vz[p + 1] - vz0) * a.r20x;
        const float dyVx = (Vx[p + Y] - vx0) * a.r20y;
        const float dyVz = (Vz[p + Y] - vz0) * a.r20y;
        const float dzVx = (Vx[p + P] - vx0) * a.r20z;
        const float dzVy = (Vy[p + P] - vy0) * a.r20z;
        const float m000 = mu[p], m100 = mu[p + P], m010 = mu[p + 1], m110 = mu[p + P + 1];
        const float m001 = mu[p + Y], m101 = mu[p + P + Y], m011 = mu[p + 1 + Y];
        const float muxz = hmean4(m000, m100, m010, m110, EPSL);
        const float muxy = hmean4(m000, m010, m001, m011, EPSL);
        const float muyz = hmean4(m000, m100, m001, m101, EPSL);
        const float ayVx = ade[A_ayVx * n + t], azVx = ade[A_azVx * n + t];
        const float axVy = ade[A_axVy * n + t], azVy = ade[A_azVy * n + t];
        const float axVz = ade[A_axVz * n + t], ayVz = ade[A_ayVz * n + t];
        a.f.Syz[p] = a.f.Syz[p] + muyz * (gye.x * dyVz + gze.x * dzVy + gye.y * ayVz + gze.y * azVy) * dt;
        a.f.Sxz[p] = a.f.Sxz[p] + muxz * (gxe.x * dxVz + gze.x * dzVx + gxe.y * axVz + gze.y * azVx) * dt;
        a.f.Sxy[p] = a.f.Sxy[p] + muxy * (gxe.x * dxVy + gye.x * dyVx + gxe.y * axVy + gye.y * ayVx) * dt;
        ade[A_ayVx * n + t] = gye.z * ayVx + gye.w * dyVx * dt;
        ade[A_azVx * n + t] = gze.z * azVx + gze.w * dzVx * dt;
        ade[A_axVy * n + t] = gxe.z * axVy + gxe.w * dxVy * dt;
        ade[A_azVy * n + t] = gze.z * azVy + gze.w * dzVy * dt;
        ade[A_axVz * n + t] = gxe.z * axVz + gxe.w * dxVz * dt;
        ade[A_ayVz * n + t] = gye.z * ayVz + gye.w * dyVz * dt;
    }
}

static void pmlargs_fill(swpc_ctx *c, PmlArgs &a)
{
    a.g = c->g;
    a.f.Vx = c->fields[SWPC_VX]; a.f.Vy = c->fields[SWPC_VY]; a.f.Vz = c->fields[SWPC_VZ];
    a.f.Sxx = c->fields[SWPC_SXX]; a.f.Syy = c->fields[SWPC_SYY]; a.f.Szz = c->fields[SWPC_SZZ];
    a.f.Syz = c->fields[SWPC_SYZ]; a.f.Sxz = c->fields[SWPC_SXZ]; a.f.Sxy = c->fields[SWPC_SXY];
    a.m.rho = c->rho; a.m.lam = c->lam; a.m.mu = c->mu; a.m.taup = c->taup; a.m.taus = c->taus;
    a.m.kfs_top = c->kfs_top; a.m.kfs_bot = c->kfs_bot; a.m.kob_top = c->kob_top; a.m.kob_bot = c->kob_bot;
    a.m.kob = c->kob;
    a.ade = c->ade;
    a.npml = c->npml;
    for (int q = 0; q < 5; ++q) a.box[q] = c->box[q];
    a.gxc = c->gxc; a.gxe = c->gxe; a.gyc = c->gyc; a.gye = c->gye; a.gzc = c->gzc; a.gze = c->gze;
    a.r20x = 1.0f / (float)c->desc.dx;
    a.r20y = 1.0f / (float)c->desc.dy;
    a.r20z = 1.0f / (float)c->desc.dz;
    a.dt = c->desc.dt;
}

hipError_t launch_pml_vel(swpc_ctx *c)
{
    if (c->npml <= 0) return hipSuccess;
    PmlArgs a;
    pmlargs_fill(c, a);
    const long long nb = (c->npml + 255) / 256;
    hipLaunchKernelGGL(k_pml_vel, dim3((unsigned)nb), dim3(256), 0, c->stream, a);
    return hipGetLastError();
}

hipError_t launch_pml_stress(swpc_ctx *c)
{
    if (c->npml <= 0) return hipSuccess;
    PmlArgs a;
    pmlargs_fill(c, a);
    const long long nb = (c->npml + 255) / 256;
    hipLaunchKernelGGL(k_pml_stress, dim3((unsigned)nb), dim3(256), 0, c->stream, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Cerjan sponge (m_absorb_c.f90:113-198).  Profiles carry the 3-cell halo:
// entry of cell n is at n + 2.
// ---------------------------------------------------------------------------
struct CerjanArgs {
    Grid g;
    FieldPtrs f;
    const float *gx_c, *gx_b, *gy_c, *gy_b, *gz_c, *gz_b;
    int kend;
};

__global__ void __launch_bounds__(256) k_cerjan_stress(CerjanArgs a)
{
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 1 + blockIdx.y;
    const int k = 1 + blockIdx.z;
    if (i > a.g.nx || k > a.kend) return;
    const long long p = gidx(a.g, i, j, k);
    const float gxc = a.gx_c[i + 2], gxb = a.gx_b[i + 2];
    const float gyc = a.gy_c[j + 2], gyb = a.gy_b[j + 2];
    const float gzc = a.gz_c[k + 2], gzb = a.gz_b[k + 2];
    const float gcc = gxc * gyc * gzc;
    a.f.Sxx[p] = a.f.Sxx[p] * gcc;
    a.f.Syy[p] = a.f.Syy[p] * gcc;
    a.f.Szz[p] = a.f.Szz[p] * gcc;
    a.f.Syz[p] = a.f.Syz[p] * gxc * gyb * gzb;
    a.f.Sxz[p] = a.f.Sxz[p] * gxb * gyc * gzb;
    a.f.Sxy[p] = a.f.Sxy[p] * gxb * gyb * gzc;
}

__global__ void __launch_bounds__(256) k_cerjan_vel(CerjanArgs a)
{
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 1 + blockIdx.y;
    const int k = 1 + blockIdx.z;
    if (i > a.g.nx || k > a.kend) return;
    const long long p = gidx(a.g, i, j, k);
    const float gxc = a.gx_c[i + 2], gxb = a.gx_b[i + 2];
    const float gyc = a.gy_c[j + 2], gyb = a.gy_b[j + 2];
    const float gzc = a.gz_c[k + 2], gzb = a.gz_b[k + 2];
    a.f.Vx[p] = a.f.Vx[p] * gxb * gyc * gzc;
    a.f.Vy[p] = a.f.Vy[p] * gxc * gyb * gzc;
    a.f.Vz[p] = a.f.Vz[p] * gxc * gyc * gzb;
}

static void cerjan_fill(swpc_ctx *c, CerjanArgs &a, int kend)
{
    a.g = c->g;
    a.f.Vx = c->fields[SWPC_VX]; a.f.Vy = c->fields[SWPC_VY]; a.f.Vz = c->fields[SWPC_VZ];
    a.f.Sxx = c->fields[SWPC_SXX]; a.f.Syy = c->fields[SWPC_SYY]; a.f.Szz = c->fields[SWPC_SZZ];
    a.f.Syz = c->fields[SWPC_SYZ]; a.f.Sxz = c->fields[SWPC_SXZ]; a.f.Sxy = c->fields[SWPC_SXY];
    a.gx_c = c->cgx_c; a.gx_b = c->cgx_b; a.gy_c = c->cgy_c; a.gy_b = c->cgy_b; a.gz_c = c->cgz_c; a.gz_b = c->cgz_b;
    a.kend = kend;
}

hipError_t launch_cerjan_stress(swpc_ctx *c)
{
    CerjanArgs a;
    cerjan_fill(c, a, c->g.ke);
    dim3 grid((c->g.nx + 255) / 256, c->g.ny, c->g.ke);
    hipLaunchKernelGGL(k_cerjan_stress, grid, dim3(256), 0, c->stream, a);
    return hipGetLastError();
}

hipError_t launch_cerjan_vel(swpc_ctx *c)
{
    CerjanArgs a;
    cerjan_fill(c, a, c->g.nz);
    dim3 grid((c->g.nx + 255) / 256, c->g.ny, c->g.nz);
    hipLaunchKernelGGL(k_cerjan_vel, grid, dim3(256), 0, c->stream, a);
    return hipGetLastError();
}
