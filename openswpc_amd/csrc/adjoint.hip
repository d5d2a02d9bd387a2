// adjoint.hip -- discrete adjoint of the PML time step with its Zener memory
// variables, and the checkpointed gradient driver (swpc_adj_gradient).
//
// The forward step is linear in (V, S, R, PML auxiliaries) for a fixed
// medium; its adjoint applies the transposed updates in reverse order.  The
// memory variables of an interior cell (m_kernel.f90:222-238,319-333),
//     R'_c[m] = c1[m] R_c[m] - c2[m] E_c dt,  S'_c = S_c + (F_c + sum d1[m] R'_c[m]) dt,
// transpose to rho = aR' + d1 dt aS', aR = c1 rho, and -dt sum c2 rho on the
// derivative combination E_c; they need no forward R for the gradients.
// Per phase:
//   1. density kernel (one thread per cell): the adjoint of every spatial
//      derivative the cell's forward update used (9 per phase), the updated
//      auxiliary adjoints, and the cell's gradient terms;
//   2. gather kernel (one thread per field point): each point collects the
//      densities of the cells whose stencils read it, with those cells'
//      coefficients -- the band-switched 4th-order pair in the interior,
//      1/h in the PML.
// Gathers replace scatters, so no atomics and a fixed summation order.
// The forward state the gradients need (V_n and its auxiliaries for the
// stress update, S_{n+1} and the stress auxiliaries for the velocity
// update) is recomputed segment by segment from checkpoints taken every
// `ckp` steps of a first forward pass.
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "swpc_internal.h"

namespace {

struct AdjArgs {
    Grid g;
    Coef4 c;
    float dt, hx, hy, hz;
    const float *rho, *lam, *mu, *taup, *taus;
    const int *kfs_top, *kfs_bot, *kob_top, *kob_bot;
    float *aR;          // memory-variable adjoints (6 components x nm, rn cells each)
    long long rn;
    int nm;
    float c1[SWPC_MAXNM], c2[SWPC_MAXNM], d1[SWPC_MAXNM], d2;
    PmlBox box[5];
    long long npml;
    int has_pml;
    const float4 *gxc, *gxe, *gyc, *gye, *gzc, *gze;
    float *aF[9];       // adjoint fields: Vx Vy Vz Sxx Syy Szz Syz Sxz Sxy
    float *aA;          // adjoint auxiliaries, 18 x npml
    float *D[9];        // per-cell derivative adjoints
    float *q[6];        // per-cell gradient terms gathered in phase 2 (velocity phase: 3, stress phase: 6)
    const unsigned char *cls;  // class of every grid point (k_adj_class)
    float tA[3][4], tB[3][4];  // per axis and class: near (p40 or 1/h) and far (p41) weights
    const float *fF[9]; // forward fields of the step (same order as aF)
    const float *fA;    // forward auxiliaries of the phase, 9 x npml
    float *g_lam, *g_mu, *g_rho;
};

enum { FVX = 0, FVY, FVZ, FSXX, FSYY, FSZZ, FSYZ, FSXZ, FSXY };

__device__ static inline bool in_band(int k, int kft, int kfb, int kot, int kob)
{
    const int a = (k - kft) * (kfb - k);
    const int b = (k - kot) * (kob - k);
    return (a > b ? a : b) >= 0;
}

__device__ static inline bool is_int(const Grid &g, int i, int j, int k)
{
    return i >= g.ib && i <= g.ie && j >= g.jb && j <= g.je && k >= 1 && k <= g.ke;
}

// compact auxiliary index of a PML cell, -1 elsewhere
__device__ static inline long long ade_index(const AdjArgs &a, int i, int j, int k)
{
    const Grid &g = a.g;
    if (!a.has_pml || i < 1 || i > g.nx || j < 1 || j > g.ny || k < 1 || k > g.nz || is_int(g, i, j, k)) return -1;
    const int na = g.na;
    int q;
    if (j <= na) q = 0;
    else if (j > g.ny - na) q = 1;
    else if (i <= na) q = 2;
    else if (i > g.nx - na) q = 3;
    else q = 4;
    const PmlBox b = a.box[q];
    return b.off + ((long long)(k - b.k0) * b.nyb + (j - b.j0)) * b.nxb + (i - b.i0);
}

// 4th-order pair of a cell along an axis (the forward kernel's band switch)
__device__ static inline void pair_at(const AdjArgs &a, int axis, int i, int j, int k, float &p40, float &p41)
{
    const long long c2 = g2idx(a.g, i, j);
    const int in = in_band(k, a.kfs_top[c2], a.kfs_bot[c2], a.kob_top[c2], a.kob_bot[c2]) ? 1 : 0;
    if (axis == 0) { p40 = a.c.re40x[in]; p41 = a.c.re41x[in]; }
    else if (axis == 1) { p40 = a.c.re40y[in]; p41 = a.c.re41y[in]; }
    else { p40 = a.c.re40z[in]; p41 = a.c.re41z[in]; }
}

// 0 = not updated, 1 = interior (4th order), 2 = PML (2nd order)
__device__ static inline int cell_kind(const AdjArgs &a, int i, int j, int k)
{
    const Grid &g = a.g;
    if (i < 1 || i > g.nx || j < 1 || j > g.ny || k < 1 || k > g.nz) return 0;
    if (is_int(g, i, j, k)) return 1;
    return a.has_pml ? 2 : 0;
}

// Class of every grid point, for the gathers: 0 outside the updated cells
// (halo), 1 interior off the free-surface band, 2 interior in the band, 3 PML.
// A cell's weights in a gather depend only on its class and the axis.
__global__ void k_adj_class(AdjArgs a, unsigned char *cls)
{
    const Grid &g = a.g;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < g.n; p += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(p / g.plane) - 2;
        const long long r = p % g.plane;
        const int j = (int)(r / g.pitch) - 2;
        const int i = (int)(r % g.pitch) - g.ox + 1;
        const int kind = cell_kind(a, i, j, k);
        unsigned char c = 0;
        if (kind == 1) {
            const long long c2 = g2idx(g, i, j);
            c = in_band(k, a.kfs_top[c2], a.kfs_bot[c2], a.kob_top[c2], a.kob_bot[c2]) ? 2 : 1;
        } else if (kind == 2) {
            c = 3;
        }
        cls[p] = c;
    }
}

// weight of a class from a per-class table (t[0] = 0 for points outside)
__device__ static inline float wt(const float t[4], int c) { return c == 1 ? t[1] : c == 2 ? t[2] : c == 3 ? t[3] : t[0]; }

// Gathers of one axis (stride s): the densities of the cells whose
// derivative stencils read point p, times those cells' weights (A: the
// near pair p40 or 1/h, B: the far pair p41 or 0).  Points outside the
// updated cells hold density 0 and class 0.  The weights of the five
// points p-2s .. p+2s are read once per axis and shared by its gathers.
struct AxisW {
    float A[5], B[5];
    unsigned pml;   // bit o: point p + (o - 2) s is a PML cell
};
__device__ static inline AxisW axis_w(const unsigned char *cls, long long p, long long s, const float tA[4],
                                      const float tB[4])
{
    AxisW w;
    w.pml = 0;
#pragma unroll
    for (int o = 0; o < 5; ++o) {
        const int c = cls[p + (o - 2) * s];
        w.A[o] = wt(tA, c);
        w.B[o] = wt(tB, c);
        w.pml |= (c == 3 ? 1u : 0u) << o;
    }
    return w;
}
// forward-type derivative (cell c reads c+1, c, c+2, c-1)
__device__ static inline float gat_fwd(const float *D, long long p, long long s, const AxisW &w)
{
    return w.A[1] * D[p - s] - w.A[2] * D[p] - w.B[0] * D[p - 2 * s] + w.B[3] * D[p + s];
}
// backward-type derivative (cell c reads c, c-1, c+1, c-2)
__device__ static inline float gat_bwd(const float *D, long long p, long long s, const AxisW &w)
{
    return w.A[2] * D[p] - w.A[3] * D[p + s] - w.B[1] * D[p - s] + w.B[4] * D[p + 2 * s];
}

// The same gathers for a density that interior cells share with another
// one (the velocity update's three derivatives of one component all carry
// that component's weight; the stress update's two derivatives of one shear
// component carry its weight): the density kernels store Dp for PML cells
// only, and the gather reads point o from Dp where it is a PML cell and from
// Di, which holds the same value, elsewhere.  Points outside the cells hold 0
// in both.
__device__ static inline float ld_sel(const float *Dp, const float *Di, long long p, long long s, int o,
                                      const AxisW &w)
{
    return ((w.pml >> o) & 1u ? Dp : Di)[p + (o - 2) * s];
}
__device__ static inline float gat_fwd_s(const float *Dp, const float *Di, long long p, long long s, const AxisW &w)
{
    return w.A[1] * ld_sel(Dp, Di, p, s, 1, w) - w.A[2] * ld_sel(Dp, Di, p, s, 2, w)
         - w.B[0] * ld_sel(Dp, Di, p, s, 0, w) + w.B[3] * ld_sel(Dp, Di, p, s, 3, w);
}
__device__ static inline float gat_bwd_s(const float *Dp, const float *Di, long long p, long long s, const AxisW &w)
{
    return w.A[2] * ld_sel(Dp, Di, p, s, 2, w) - w.A[3] * ld_sel(Dp, Di, p, s, 3, w)
         - w.B[1] * ld_sel(Dp, Di, p, s, 1, w) + w.B[4] * ld_sel(Dp, Di, p, s, 4, w);
}

#define AT(F, di, dj, dk) (F)[gidx(g, i + (di), j + (dj), k + (dk))]
#define BWD4(F, di, dj, dk, p40, p41) ((AT(F, 0, 0, 0) - AT(F, -(di), -(dj), -(dk))) * (p40) \
                                       - (AT(F, di, dj, dk) - AT(F, -2 * (di), -2 * (dj), -2 * (dk))) * (p41))
#define FWD4(F, di, dj, dk, p40, p41) ((AT(F, di, dj, dk) - AT(F, 0, 0, 0)) * (p40) \
                                       - (AT(F, 2 * (di), 2 * (dj), 2 * (dk)) - AT(F, -(di), -(dj), -(dk))) * (p41))
#define BWD2(F, di, dj, dk, h) ((AT(F, 0, 0, 0) - AT(F, -(di), -(dj), -(dk))) * (h))
#define FWD2(F, di, dj, dk, h) ((AT(F, di, dj, dk) - AT(F, 0, 0, 0)) * (h))

__device__ static inline float4 prof(const float4 *p, int n) { return p[n - 1]; }

// One thread per cell (i, j, k) of the updated box, 64 x 4 cells per block.
// The blocks form a 1D grid, plane after plane, each plane's tiles padded to
// a multiple of 8 blocks: block b goes to XCD b % 8, so a tile and the same
// tile one plane up or down share an XCD and its L2, and each plane still
// spreads over all eight XCDs.  (A thread marching over several planes keeps
// the 64-bit addresses of every array and offset live across its loop and
// halves the occupancy; keeping whole planes on one XCD overflows its L2
// with the z neighbours of nine density arrays.  Both measured slower.
// Giving each XCD a contiguous eighth of a plane's tiles keeps halo rows in
// one L2: 2.79 against 2.84 ms per gradient step at 256 x 256 x 192, but
// 10.04 against 9.93 ms per source-step at 384 x 384 x 256, NM = 3.  With
// the absorber tile rows dealt out one tile per XCD and equal runs of the
// rest, the same: 2.83 against 2.88, 10.08 against 9.89.)
#define ADJ_CELL(a)                                                \
    const Grid &g = (a).g;                                         \
    const int nbx_ = (g.nx + 63) >> 6, nby_ = (g.ny + 3) >> 2;     \
    const int tpp_ = (nbx_ * nby_ + 7) & ~7;                       \
    const int k = 1 + (int)(blockIdx.x / tpp_);                    \
    const int t_ = (int)(blockIdx.x % tpp_);                       \
    if (t_ >= nbx_ * nby_) return;                                 \
    const int i = 1 + (t_ % nbx_) * 64 + threadIdx.x;              \
    const int j = 1 + (t_ / nbx_) * 4 + threadIdx.y;               \
    if (i > g.nx || j > g.ny) return;

// adjoint of the velocity update: densities, stress-auxiliary adjoints, rho terms
// (the stores come after every load, which the compiler cannot reorder)
__global__ void __launch_bounds__(256) k_adj_vel_dens(AdjArgs a)
{
    ADJ_CELL(a) {
        const int kind = cell_kind(a, i, j, k);
        const long long p = gidx(g, i, j, k);
        const float r0 = a.rho[p];
        const float sx = r0 + AT(a.rho, 1, 0, 0), sy = r0 + AT(a.rho, 0, 1, 0), sz = r0 + AT(a.rho, 0, 0, 1);
        const float vx = a.aF[FVX][p], vy = a.aF[FVY][p], vz = a.aF[FVZ][p];
        const float wx = 2.0f / sx * a.dt * vx, wy = 2.0f / sy * a.dt * vy, wz = 2.0f / sz * a.dt * vz;
        const float *Sxx = a.fF[FSXX], *Syy = a.fF[FSYY], *Szz = a.fF[FSZZ];
        const float *Syz = a.fF[FSYZ], *Sxz = a.fF[FSXZ], *Sxy = a.fF[FSXY];
        float fx, fy, fz;
        float d[9];
        if (kind == 1) {
            float ax, bx, ay, by, az, bz;
            pair_at(a, 0, i, j, k, ax, bx);
            pair_at(a, 1, i, j, k, ay, by);
            pair_at(a, 2, i, j, k, az, bz);
            fx = FWD4(Sxx, 1, 0, 0, ax, bx) + BWD4(Sxy, 0, 1, 0, ay, by) + BWD4(Sxz, 0, 0, 1, az, bz);
            fy = BWD4(Sxy, 1, 0, 0, ax, bx) + FWD4(Syy, 0, 1, 0, ay, by) + BWD4(Syz, 0, 0, 1, az, bz);
            fz = BWD4(Sxz, 1, 0, 0, ax, bx) + BWD4(Syz, 0, 1, 0, ay, by) + FWD4(Szz, 0, 0, 1, az, bz);
            d[0] = wx; d[1] = wx; d[2] = wx;
            d[3] = wy; d[4] = wy; d[5] = wy;
            d[6] = wz; d[7] = wz; d[8] = wz;
        } else {
            const long long q = ade_index(a, i, j, k);
            const long long n = a.npml;
            const float4 gp[9] = {prof(a.gxe, i), prof(a.gyc, j), prof(a.gzc, k), prof(a.gxc, i), prof(a.gye, j),
                                  prof(a.gzc, k), prof(a.gxc, i), prof(a.gyc, j), prof(a.gze, k)};
            const float w[9] = {wx, wx, wx, wy, wy, wy, wz, wz, wz};
            const float der[9] = {FWD2(Sxx, 1, 0, 0, a.hx), BWD2(Sxy, 0, 1, 0, a.hy), BWD2(Sxz, 0, 0, 1, a.hz),
                                  BWD2(Sxy, 1, 0, 0, a.hx), FWD2(Syy, 0, 1, 0, a.hy), BWD2(Syz, 0, 0, 1, a.hz),
                                  BWD2(Sxz, 1, 0, 0, a.hx), BWD2(Syz, 0, 1, 0, a.hy), FWD2(Szz, 0, 0, 1, a.hz)};
            float ahat[9], fa[9];
#pragma unroll
            for (int c = 0; c < 9; ++c) {
                ahat[c] = a.aA[(long long)(A_axSxx + c) * n + q];
                fa[c] = a.fA[(long long)c * n + q];
            }
            float f[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int c = 0; c < 9; ++c) {
                d[c] = gp[c].x * w[c] + gp[c].w * a.dt * ahat[c];
                f[c / 3] += gp[c].x * der[c] + gp[c].y * fa[c];
            }
#pragma unroll
            for (int c = 0; c < 9; ++c) a.aA[(long long)(A_axSxx + c) * n + q] = gp[c].y * w[c] + gp[c].z * ahat[c];
            fx = f[0]; fy = f[1]; fz = f[2];
        }
        // interior cells: d[0..2], d[3..5], d[6..8] are one value each, kept in D0, D3, D6
        if (kind == 1) {
            a.D[0][p] = d[0]; a.D[3][p] = d[3]; a.D[6][p] = d[6];
        } else {
#pragma unroll
            for (int c = 0; c < 9; ++c) a.D[c][p] = d[c];
        }
        a.q[0][p] = -2.0f / (sx * sx) * a.dt * fx * vx;
        a.q[1][p] = -2.0f / (sy * sy) * a.dt * fy * vy;
        a.q[2][p] = -2.0f / (sz * sz) * a.dt * fz * vz;
    }
}

// adjoint of the velocity update: stress adjoints and the rho gradient
// (every load before the stores: the class bytes may alias anything)
__global__ void __launch_bounds__(256) k_adj_vel_gather(AdjArgs a)
{
    ADJ_CELL(a) {
        const long long p = gidx(g, i, j, k);
        const long long sx = 1, sy = g.pitch, sz = g.plane;
        const AxisW wx = axis_w(a.cls, p, sx, a.tA[0], a.tB[0]);
        const AxisW wy = axis_w(a.cls, p, sy, a.tA[1], a.tB[1]);
        const AxisW wz = axis_w(a.cls, p, sz, a.tA[2], a.tB[2]);
        const float dxx = gat_fwd(a.D[0], p, sx, wx);
        const float dyy = gat_fwd_s(a.D[4], a.D[3], p, sy, wy);
        const float dzz = gat_fwd_s(a.D[8], a.D[6], p, sz, wz);
        const float dxy = gat_bwd_s(a.D[1], a.D[0], p, sy, wy) + gat_bwd(a.D[3], p, sx, wx);
        const float dxz = gat_bwd_s(a.D[2], a.D[0], p, sz, wz) + gat_bwd(a.D[6], p, sx, wx);
        const float dyz = gat_bwd_s(a.D[5], a.D[3], p, sz, wz) + gat_bwd_s(a.D[7], a.D[6], p, sy, wy);
        // each cell's density enters the buoyancy of its own point and of the
        // point one cell back along each axis (q is 0 outside the cells)
        float gr = a.q[0][p] + a.q[1][p] + a.q[2][p];
        gr += a.q[0][p - sx];
        gr += a.q[1][p - sy];
        gr += a.q[2][p - sz];
        const float oxx = a.aF[FSXX][p], oyy = a.aF[FSYY][p], ozz = a.aF[FSZZ][p];
        const float oxy = a.aF[FSXY][p], oxz = a.aF[FSXZ][p], oyz = a.aF[FSYZ][p], org = a.g_rho[p];
        a.aF[FSXX][p] = oxx + dxx;
        a.aF[FSYY][p] = oyy + dyy;
        a.aF[FSZZ][p] = ozz + dzz;
        a.aF[FSXY][p] = oxy + dxy;
        a.aF[FSXZ][p] = oxz + dxz;
        a.aF[FSYZ][p] = oyz + dyz;
        a.g_rho[p] = org + gr;
    }
}

// harmonic mean of four moduli (the forward's shear modulus, m_kernel.f90:293-309)
__device__ static inline float hmean4(float a, float b, float c, float d)
{
    const float eps = 1.1920928955078125e-07f;
    return 4.0f * a * b * c * d / (a * b * c + a * b * d + a * c * d + b * c * d + eps);
}

// its derivatives with respect to each of the four moduli
__device__ static inline void hmean4_w(float a, float b, float c, float d, float gr[4])
{
    const float eps = 1.1920928955078125e-07f;
    const float N = 4.0f * a * b * c * d, Dn = a * b * c + a * b * d + a * c * d + b * c * d + eps;
    const float r = 1.0f / (Dn * Dn);
    gr[0] = (4.0f * b * c * d * Dn - N * (b * c + b * d + c * d)) * r;
    gr[1] = (4.0f * a * c * d * Dn - N * (a * c + a * d + c * d)) * r;
    gr[2] = (4.0f * a * b * d * Dn - N * (a * b + a * d + b * d)) * r;
    gr[3] = (4.0f * a * b * c * Dn - N * (a * b + a * c + b * c)) * r;
}

// adjoint of the stress update: densities, velocity-auxiliary adjoints,
// lam and mu gradients of the cell, and the shear-mu terms of its
// neighbours.  A shear point's modulus is the harmonic mean of four cells
// (origin o and o + one or two unit steps); the cell adds the derivative
// for its own corner to g_mu and leaves those for the other three corners
// of each of its three quads, summed by target offset, in q[0..5]:
//   q0 -> o+z, q1 -> o+y, q2 -> o+x, q3 -> o+y+z, q4 -> o+x+z, q5 -> o+x+y.
// NM memory variables per component (a template argument, so their loops
// unroll); the stores come after every load.
template <int NM>
__global__ void __launch_bounds__(256) k_adj_stress_dens(AdjArgs a)
{
    ADJ_CELL(a) {
        const int kind = cell_kind(a, i, j, k);
        const long long p = gidx(g, i, j, k);
        const float lam = a.lam[p], mu = a.mu[p];
        const float sxx = a.aF[FSXX][p], syy = a.aF[FSYY][p], szz = a.aF[FSZZ][p];
        const float X = a.dt * ((lam + 2.0f * mu) * sxx + lam * (syy + szz));
        const float Y = a.dt * ((lam + 2.0f * mu) * syy + lam * (sxx + szz));
        const float Z = a.dt * ((lam + 2.0f * mu) * szz + lam * (sxx + syy));
        const float *mup = a.mu;
        const float m001 = AT(mup, 0, 0, 1), m010 = AT(mup, 0, 1, 0), m011 = AT(mup, 0, 1, 1);
        const float m100 = AT(mup, 1, 0, 0), m101 = AT(mup, 1, 0, 1), m110 = AT(mup, 1, 1, 0);
        const float hyz = hmean4(mu, m001, m010, m011);
        const float hxz = hmean4(mu, m001, m100, m101);
        const float hxy = hmean4(mu, m100, m010, m110);
        const float syz = a.aF[FSYZ][p], sxz = a.aF[FSXZ][p], sxy = a.aF[FSXY][p];
        const float Wyz = a.dt * hyz * syz, Wxz = a.dt * hxz * sxz, Wxy = a.dt * hxy * sxy;
        const float *Vx = a.fF[FVX], *Vy = a.fF[FVY], *Vz = a.fF[FVZ];
        const float glo = a.g_lam[p], gmo = a.g_mu[p];
        float ex, ey, ez, tyz, txz, txy;
        float qyz, qxz, qxy;   // shear-mu terms of the three quads
        float d[9], dgl, dgm;
        if (kind == 1) {
            float ax, bx, ay, by, az, bz;
            pair_at(a, 0, i, j, k, ax, bx);
            pair_at(a, 1, i, j, k, ay, by);
            pair_at(a, 2, i, j, k, az, bz);
            ex = BWD4(Vx, 1, 0, 0, ax, bx);
            ey = BWD4(Vy, 0, 1, 0, ay, by);
            ez = BWD4(Vz, 0, 0, 1, az, bz);
            tyz = FWD4(Vz, 0, 1, 0, ay, by) + FWD4(Vy, 0, 0, 1, az, bz);
            txz = FWD4(Vz, 1, 0, 0, ax, bx) + FWD4(Vx, 0, 0, 1, az, bz);
            txy = FWD4(Vy, 1, 0, 0, ax, bx) + FWD4(Vx, 0, 1, 0, ay, by);
            if (NM > 0) {
                // transposed memory variables and weighted stress sums
                const float taup = a.taup[p], taus = a.taus[p];
                const float tp1 = 1.0f + taup * (1.0f + a.d2), ts1 = 1.0f + taus * (1.0f + a.d2);
                const float as6[6] = {sxx, syy, szz, syz, sxz, sxy};
                const long long r = ((long long)(k - g.kb) * g.nyk + (j - g.jb)) * g.rpitch + (i - 1);
                float rh[6][NM > 0 ? NM : 1];
#pragma unroll
                for (int m = 0; m < NM; ++m)
#pragma unroll
                    for (int cc = 0; cc < 6; ++cc) rh[cc][m] = a.aR[((long long)cc * NM + m) * a.rn + r];
                float sig[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int m = 0; m < NM; ++m)
#pragma unroll
                    for (int cc = 0; cc < 6; ++cc) {
                        rh[cc][m] = rh[cc][m] + a.d1[m] * a.dt * as6[cc];
                        sig[cc] -= a.dt * a.c2[m] * rh[cc][m];
                    }
#pragma unroll
                for (int m = 0; m < NM; ++m)
#pragma unroll
                    for (int cc = 0; cc < 6; ++cc) a.aR[((long long)cc * NM + m) * a.rn + r] = a.c1[m] * rh[cc][m];
                float P[3], Q[6];
#pragma unroll
                for (int cc = 0; cc < 6; ++cc) {
                    if (cc < 3) P[cc] = a.dt * as6[cc] * tp1 + sig[cc] * taup;
                    Q[cc] = a.dt * as6[cc] * ts1 + sig[cc] * taus;
                }
                const float sP = P[0] + P[1] + P[2], l2m = lam + 2.0f * mu, d3 = ex + ey + ez;
                d[0] = l2m * sP - 2.0f * mu * (Q[1] + Q[2]);
                d[1] = l2m * sP - 2.0f * mu * (Q[0] + Q[2]);
                d[2] = l2m * sP - 2.0f * mu * (Q[0] + Q[1]);
                d[3] = hyz * Q[3]; d[4] = hyz * Q[3];
                d[5] = hxz * Q[4]; d[6] = hxz * Q[4];
                d[7] = hxy * Q[5]; d[8] = hxy * Q[5];
                dgl = d3 * sP;
                dgm = 2.0f * d3 * sP - 2.0f * ((d3 - ex) * Q[0] + (d3 - ey) * Q[1] + (d3 - ez) * Q[2]);
                qyz = tyz * Q[3];
                qxz = txz * Q[4];
                qxy = txy * Q[5];
            } else {
                d[0] = X; d[1] = Y; d[2] = Z;
                d[3] = Wyz; d[4] = Wyz; d[5] = Wxz; d[6] = Wxz; d[7] = Wxy; d[8] = Wxy;
                dgl = a.dt * (ex + ey + ez) * (sxx + syy + szz);
                dgm = 2.0f * a.dt * (ex * sxx + ey * syy + ez * szz);
                qyz = a.dt * tyz * syz;
                qxz = a.dt * txz * sxz;
                qxy = a.dt * txy * sxy;
            }
        } else {
            const long long q = ade_index(a, i, j, k);
            const long long n = a.npml;
            // derivatives dxVx dyVy dzVz dyVz dzVy dxVz dzVx dxVy dyVx and their auxiliaries
            const float4 gp[9] = {prof(a.gxc, i), prof(a.gyc, j), prof(a.gzc, k), prof(a.gye, j), prof(a.gze, k),
                                  prof(a.gxe, i), prof(a.gze, k), prof(a.gxe, i), prof(a.gye, j)};
            const int ai[9] = {A_axVx, A_ayVy, A_azVz, A_ayVz, A_azVy, A_axVz, A_azVx, A_axVy, A_ayVx};
            const float w[9] = {X, Y, Z, Wyz, Wyz, Wxz, Wxz, Wxy, Wxy};
            const float der[9] = {BWD2(Vx, 1, 0, 0, a.hx), BWD2(Vy, 0, 1, 0, a.hy), BWD2(Vz, 0, 0, 1, a.hz),
                                  FWD2(Vz, 0, 1, 0, a.hy), FWD2(Vy, 0, 0, 1, a.hz), FWD2(Vz, 1, 0, 0, a.hx),
                                  FWD2(Vx, 0, 0, 1, a.hz), FWD2(Vy, 1, 0, 0, a.hx), FWD2(Vx, 0, 1, 0, a.hy)};
            float ahat[9], fa[9], t[9];
#pragma unroll
            for (int c = 0; c < 9; ++c) {
                ahat[c] = a.aA[(long long)ai[c] * n + q];
                fa[c] = a.fA[(long long)ai[c] * n + q];
            }
#pragma unroll
            for (int c = 0; c < 9; ++c) {
                d[c] = gp[c].x * w[c] + gp[c].w * a.dt * ahat[c];
                t[c] = gp[c].x * der[c] + gp[c].y * fa[c];
            }
#pragma unroll
            for (int c = 0; c < 9; ++c) a.aA[(long long)ai[c] * n + q] = gp[c].y * w[c] + gp[c].z * ahat[c];
            ex = t[0]; ey = t[1]; ez = t[2];
            tyz = t[3] + t[4]; txz = t[5] + t[6]; txy = t[7] + t[8];
            dgl = a.dt * (ex + ey + ez) * (sxx + syy + szz);
            dgm = 2.0f * a.dt * (ex * sxx + ey * syy + ez * szz);
            qyz = a.dt * tyz * syz;
            qxz = a.dt * txz * sxz;
            qxy = a.dt * txy * sxy;
        }
        // quads: yz (o, o+z, o+y, o+y+z), xz (o, o+z, o+x, o+x+z), xy (o, o+x, o+y, o+x+y);
        // their weights are taken here, late, to keep registers free above
        float gyz[4], gxz[4], gxy[4];
        hmean4_w(mu, m001, m010, m011, gyz);
        hmean4_w(mu, m001, m100, m101, gxz);
        hmean4_w(mu, m100, m010, m110, gxy);
        // interior cells: d[3] = d[4], d[5] = d[6], d[7] = d[8], kept in D3, D5, D7
        if (kind == 1) {
            a.D[0][p] = d[0]; a.D[1][p] = d[1]; a.D[2][p] = d[2];
            a.D[3][p] = d[3]; a.D[5][p] = d[5]; a.D[7][p] = d[7];
        } else {
#pragma unroll
            for (int c = 0; c < 9; ++c) a.D[c][p] = d[c];
        }
        a.g_lam[p] = glo + dgl;
        a.g_mu[p] = (gmo + dgm) + (qyz * gyz[0] + qxz * gxz[0] + qxy * gxy[0]);
        a.q[0][p] = qyz * gyz[1] + qxz * gxz[1];
        a.q[1][p] = qyz * gyz[2] + qxy * gxy[2];
        a.q[2][p] = qxz * gxz[2] + qxy * gxy[1];
        a.q[3][p] = qyz * gyz[3];
        a.q[4][p] = qxz * gxz[3];
        a.q[5][p] = qxy * gxy[3];
    }
}

// adjoint of the stress update: velocity adjoints and the shear terms of
// the mu gradient left by the neighbouring cells (0 outside the cells)
__global__ void __launch_bounds__(256) k_adj_stress_gather(AdjArgs a)
{
    ADJ_CELL(a) {
        const long long p = gidx(g, i, j, k);
        const long long sx = 1, sy = g.pitch, sz = g.plane;
        const AxisW wx = axis_w(a.cls, p, sx, a.tA[0], a.tB[0]);
        const AxisW wy = axis_w(a.cls, p, sy, a.tA[1], a.tB[1]);
        const AxisW wz = axis_w(a.cls, p, sz, a.tA[2], a.tB[2]);
        const float vx = gat_bwd(a.D[0], p, sx, wx) + gat_fwd_s(a.D[8], a.D[7], p, sy, wy) +
                         gat_fwd_s(a.D[6], a.D[5], p, sz, wz);
        const float vy = gat_bwd(a.D[1], p, sy, wy) + gat_fwd(a.D[7], p, sx, wx) +
                         gat_fwd_s(a.D[4], a.D[3], p, sz, wz);
        const float vz = gat_bwd(a.D[2], p, sz, wz) + gat_fwd(a.D[5], p, sx, wx) + gat_fwd(a.D[3], p, sy, wy);
        const float gm = a.q[0][p - sz] + a.q[1][p - sy] + a.q[2][p - sx] + a.q[3][p - sy - sz] + a.q[4][p - sx - sz] +
                         a.q[5][p - sx - sy];
        const float ox = a.aF[FVX][p], oy = a.aF[FVY][p], oz = a.aF[FVZ][p], om = a.g_mu[p];
        a.aF[FVX][p] = ox + vx;
        a.aF[FVY][p] = oy + vy;
        a.aF[FVZ][p] = oz + vz;
        a.g_mu[p] = om + gm;
    }
}

// adjoint of the velocity traces at one recorded step (m_wav.f90:517-531).
// One thread per distinct target point (field, cell): it adds the residuals
// of the stations that read the point, in station order, so stations that
// share points add deterministically.  tgt[t] = field << 40 | gidx;
// the contributions of target t are src[ptr[t] .. ptr[t+1]), each a
// station-component w = 3 * station + component.
__global__ void k_adj_record(AdjArgs a, int ntgt, const long long *tgt, const int *ptr, const int *src,
                             const float *wav, const float *obs, int ntw, int itw, float sc)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntgt) return;
    const int f = (int)(tgt[t] >> 40);
    float *dst = (f == FVX ? a.aF[FVX] : f == FVY ? a.aF[FVY] : a.aF[FVZ]) + (tgt[t] & ((1LL << 40) - 1));
    float acc = *dst;
    for (int e = ptr[t]; e < ptr[t + 1]; ++e) {
        const long long w = (long long)src[e] * ntw + itw;
        const float r = sc * (wav[w] - obs[w]);
        if (f == FVZ) acc -= r;
        else acc += r;
    }
    *dst = acc;
}

__global__ void k_fill(float *p, long long n)
{
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        p[t] = 0.0f;
}

// sum += part, one rounding per element
__global__ void k_add(float *sum, const float *part, long long n)
{
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        sum[t] = sum[t] + part[t];
}

}  // namespace

// ---------------------------------------------------------------------------
// driver
// ---------------------------------------------------------------------------
// Sizes of one context's adjoint, in floats.
struct AdjSizes {
    long long nf, na18, nR;
    long long state() const { return 9 * nf + na18 + nR; }  // checkpoint: V, S, auxiliaries, R
    long long sstep() const { return 9 * nf + na18; }       // per-step record: V_n, S_{n+1}, aV_n, aS_n
    long long cls() const { return (nf + 3) / 4; }           // point classes, one byte each
    long long base() const { return 30 * nf + cls() + na18 + nR; }  // adjoint fields, densities, q, gradients, sums, classes, aR
};

namespace {

hipError_t zero(float *p, long long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill, dim3(1024), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

AdjSizes sizes_of(const swpc_ctx *c)
{
    AdjSizes z;
    z.nf = c->g.n;
    z.na18 = (long long)A_N * c->npml;
    z.nR = 6LL * c->desc.nm * c->g.rn;
    return z;
}

// Two-level checkpoint scheme: outer checkpoints every c1 steps of the first
// forward pass; each outer segment is recomputed once to place inner
// checkpoints every c2 steps, and each inner segment once more into the
// per-step records the adjoint reads.  c2 == c1 is the single-level scheme.
// The first pass itself places the last outer segment's inner checkpoints and
// records its last inner segment, so that segment is not run again but for
// its other inner segments' records.
long long scheme_fwd_steps(int nt, int c1, int c2)
{
    long long f = nt;  // the recording pass
    for (int n0 = 0; n0 < nt; n0 += c1) {
        const int L = nt - n0 < c1 ? nt - n0 : c1;
        f += L;
        if (c2 < L) f += (long long)((L - 1) / c2) * c2;
    }
    return f - (nt - (long long)((nt - 1) / c1) * c1);   // the last outer segment's length
}

long long scheme_ck_slots(int nt, int c1, int c2)
{
    const long long outer = (nt + c1 - 1) / c1;
    const int L = nt < c1 ? nt : c1;
    return outer + (c2 < L ? (L - 1) / c2 : 0);  // inner segment 0 starts from its outer checkpoint
}

long long scheme_seg_slots(int nt, int c1, int c2)
{
    int m = c2 < c1 ? c2 : c1;
    return m < nt ? m : nt;
}

long long scheme_floats(const AdjSizes &z, int nt, int c1, int c2)
{
    return z.base() + scheme_ck_slots(nt, c1, c2) * z.state() + scheme_seg_slots(nt, c1, c2) * z.sstep();
}

}  // namespace

// Device memory of the adjoint, kept by the context between calls so that
// the gradients of many sources sum on the device: the adjoint state, the
// gradient sums, and checkpoint stores that grow to the largest scheme used.
struct AdjWorkspace {
    AdjSizes z;
    float *adjF = nullptr, *adjA = nullptr, *dens = nullptr, *qq = nullptr, *adjR = nullptr;
    float *gsrc = nullptr;  // one source's gradients, summed over its steps
    float *grad = nullptr;  // the sums over sources
    float *cls = nullptr;   // point classes (bytes)
    float *ck = nullptr, *seg = nullptr, *obs = nullptr;
    // record table for 6 nst targets: tgt (long long) | ptr (6 nst + 1 int) | src (6 nst int)
    float *rec = nullptr;
    long long ck_n = 0, seg_n = 0, obs_n = 0, rec_n = 0;
    long long bytes = 0;
    int nsum = 0;  // sources summed since the last reset
};


namespace {

void wfree(swpc_ctx *c, float *&p, long long n)
{
    if (p) {
        if (!arena_free(p)) (void)hipFree(p);
        c->adj_ws->bytes -= n * (long long)sizeof(float);
        c->bytes -= n * (long long)sizeof(float);
    }
    p = nullptr;
}

// n floats; with a striped arena on the device, in segments of seg floats
// that rotate over its HBM regions (seg = nf: one field per segment, so the
// fields a kernel streams together sit in different regions), else hipMalloc
bool walloc(swpc_ctx *c, float *&p, long long n, long long seg = 0)
{
    p = nullptr;
    if (n <= 0) return true;
    p = (float *)arena_alloc(c->device, c, (size_t)n * sizeof(float), nullptr, 0, (size_t)seg * sizeof(float));
    if (!p && hipMalloc((void **)&p, (size_t)n * sizeof(float)) != hipSuccess) {
        (void)hipGetLastError();
        p = nullptr;
        return false;
    }
    c->adj_ws->bytes += n * (long long)sizeof(float);
    c->bytes += n * (long long)sizeof(float);
    return true;
}

// a store of at least n floats (contents not kept)
bool wgrow(swpc_ctx *c, float *&p, long long &have, long long n, long long seg = 0)
{
    if (have >= n && p) return true;
    wfree(c, p, have);
    have = 0;
    if (!walloc(c, p, n, seg)) return false;
    have = n;
    return true;
}

// the workspace with its fixed part; the sums start at zero
int ws_open(swpc_ctx *c)
{
    if (c->adj_ws) return SWPC_OK;
    c->adj_ws = new AdjWorkspace();
    AdjWorkspace &w = *c->adj_ws;
    w.z = sizes_of(c);
    const long long nf = w.z.nf;
    if (!walloc(c, w.adjF, 9 * nf, nf) || !walloc(c, w.adjA, w.z.na18, c->npml) || !walloc(c, w.dens, 9 * nf, nf) ||
        !walloc(c, w.qq, 6 * nf, nf) || !walloc(c, w.gsrc, 3 * nf, nf) || !walloc(c, w.grad, 3 * nf, nf) ||
        !walloc(c, w.cls, w.z.cls()) || !walloc(c, w.adjR, w.z.nR, c->g.rn)) {
        swpc_adj_release(c);
        return ctx_fail(c, SWPC_ENOMEM, "adjoint: device memory for the adjoint fields");
    }
    // the density kernels write every cell and never the halo, where the
    // gathers then read zeros
    if (zero(w.grad, 3 * nf, c->stream) != hipSuccess || zero(w.dens, 9 * nf, c->stream) != hipSuccess ||
        zero(w.qq, 6 * nf, c->stream) != hipSuccess)
        return SWPC_EHIP;
    return SWPC_OK;
}

int check_ready(swpc_ctx *c)
{
    if (c->desc.abc != SWPC_ABC_PML || c->npml == 0)
        return ctx_fail(c, SWPC_ESTATE, "adjoint: the adjoint covers the PML scheme");
    if (!c->has_medium || !c->has_pml || (c->desc.nm > 0 && !c->has_visco))
        return ctx_fail(c, SWPC_ESTATE, "adjoint: medium, visco coefficients or PML not set");
    if (c->nst <= 0 || !(c->wav_what & SWPC_WAV_V) || !c->d_wav[0])
        return ctx_fail(c, SWPC_ESTATE, "adjoint: velocity receivers not set");
    if (c->snap_ntask > 0) return ctx_fail(c, SWPC_ESTATE, "adjoint: disable sections first");
    return SWPC_OK;
}

}  // namespace

extern "C" int swpc_adj_release(swpc_ctx *c)
{
    if (!c) return SWPC_EINVAL;
    if (!c->adj_ws) return SWPC_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    AdjWorkspace &w = *c->adj_ws;
    const long long nf = w.z.nf;
    wfree(c, w.adjF, 9 * nf); wfree(c, w.adjA, w.z.na18); wfree(c, w.dens, 9 * nf);
    wfree(c, w.qq, 6 * nf); wfree(c, w.gsrc, 3 * nf); wfree(c, w.grad, 3 * nf); wfree(c, w.adjR, w.z.nR);
    wfree(c, w.cls, w.z.cls());
    wfree(c, w.ck, w.ck_n); wfree(c, w.seg, w.seg_n); wfree(c, w.obs, w.obs_n); wfree(c, w.rec, w.rec_n);
    delete c->adj_ws;
    c->adj_ws = nullptr;
    return SWPC_OK;
}

extern "C" int swpc_adj_plan(swpc_ctx *c, int nt, int64_t budget, int *ckp_outer, int *ckp_inner, int64_t *bytes,
                             int64_t *fwd_steps)
{
    if (!c) return SWPC_EINVAL;
    if (nt < 1) return ctx_fail(c, SWPC_EINVAL, "swpc_adj_plan: nt < 1");
    if (budget <= 0) {
        size_t fr = 0, tot = 0;
        if (hipSetDevice(c->device) != hipSuccess || hipMemGetInfo(&fr, &tot) != hipSuccess) return SWPC_EHIP;
        // each store comes from one pool: the striped arena when the device
        // has one and it holds the store, else hipMalloc; plan within the
        // larger of the two, less the arena's granule rounding (under one
        // 256 MiB granule per allocation, sixteen granules in all)
        const int64_t held = c->adj_ws ? c->adj_ws->bytes : 0;
        const int64_t arena = arena_granule_bytes(c->device, c);
        const int64_t from_free = (int64_t)fr + held - (1LL << 30);
        const int64_t from_arena = arena > 0 ? arena + held - (16LL << 28) : 0;
        budget = from_free > from_arena ? from_free : from_arena;
    }
    const AdjSizes z = sizes_of(c);
    const long long cap = budget / (long long)sizeof(float);
    int b1 = 0, b2 = 0;
    long long bf = 0, bm = 0;
    // fewest forward steps, then least memory; every (c1, c2 <= c1)
    for (int c1 = 1; c1 <= nt; ++c1) {
        for (int c2 = c1; c2 >= 1; --c2) {
            const long long m = scheme_floats(z, nt, c1, c2);
            if (m > cap) continue;
            const long long f = scheme_fwd_steps(nt, c1, c2);
            if (!b1 || f < bf || (f == bf && m < bm)) { b1 = c1; b2 = c2; bf = f; bm = m; }
            break;  // a smaller c2 only recomputes more
        }
    }
    if (!b1) return ctx_fail(c, SWPC_ENOMEM, "swpc_adj_plan: no checkpoint scheme fits the budget");
    if (ckp_outer) *ckp_outer = b1;
    if (ckp_inner) *ckp_inner = b2;
    if (bytes) *bytes = bm * (int64_t)sizeof(float);
    if (fwd_steps) *fwd_steps = bf;
    return SWPC_OK;
}

extern "C" int swpc_adj_reset(swpc_ctx *c)
{
    if (!c) return SWPC_EINVAL;
    if (hipSetDevice(c->device) != hipSuccess) return SWPC_EHIP;
    int rc = ws_open(c);
    if (rc != SWPC_OK) return rc;
    if (zero(c->adj_ws->grad, 3 * c->adj_ws->z.nf, c->stream) != hipSuccess) return SWPC_EHIP;
    c->adj_ws->nsum = 0;
    return SWPC_OK;
}

extern "C" int swpc_adj_accumulate(swpc_ctx *c, int nt, const float *obs, int ckp_outer, int ckp_inner,
                                   double *misfit)
{
    if (!c) return SWPC_EINVAL;
    if (!obs || nt < 1 || ckp_outer < 1 || ckp_inner < 0)
        return ctx_fail(c, SWPC_EINVAL, "swpc_adj_accumulate: bad arguments");
    int rc = check_ready(c);
    if (rc != SWPC_OK) return rc;
    if (hipSetDevice(c->device) != hipSuccess) return SWPC_EHIP;
    if ((rc = ws_open(c)) != SWPC_OK) return rc;
    AdjWorkspace &w = *c->adj_ws;
    const int c1 = ckp_outer;
    const int c2 = (ckp_inner == 0 || ckp_inner >= c1) ? c1 : ckp_inner;
    hipStream_t s = c->stream;
    const Grid &g = c->g;
    const long long nf = w.z.nf, na18 = w.z.na18, na9 = 9 * c->npml, nR = w.z.nR;
    const long long per_state = w.z.state(), sstep = w.z.sstep();
    const long long nck = scheme_ck_slots(nt, c1, c2), nout = (nt + c1 - 1) / c1;
    const long long nobs = (long long)c->ntw * 3 * c->nst;
    const float *seg_before = w.seg;
    if (!wgrow(c, w.ck, w.ck_n, nck * per_state, nf) ||
        !wgrow(c, w.seg, w.seg_n, scheme_seg_slots(nt, c1, c2) * sstep, nf) ||
        !wgrow(c, w.obs, w.obs_n, nobs))
        return ctx_fail(c, SWPC_ENOMEM, "swpc_adj_accumulate: device memory for the checkpoints "
                                        "(swpc_adj_plan finds a scheme that fits)");
    // the steps recorded in place of the records never write the halo of a
    // field, which the next step's stencils read as zero: a new store starts
    // at zero and its halos stay so
    if (w.seg != seg_before && zero(w.seg, w.seg_n, s) != hipSuccess) return SWPC_EHIP;
    float *outer_ck = w.ck, *inner_ck = w.ck + nout * per_state;
    if (hipMemcpyAsync(w.obs, obs, nobs * sizeof(float), hipMemcpyHostToDevice, s) != hipSuccess) return SWPC_EHIP;

    auto save_state = [&](float *dst) -> hipError_t {
        for (int q = 0; q < SWPC_NFIELD; ++q) {
            hipError_t e = hipMemcpyAsync(dst + q * nf, c->fields[q], nf * sizeof(float), hipMemcpyDeviceToDevice, s);
            if (e != hipSuccess) return e;
        }
        hipError_t e = hipMemcpyAsync(dst + 9 * nf, c->ade, na18 * sizeof(float), hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess || !nR) return e;
        return hipMemcpyAsync(dst + 9 * nf + na18, c->R, nR * sizeof(float), hipMemcpyDeviceToDevice, s);
    };
    auto load_state = [&](const float *src) -> hipError_t {
        for (int q = 0; q < SWPC_NFIELD; ++q) {
            hipError_t e = hipMemcpyAsync(c->fields[q], src + q * nf, nf * sizeof(float), hipMemcpyDeviceToDevice, s);
            if (e != hipSuccess) return e;
        }
        hipError_t e = hipMemcpyAsync(c->ade, src + 9 * nf, na18 * sizeof(float), hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess || !nR) return e;
        return hipMemcpyAsync(c->R, src + 9 * nf + na18, nR * sizeof(float), hipMemcpyDeviceToDevice, s);
    };
    // one forward step n (time step it = n + 1), the order of main.f90:119-139
    auto stress_phase = [&](int it) -> hipError_t {
        hipError_t e = launch_stress_side(c, nullptr);
        if (e == hipSuccess) e = launch_stress_cols(c, 3);
        if (e == hipSuccess) e = launch_stressglut(c, it, nullptr);
        return e;
    };
    auto vel_phase = [&](int it) -> hipError_t {
        hipError_t e = launch_vel_side(c, nullptr);
        if (e == hipSuccess) e = launch_vel_cols(c, 3);
        if (e == hipSuccess) e = launch_bodyforce(c, it, nullptr);
        return e;
    };

    // Records of steps n0 .. n1-1 from the current state.  Record n holds
    // V_n (3 fields) | S_{n+1} (6 fields) | aV_n, aS_n (18 x npml, laid out as
    // the auxiliaries).  The steps run out of place: step n reads V_n, S_n and
    // the auxiliaries from the records and writes S_{n+1} into record n and
    // V_{n+1} and the new auxiliaries into record n + 1 (the last step into
    // the context's own arrays), so no field is copied but those of step n0.
    // Variant bit 4 keeps the copying form (the same bits; for checks).
    // sample: the receivers sample every step first (the first forward pass).
    auto record = [&](int n0, int n1, bool sample) -> hipError_t {
        hipError_t e = hipSuccess;
        if (c->variant & 16) {
            for (int n = n0; n < n1 && e == hipSuccess; ++n) {
                float *st = w.seg + (size_t)(n - n0) * sstep;
                const int it = n + 1;
                if (sample) e = launch_wav_store(c, it, nullptr);
                for (int q = 0; q < 3 && e == hipSuccess; ++q)
                    e = hipMemcpyAsync(st + q * nf, c->fields[q], nf * sizeof(float), hipMemcpyDeviceToDevice, s);
                if (e == hipSuccess)
                    e = hipMemcpyAsync(st + 9 * nf, c->ade, na9 * sizeof(float), hipMemcpyDeviceToDevice, s);
                if (e == hipSuccess) e = stress_phase(it);
                for (int q = 3; q < 9 && e == hipSuccess; ++q)
                    e = hipMemcpyAsync(st + q * nf, c->fields[q], nf * sizeof(float), hipMemcpyDeviceToDevice, s);
                if (e == hipSuccess)
                    e = hipMemcpyAsync(st + 9 * nf + na9, c->ade + na9, na9 * sizeof(float), hipMemcpyDeviceToDevice, s);
                if (e == hipSuccess) e = vel_phase(it);
            }
            return e;
        }
        auto V = [&](int n) { return w.seg + (size_t)(n - n0) * sstep; };
        auto S = [&](int n) { return V(n) + 3 * nf; };
        auto A = [&](int n) { return V(n) + 9 * nf; };
        float *real[SWPC_NFIELD];
        for (int q = 0; q < SWPC_NFIELD; ++q) real[q] = c->fields[q];
        float *const real_ade = c->ade;
        for (int q = 0; q < 3 && e == hipSuccess; ++q)
            e = hipMemcpyAsync(V(n0) + q * nf, real[q], nf * sizeof(float), hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(A(n0), real_ade, na18 * sizeof(float), hipMemcpyDeviceToDevice, s);
        for (int n = n0; n < n1 && e == hipSuccess; ++n) {
            const int it = n + 1;
            const bool last = n == n1 - 1;
            float *aux_out = last ? real_ade : A(n + 1);
            if (sample) {
                // the receivers read V_n and S_n where this step finds them
                for (int q = 0; q < 3; ++q) c->fields[q] = V(n) + q * nf;
                for (int q = 3; q < 9; ++q) c->fields[q] = n == n0 ? real[q] : S(n - 1) + (q - 3) * nf;
                e = launch_wav_store(c, it, nullptr);
                if (e != hipSuccess) break;
            }
            // stress phase: V_n read, S_n -> S_{n+1} in record n, aV_n -> aV_{n+1}
            for (int q = 0; q < 3; ++q) c->fields_in[q] = c->fields[q] = V(n) + q * nf;
            for (int q = 3; q < 9; ++q) {
                c->fields[q] = S(n) + (q - 3) * nf;
                c->fields_in[q] = n == n0 ? real[q] : S(n - 1) + (q - 3) * nf;
            }
            c->ade = aux_out;
            c->ade_in = A(n);
            e = stress_phase(it);
            // velocity phase: S_{n+1} read, V_n -> V_{n+1}, aS_n -> aS_{n+1}
            for (int q = 0; q < 3; ++q) {
                c->fields_in[q] = V(n) + q * nf;
                c->fields[q] = last ? real[q] : V(n + 1) + q * nf;
            }
            for (int q = 3; q < 9; ++q) c->fields_in[q] = c->fields[q] = S(n) + (q - 3) * nf;
            if (e == hipSuccess) e = vel_phase(it);
        }
        for (int q = 0; q < SWPC_NFIELD; ++q) {
            c->fields[q] = real[q];
            c->fields_in[q] = nullptr;
        }
        c->ade = real_ade;
        c->ade_in = nullptr;
        return e;
    };
    // ---- forward pass from rest, outer checkpoints every c1 steps ----
    hipError_t e = hipSuccess;
    for (int q = 0; q < SWPC_NFIELD && e == hipSuccess; ++q) e = zero(c->fields[q], nf, s);
    if (e == hipSuccess) e = zero(c->ade, na18, s);
    if (e == hipSuccess && nR) e = zero(c->R, nR, s);
    if (e == hipSuccess && c->d_uacc) e = zero(c->d_uacc, 3LL * c->nst, s);
    if (e == hipSuccess && c->d_eacc) e = zero(c->d_eacc, 6LL * c->nst, s);
    // The backward walk starts with the last outer segment: this pass also
    // lays its inner checkpoints and records its last inner segment, which
    // the walk would otherwise recompute (scheme_fwd_steps counts it so).
    const int N0L = (int)(nout - 1) * c1;
    const int ninL = c2 < nt - N0L ? (nt - N0L - 1) / c2 : 0;
    const int n0L = N0L + ninL * c2;
    for (int n = 0; n < n0L && e == hipSuccess; ++n) {
        const int it = n + 1;
        if (n % c1 == 0) e = save_state(outer_ck + (size_t)(n / c1) * per_state);
        if (e == hipSuccess && n > N0L && (n - N0L) % c2 == 0)
            e = save_state(inner_ck + (size_t)((n - N0L) / c2 - 1) * per_state);
        if (e == hipSuccess) e = launch_wav_store(c, it, nullptr);
        if (e == hipSuccess) e = stress_phase(it);
        if (e == hipSuccess) e = vel_phase(it);
    }
    if (e == hipSuccess) e = record(n0L, nt, true);
    if (e != hipSuccess) return ctx_fail(c, SWPC_EHIP, hipGetErrorString(e));

    // ---- misfit on the host (double) ----
    std::vector<float> syn(nobs);
    if (hipMemcpyAsync(syn.data(), c->d_wav[0], nobs * sizeof(float), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return SWPC_EHIP;
    const int nrec = (nt - 1) / c->ntdec_w + 1 < c->ntw ? (nt - 1) / c->ntdec_w + 1 : c->ntw;
    double J = 0.0;
    for (int st = 0; st < c->nst; ++st)
        for (int cc = 0; cc < 3; ++cc)
            for (int t = 0; t < nrec; ++t) {
                const size_t w3 = ((size_t)st * 3 + cc) * c->ntw + t;
                const double r = (double)syn[w3] - (double)obs[w3];
                J += 0.5 * r * r;
            }
    if (misfit) *misfit = J;

    // ---- backward pass ----
    AdjArgs a;
    a.g = g;
    a.c = c->coef;
    a.dt = c->desc.dt;
    a.hx = 1.0f / (float)c->desc.dx;
    a.hy = 1.0f / (float)c->desc.dy;
    a.hz = 1.0f / (float)c->desc.dz;
    a.rho = c->rho; a.lam = c->lam; a.mu = c->mu; a.taup = c->taup; a.taus = c->taus;
    a.aR = w.adjR;
    a.rn = g.rn;
    a.nm = c->desc.nm;
    for (int m = 0; m < SWPC_MAXNM; ++m) {
        a.c1[m] = c->visco.c1[m]; a.c2[m] = c->visco.c2[m]; a.d1[m] = c->visco.d1[m];
    }
    a.d2 = c->visco.d2;
    a.kfs_top = c->kfs_top; a.kfs_bot = c->kfs_bot; a.kob_top = c->kob_top; a.kob_bot = c->kob_bot;
    for (int q = 0; q < 5; ++q) a.box[q] = c->box[q];
    a.npml = c->npml;
    a.has_pml = 1;
    a.gxc = c->gxc; a.gxe = c->gxe; a.gyc = c->gyc; a.gye = c->gye; a.gzc = c->gzc; a.gze = c->gze;
    for (int q = 0; q < 9; ++q) {
        a.aF[q] = w.adjF + q * nf;
        a.D[q] = w.dens + q * nf;
    }
    for (int q = 0; q < 6; ++q) a.q[q] = w.qq + q * nf;
    a.cls = (const unsigned char *)w.cls;
    {
        const float h[3] = {a.hx, a.hy, a.hz};
        const float *p40[3] = {c->coef.re40x, c->coef.re40y, c->coef.re40z};
        const float *p41[3] = {c->coef.re41x, c->coef.re41y, c->coef.re41z};
        for (int ax = 0; ax < 3; ++ax) {
            const float ta[4] = {0.0f, p40[ax][0], p40[ax][1], h[ax]}, tb[4] = {0.0f, p41[ax][0], p41[ax][1], 0.0f};
            for (int q = 0; q < 4; ++q) { a.tA[ax][q] = ta[q]; a.tB[ax][q] = tb[q]; }
        }
    }
    a.aA = w.adjA;
    a.g_lam = w.gsrc; a.g_mu = w.gsrc + nf; a.g_rho = w.gsrc + 2 * nf;
    if (zero(w.adjF, 9 * nf, s) != hipSuccess || zero(w.adjA, na18, s) != hipSuccess ||
        zero(w.gsrc, 3 * nf, s) != hipSuccess || (nR && zero(w.adjR, nR, s) != hipSuccess))
        return SWPC_EHIP;
    hipLaunchKernelGGL(k_adj_class, dim3(1024), dim3(256), 0, s, a, (unsigned char *)w.cls);
    if (hipGetLastError() != hipSuccess) return SWPC_EHIP;
    const dim3 blk(64, 4, 1), grd(((((g.nx + 63) / 64) * ((g.ny + 3) / 4) + 7) & ~7) * g.nz);
    // record table: the distinct (field, point) targets of the stations'
    // trace adjoints, each with its station-components in station order
    const int ntgt_max = 6 * c->nst;
    if (!wgrow(c, w.rec, w.rec_n, 4LL * ntgt_max + 1))
        return ctx_fail(c, SWPC_ENOMEM, "swpc_adj_accumulate: device memory for the record table");
    long long *rec_tgt = (long long *)w.rec;
    int *rec_ptr = (int *)(w.rec + 2LL * ntgt_max);
    int *rec_src = rec_ptr + ntgt_max + 1;
    int ntgt = 0;
    {
        std::vector<int> hi(c->nst), hj(c->nst), hk(c->nst);
        if (hipMemcpy(hi.data(), c->d_ist, c->nst * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(hj.data(), c->d_jst, c->nst * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(hk.data(), c->d_kst, c->nst * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
            return SWPC_EHIP;
        std::map<long long, int> where;
        std::vector<long long> tg;
        std::vector<std::vector<int>> lists;
        for (int n = 0; n < c->nst; ++n) {
            const int i = hi[n], j = hj[n], k = hk[n];
            const struct { int f, i, j, k, cmp; } pts[6] = {{FVX, i, j, k, 0}, {FVX, i - 1, j, k, 0},
                                                            {FVY, i, j, k, 1}, {FVY, i, j - 1, k, 1},
                                                            {FVZ, i, j, k, 2}, {FVZ, i, j, k - 1, 2}};
            for (const auto &pt : pts) {
                const long long key = (long long)pt.f << 40 | gidx(g, pt.i, pt.j, pt.k);
                auto it = where.find(key);
                int id;
                if (it == where.end()) {
                    id = (int)tg.size();
                    where[key] = id;
                    tg.push_back(key);
                    lists.emplace_back();
                } else {
                    id = it->second;
                }
                lists[id].push_back(3 * n + pt.cmp);
            }
        }
        ntgt = (int)tg.size();
        std::vector<int> ptr(ntgt + 1, 0), src;
        for (int t = 0; t < ntgt; ++t) {
            ptr[t + 1] = ptr[t] + (int)lists[t].size();
            src.insert(src.end(), lists[t].begin(), lists[t].end());
        }
        if (hipMemcpy(rec_tgt, tg.data(), ntgt * sizeof(long long), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(rec_ptr, ptr.data(), (ntgt + 1) * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(rec_src, src.data(), src.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
            return SWPC_EHIP;
    }
    const float sc = 0.5f * c->wav_M0 * c->wav_UC * 1e9f;
    // the adjoint of steps n1-1 down to n0 from their records
    auto backward = [&](int n0, int n1) -> hipError_t {
        hipError_t e = hipSuccess;
        for (int n = n1 - 1; n >= n0 && e == hipSuccess; --n) {
            const float *st = w.seg + (size_t)(n - n0) * sstep;
            for (int q = 0; q < 9; ++q) a.fF[q] = st + q * nf;
            a.fA = st + 9 * nf + na9;      // stress auxiliaries aS_n
            hipLaunchKernelGGL(k_adj_vel_dens, grd, blk, 0, s, a);
            hipLaunchKernelGGL(k_adj_vel_gather, grd, blk, 0, s, a);
            a.fA = st + 9 * nf;            // velocity auxiliaries aV_n
            switch (a.nm) {
#define SWPC_ADJ_CASE(NMV) \
    case NMV: hipLaunchKernelGGL((k_adj_stress_dens<NMV>), grd, blk, 0, s, a); break;
                SWPC_ADJ_CASE(0) SWPC_ADJ_CASE(1) SWPC_ADJ_CASE(2) SWPC_ADJ_CASE(3) SWPC_ADJ_CASE(4)
                SWPC_ADJ_CASE(5) SWPC_ADJ_CASE(6) SWPC_ADJ_CASE(7) SWPC_ADJ_CASE(8)
#undef SWPC_ADJ_CASE
            default: break;
            }
            hipLaunchKernelGGL(k_adj_stress_gather, grd, blk, 0, s, a);
            if (n % c->ntdec_w == 0 && n / c->ntdec_w < c->ntw)
                hipLaunchKernelGGL(k_adj_record, dim3((ntgt + 63) / 64), dim3(64), 0, s, a, ntgt, rec_tgt, rec_ptr,
                                   rec_src, c->d_wav[0], w.obs, c->ntw, n / c->ntdec_w, sc);
            e = hipGetLastError();
        }
        return e;
    };
    for (int sg = (int)nout - 1; sg >= 0 && e == hipSuccess; --sg) {
        const int N0 = sg * c1, N1 = (N0 + c1 < nt) ? N0 + c1 : nt;
        const int nin = c2 < N1 - N0 ? (N1 - N0 - 1) / c2 : 0;  // inner segments after the first
        const bool laid = sg == (int)nout - 1;   // by the forward pass: inner checkpoints, last records
        if (!laid) {
            e = load_state(outer_ck + (size_t)sg * per_state);
            // inner checkpoints: the fields end at the start of the last inner segment
            for (int n = N0; n < N0 + nin * c2 && e == hipSuccess; ++n) {
                if (n > N0 && (n - N0) % c2 == 0) e = save_state(inner_ck + (size_t)((n - N0) / c2 - 1) * per_state);
                if (e == hipSuccess) e = stress_phase(n + 1);
                if (e == hipSuccess) e = vel_phase(n + 1);
            }
        }
        for (int is = nin; is >= 0 && e == hipSuccess; --is) {
            const int n0 = N0 + is * c2, n1 = (n0 + c2 < N1) ? n0 + c2 : N1;
            if (is < nin) e = load_state(is == 0 ? outer_ck + (size_t)sg * per_state : inner_ck + (size_t)(is - 1) * per_state);
            if (e == hipSuccess && !(laid && is == nin)) e = record(n0, n1, false);
            if (e == hipSuccess) e = backward(n0, n1);
        }
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_add, dim3(1024), dim3(256), 0, s, w.grad, w.gsrc, 3 * nf);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return ctx_fail(c, SWPC_EHIP, hipGetErrorString(e));
    if (hipStreamSynchronize(s) != hipSuccess) return SWPC_EHIP;
    w.nsum += 1;
    return SWPC_OK;
}

extern "C" int swpc_adj_get_gradient(swpc_ctx *c, float *g_lam, float *g_mu, float *g_rho)
{
    if (!c) return SWPC_EINVAL;
    if (!c->adj_ws || c->adj_ws->nsum == 0)
        return ctx_fail(c, SWPC_ESTATE, "swpc_adj_get_gradient: no gradient summed since the last reset");
    if (hipSetDevice(c->device) != hipSuccess) return SWPC_EHIP;
    hipStream_t s = c->stream;
    const long long nf = c->adj_ws->z.nf;
    // gradients out in the reference (k,i,j) halo layout
    float *outs[3] = {g_lam, g_mu, g_rho};
    for (int q = 0; q < 3; ++q) {
        if (!outs[q]) continue;
        if (launch_dev_to_ref(c, c->adj_ws->grad + q * nf, c->staging, s) != hipSuccess ||
            hipMemcpyAsync(outs[q], c->staging, (size_t)c->staging_n * sizeof(float), hipMemcpyDeviceToHost, s) !=
                hipSuccess)
            return SWPC_EHIP;
    }
    if (hipStreamSynchronize(s) != hipSuccess) return SWPC_EHIP;
    return SWPC_OK;
}

extern "C" int swpc_adj_gradient(swpc_ctx *c, int nt, const float *obs, int ckp, double *misfit, float *g_lam,
                                 float *g_mu, float *g_rho)
{
    if (!c) return SWPC_EINVAL;
    if (!obs || nt < 1 || ckp < 1) return ctx_fail(c, SWPC_EINVAL, "swpc_adj_gradient: bad arguments");
    int rc = check_ready(c);
    if (rc == SWPC_OK) rc = swpc_adj_reset(c);
    if (rc == SWPC_OK) rc = swpc_adj_accumulate(c, nt, obs, ckp, ckp, misfit);
    if (rc == SWPC_OK) rc = swpc_adj_get_gradient(c, g_lam, g_mu, g_rho);
    if (rc == SWPC_OK) rc = swpc_adj_release(c);
    return rc;
}
