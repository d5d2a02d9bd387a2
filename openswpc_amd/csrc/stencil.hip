// stencil.hip -- velocity/stress updates of OpenSWPC's 3D loop on gfx950.
//
// Each phase is two z-marching kernels: the interior (4th order + Zener
// memory variables, m_kernel) and the ADE-CFS PML (2nd order, m_absorb_p).
// A thread owns one (i,j) column of a 64 x BY tile of full-width rows whose
// first cell sits on a 128-byte line, and marches a chunk of planes along z,
// keeping the z-neighbours of its column in a rotating register queue while
// the x/y neighbours come from the plane the tile is streaming (L1/L2
// resident).  Loop bodies are branch free so every load of an iteration is
// issued at once.  Every arithmetic expression keeps the reference's FP32
// evaluation order; the library is compiled with -ffp-contract=off so no FMA
// changes the rounding.
//
// Cerjan tapers (m_absorb_c) are separate per-cell sweeps, as in the
// reference.
#include <vector>

#include "swpc_internal.h"

#define BX 64
#define MASK_INT 1
#define MASK_PML 2


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

typedef float f4v __attribute__((ext_vector_type(4)));
typedef float f2v __attribute__((ext_vector_type(2)));

template <bool NT, typename T>
__device__ static inline T ldv(const T *p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT, typename T>
__device__ static inline void stv(T *p, T v)
{
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// The 6*NM memory variables of one cell, value v = c*NM + m for component
// c = xx yy zz yz xz xy and mechanism m, packed as 16-byte vectors: value v
// lives in vector v/4 of stream v/4 (streams of rn float4 each), the last
// two values (6*NM mod 4 = 2) in a float2 stream after them.  One cell's
// values are loaded and stored together, so a wave moves 1 KB per access and
// the kernel keeps 6*NM/4 streams open instead of 6*NM.
template <int NM>
struct RCell {
    static constexpr int NV = 6 * NM, N4 = NV / 4, N2 = (NV % 4) / 2;
    float v[NV > 0 ? NV : 1];

    template <bool NT>
    __device__ __forceinline__ void load(const float *R, long long rn, long long r)
    {
#pragma unroll
        for (int s = 0; s < N4; ++s) {
            const f4v t = ldv<NT>((const f4v *)R + s * rn + r);
            v[4 * s] = t.x; v[4 * s + 1] = t.y; v[4 * s + 2] = t.z; v[4 * s + 3] = t.w;
        }
        if constexpr (N2 == 1) {
            const f2v t = ldv<NT>((const f2v *)(R + 4 * N4 * rn) + r);
            v[4 * N4] = t.x; v[4 * N4 + 1] = t.y;
        }
    }
    template <bool NT>
    __device__ __forceinline__ void store(float *R, long long rn, long long r) const
    {
#pragma unroll
        for (int s = 0; s < N4; ++s) {
            f4v t;
            t.x = v[4 * s]; t.y = v[4 * s + 1]; t.z = v[4 * s + 2]; t.w = v[4 * s + 3];
            stv<NT>((f4v *)R + s * rn + r, t);
        }
        if constexpr (N2 == 1) {
            f2v t;
            t.x = v[4 * N4]; t.y = v[4 * N4 + 1];
            stv<NT>((f2v *)(R + 4 * N4 * rn) + r, t);
        }
    }
};

// Streams read or written once per step (memory variables, stress, the
// per-cell medium) can bypass cache retention so the L2 keeps the V and mu
// rows that neighbouring tiles re-read.
template <bool NT>
__device__ static inline float ld1(const float *p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT>
__device__ static inline void st1(float *p, float v)
{
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

struct StepArgs {
    Grid g;
    FieldPtrs f;
    // where the updated fields' old values are read: f itself (in place), or
    // other buffers, so a step can write its results next to its inputs
    FieldPtrs fin;
    MediumPtrs m;
    float *R;
    ViscoCoef v;
    Coef4 c;
    float dt;
    // PML
    float *ade;
    const float *ade_in;  // old auxiliaries (ade when in place)
    long long npml;
    PmlBox box[5];
    const float4 *gxc, *gxe, *gyc, *gye, *gzc, *gze;
    float r20x, r20y, r20z;
    const int4 *tiles;   // side-column PML tiles: (i0 | width << 20, j0, k0, k1)
    int ntiles;
    int has_pml, mask;
    // interior launch shape
    const int *zc;       // first plane of each k chunk (nbz + 1 entries, last = nz + 1)
    int bx0, nbx, nby, nbz, by;
    int bxw, byi;        // interior tiles: bxw (64 or 32) cells wide, byi rows (bxw * byi = 64 * by threads)
};

// ---------------------------------------------------------------------------
// ADE-CFS PML cell updates (m_absorb_p).  A column's auxiliaries are numbered
// compactly box by box; the auxiliary of plane k is q = abase + k*astride.
// ---------------------------------------------------------------------------
__device__ static inline PmlBox pick_box(const StepArgs &a, int q)
{
    return q == 0 ? a.box[0] : q == 1 ? a.box[1] : q == 2 ? a.box[2] : q == 3 ? a.box[3] : a.box[4];
}

// absorb_p__update_stress for one cell, m_absorb_p.f90:455-517.  The caller
// passes its z-queue values: V at k (and k+1 for Vx, Vy; k-1 for Vz) and mu
// on planes k, k+1.
template <bool NT, bool OOP>
__device__ __forceinline__ void pml_stress_cell(const StepArgs &a, long long p, long long q, int k, float vx_0,
                                                float vx_p1, float vy_0, float vy_p1, float vz_m1, float vz_0,
                                                float mu00, float mu00n, float mu10, float mu10n, float mu01,
                                                float mu01n, float mu11, float4 gxc, float4 gxe, float4 gyc,
                                                float4 gye)
{
    const long long P = a.g.plane;
    const int Y = a.g.pitch;
    const float *__restrict__ Vx = a.f.Vx;
    const float *__restrict__ Vy = a.f.Vy;
    const float *__restrict__ Vz = a.f.Vz;
    const long long n = a.npml;
    float *__restrict__ A = a.ade + q;
    const float *__restrict__ AI = OOP ? a.ade_in + q : A;
    const float dt = a.dt;
    (void)P;
    const float vx_im1 = Vx[p - 1], vy_jm1 = Vy[p - Y];
    const float vy_ip1 = Vy[p + 1], vz_ip1 = Vz[p + 1];
    const float vx_jp1 = Vx[p + Y], vz_jp1 = Vz[p + Y];
    const float4 gzc = a.gzc[k - 1], gze = a.gze[k - 1];
    const float lam0 = ld1<NT>(a.m.lam + p);
    const float axVx = ld1<NT>(AI + A_axVx * n), ayVy = ld1<NT>(AI + A_ayVy * n), azVz = ld1<NT>(AI + A_azVz * n);
    const float ayVx = ld1<NT>(AI + A_ayVx * n), azVx = ld1<NT>(AI + A_azVx * n);
    const float axVy = ld1<NT>(AI + A_axVy * n), azVy = ld1<NT>(AI + A_azVy * n);
    const float axVz = ld1<NT>(AI + A_axVz * n), ayVz = ld1<NT>(AI + A_ayVz * n);
    float *__restrict__ Sxx = a.f.Sxx;
    float *__restrict__ Syy = a.f.Syy;
    float *__restrict__ Szz = a.f.Szz;
    float *__restrict__ Syz = a.f.Syz;
    float *__restrict__ Sxz = a.f.Sxz;
    float *__restrict__ Sxy = a.f.Sxy;
    const float *__restrict__ SxxI = OOP ? a.fin.Sxx : Sxx;
    const float *__restrict__ SyyI = OOP ? a.fin.Syy : Syy;
    const float *__restrict__ SzzI = OOP ? a.fin.Szz : Szz;
    const float *__restrict__ SyzI = OOP ? a.fin.Syz : Syz;
    const float *__restrict__ SxzI = OOP ? a.fin.Sxz : Sxz;
    const float *__restrict__ SxyI = OOP ? a.fin.Sxy : Sxy;

    // normal components
    const float dxVx = (vx_0 - vx_im1) * a.r20x;
    const float dyVy = (vy_0 - vy_jm1) * a.r20y;
    const float dzVz = (vz_0 - vz_m1) * a.r20z;
    const float lam2mu_R = (lam0 + 2.0f * mu00);
    const float lam_R = lam2mu_R - 2.0f * mu00;
    const float dxVx_ade = gxc.x * dxVx + gxc.y * axVx;
    const float dyVy_ade = gyc.x * dyVy + gyc.y * ayVy;
    const float dzVz_ade = gzc.x * dzVz + gzc.y * azVz;
    st1<NT>(Sxx + p, ld1<NT>(SxxI + p) + (lam2mu_R * dxVx_ade + lam_R * (dyVy_ade + dzVz_ade)) * dt);
    st1<NT>(Syy + p, ld1<NT>(SyyI + p) + (lam2mu_R * dyVy_ade + lam_R * (dxVx_ade + dzVz_ade)) * dt);
    st1<NT>(Szz + p, ld1<NT>(SzzI + p) + (lam2mu_R * dzVz_ade + lam_R * (dxVx_ade + dyVy_ade)) * dt);
    st1<NT>(A + A_axVx * n, gxc.z * axVx + gxc.w * dxVx * dt);
    st1<NT>(A + A_ayVy * n, gyc.z * ayVy + gyc.w * dyVy * dt);
    st1<NT>(A + A_azVz * n, gzc.z * azVz + gzc.w * dzVz * dt);

    // shear components
    const float dxVy = (vy_ip1 - vy_0) * a.r20x;
    const float dxVz = (vz_ip1 - vz_0) * a.r20x;
    const float dyVx = (vx_jp1 - vx_0) * a.r20y;
    const float dyVz = (vz_jp1 - vz_0) * a.r20y;
    const float dzVx = (vx_p1 - vx_0) * a.r20z;
    const float dzVy = (vy_p1 - vy_0) * a.r20z;
    const float muxz = hmean4(mu00, mu00n, mu10, mu10n, EPSL);
    const float muxy = hmean4(mu00, mu10, mu01, mu11, EPSL);
    const float muyz = hmean4(mu00, mu00n, mu01, mu01n, EPSL);
    st1<NT>(Syz + p, ld1<NT>(SyzI + p) + muyz * (gye.x * dyVz + gze.x * dzVy + gye.y * ayVz + gze.y * azVy) * dt);
    st1<NT>(Sxz + p, ld1<NT>(SxzI + p) + muxz * (gxe.x * dxVz + gze.x * dzVx + gxe.y * axVz + gze.y * azVx) * dt);
    st1<NT>(Sxy + p, ld1<NT>(SxyI + p) + muxy * (gxe.x * dxVy + gye.x * dyVx + gxe.y * axVy + gye.y * ayVx) * dt);
    st1<NT>(A + A_ayVx * n, gye.z * ayVx + gye.w * dyVx * dt);
    st1<NT>(A + A_azVx * n, gze.z * azVx + gze.w * dzVx * dt);
    st1<NT>(A + A_axVy * n, gxe.z * axVy + gxe.w * dxVy * dt);
    st1<NT>(A + A_azVy * n, gze.z * azVy + gze.w * dzVy * dt);
    st1<NT>(A + A_axVz * n, gxe.z * axVz + gxe.w * dxVz * dt);
    st1<NT>(A + A_ayVz * n, gye.z * ayVz + gye.w * dyVz * dt);
}

// absorb_p__update_vel for one cell, m_absorb_p.f90:267-304.  Queue values:
// Szz on k, k+1; Syz, Sxz on k-1, k; rho on k, k+1.
template <bool OOP>
__device__ __forceinline__ void pml_vel_cell(const StepArgs &a, long long p, long long q, int k, float szz_0,
                                             float szz_p1, float syz_m1, float syz_0, float sxz_m1, float sxz_0,
                                             float rho0, float rho_kp, float4 gxc, float4 gxe, float4 gyc,
                                             float4 gye)
{
    const int Y = a.g.pitch;
    const float *__restrict__ Sxx = a.f.Sxx;
    const float *__restrict__ Syy = a.f.Syy;
    const float *__restrict__ Syz = a.f.Syz;
    const float *__restrict__ Sxz = a.f.Sxz;
    const float *__restrict__ Sxy = a.f.Sxy;
    const float *__restrict__ rho = a.m.rho;
    const long long n = a.npml;
    float *__restrict__ A = a.ade + q;
    const float *__restrict__ AI = OOP ? a.ade_in + q : A;
    const float dt = a.dt;
    const float sxx_0 = Sxx[p], sxx_ip1 = Sxx[p + 1];
    const float syy_0 = Syy[p], syy_jp1 = Syy[p + Y];
    const float syz_jm1 = Syz[p - Y], sxz_im1 = Sxz[p - 1];
    const float sxy_0 = Sxy[p], sxy_im1 = Sxy[p - 1], sxy_jm1 = Sxy[p - Y];
    const float rho_ip = rho[p + 1], rho_jp = rho[p + Y];
    const float4 gzc = a.gzc[k - 1], gze = a.gze[k - 1];
    const float axSxx = ld1<true>(AI + A_axSxx * n), aySxy = ld1<true>(AI + A_aySxy * n), azSxz = ld1<true>(AI + A_azSxz * n);
    const float axSxy = ld1<true>(AI + A_axSxy * n), aySyy = ld1<true>(AI + A_aySyy * n), azSyz = ld1<true>(AI + A_azSyz * n);
    const float axSxz = ld1<true>(AI + A_axSxz * n), aySyz = ld1<true>(AI + A_aySyz * n), azSzz = ld1<true>(AI + A_azSzz * n);

    const float dxSxx = (sxx_ip1 - sxx_0) * a.r20x;
    const float dySyy = (syy_jp1 - syy_0) * a.r20y;
    const float dzSzz = (szz_p1 - szz_0) * a.r20z;
    const float dySyz = (syz_0 - syz_jm1) * a.r20y;
    const float dzSyz = (syz_0 - syz_m1) * a.r20z;
    const float dxSxz = (sxz_0 - sxz_im1) * a.r20x;
    const float dzSxz = (sxz_0 - sxz_m1) * a.r20z;
    const float dxSxy = (sxy_0 - sxy_im1) * a.r20x;
    const float dySxy = (sxy_0 - sxy_jm1) * a.r20y;
    const float bx = 2.0f / (rho0 + rho_ip);
    const float by = 2.0f / (rho0 + rho_jp);
    const float bz = 2.0f / (rho0 + rho_kp);

    float *__restrict__ Vx = a.f.Vx;
    float *__restrict__ Vy = a.f.Vy;
    float *__restrict__ Vz = a.f.Vz;
    st1<true>(Vx + p, ld1<true>((OOP ? a.fin.Vx : Vx) + p) + bx * (gxe.x * dxSxx + gyc.x * dySxy + gzc.x * dzSxz
                                                + gxe.y * axSxx + gyc.y * aySxy + gzc.y * azSxz) * dt);
    st1<true>(Vy + p, ld1<true>((OOP ? a.fin.Vy : Vy) + p) + by * (gxc.x * dxSxy + gye.x * dySyy + gzc.x * dzSyz
                                                + gxc.y * axSxy + gye.y * aySyy + gzc.y * azSyz) * dt);
    st1<true>(Vz + p, ld1<true>((OOP ? a.fin.Vz : Vz) + p) + bz * (gxc.x * dxSxz + gyc.x * dySyz + gze.x * dzSzz
                                                + gxc.y * axSxz + gyc.y * aySyz + gze.y * azSzz) * dt);
    st1<true>(A + A_axSxx * n, gxe.z * axSxx + gxe.w * dxSxx * dt);
    st1<true>(A + A_aySxy * n, gyc.z * aySxy + gyc.w * dySxy * dt);
    st1<true>(A + A_azSxz * n, gzc.z * azSxz + gzc.w * dzSxz * dt);
    st1<true>(A + A_axSxy * n, gxc.z * axSxy + gxc.w * dxSxy * dt);
    st1<true>(A + A_aySyy * n, gye.z * aySyy + gye.w * dySyy * dt);
    st1<true>(A + A_azSyz * n, gzc.z * azSyz + gzc.w * dzSyz * dt);
    st1<true>(A + A_axSxz * n, gxc.z * axSxz + gxc.w * dxSxz * dt);
    st1<true>(A + A_aySyz * n, gyc.z * aySyz + gyc.w * dySyz * dt);
    st1<true>(A + A_azSzz * n, gze.z * azSzz + gze.w * dzSzz * dt);
}

// ---------------------------------------------------------------------------
// Interior columns (m_kernel, and the bottom PML band of the same columns).
// Tiles of 64 x BY columns over full-width aligned rows; lanes outside
// [ib,ie] x [jb,je] idle.  Chunks run over k = 1..nz: planes 1..ke take the
// interior update, planes ke+1..nz (the bottom PML, kbeg_a = nz-na+1) the
// ADE-CFS update, in two branch-free loops sharing the z queues.
// ---------------------------------------------------------------------------

// kernel__update_stress, m_kernel.f90:142-348 (normal and shear parts fused:
// both read V and write disjoint arrays)
// The velocities the stress update of one cell reads: the z, x and y lines
// through the cell (m_kernel.f90:198-333)
struct VStencil {
    float vx_m1, vx_0, vx_p1, vx_p2, vx_im2, vx_im1, vx_ip1, vx_jm1, vx_jp1, vx_jp2;
    float vy_m1, vy_0, vy_p1, vy_p2, vy_im1, vy_ip1, vy_ip2, vy_jm2, vy_jm1, vy_jp1;
    float vz_m2, vz_m1, vz_0, vz_p1, vz_im1, vz_ip1, vz_ip2, vz_jm1, vz_jp1, vz_jp2;
};

// kernel__update_stress for one cell of the interior box, m_kernel.f90:198-333:
// the normal and shear stresses and their memory variables.  p: field index,
// r: memory-variable index, in: the cell is in the 2nd-order band, mu..: mu
// at (i,j), (i+1,j), (i,j+1) on planes k and k+1 (n) and at (i+1,j+1) on
// plane k.  OOP: the old stresses are read from a.fin.
template <int NM, bool NT, bool OOP>
__device__ __forceinline__ void stress_cell(const StepArgs &a, long long p, long long r, bool in, const VStencil &v,
                                            float mu00, float mu00n, float mu10, float mu10n, float mu01, float mu01n,
                                            float mu11)
{
    float *__restrict__ Sxx = a.f.Sxx;
    float *__restrict__ Syy = a.f.Syy;
    float *__restrict__ Szz = a.f.Szz;
    float *__restrict__ Syz = a.f.Syz;
    float *__restrict__ Sxz = a.f.Sxz;
    float *__restrict__ Sxy = a.f.Sxy;
    const float *__restrict__ SxxI = OOP ? a.fin.Sxx : Sxx;
    const float *__restrict__ SyyI = OOP ? a.fin.Syy : Syy;
    const float *__restrict__ SzzI = OOP ? a.fin.Szz : Szz;
    const float *__restrict__ SyzI = OOP ? a.fin.Syz : Syz;
    const float *__restrict__ SxzI = OOP ? a.fin.Sxz : Sxz;
    const float *__restrict__ SxyI = OOP ? a.fin.Sxy : Sxy;
    const float dt = a.dt;
    const float d2 = a.v.d2;
    const long long rn = a.g.rn;
    const float re40x = in ? a.c.re40x[1] : a.c.re40x[0];
    const float re41x = in ? a.c.re41x[1] : a.c.re41x[0];
    const float re40y = in ? a.c.re40y[1] : a.c.re40y[0];
    const float re41y = in ? a.c.re41y[1] : a.c.re41y[0];
    const float re40z = in ? a.c.re40z[1] : a.c.re40z[0];
    const float re41z = in ? a.c.re41z[1] : a.c.re41z[0];

    // normal components, m_kernel.f90:198-238
    const float dxVx = (v.vx_0 - v.vx_im1) * re40x - (v.vx_ip1 - v.vx_im2) * re41x;
    const float dyVy = (v.vy_0 - v.vy_jm1) * re40y - (v.vy_jp1 - v.vy_jm2) * re41y;
    const float dzVz = (v.vz_0 - v.vz_m1) * re40z - (v.vz_p1 - v.vz_m2) * re41z;

    const float mu2 = 2.0f * mu00;
    const float lam2mu = ld1<NT>(a.m.lam + p) + mu2;
    const float taup1 = ld1<NT>(a.m.taup + p);
    const float taus1 = ld1<NT>(a.m.taus + p);
    const float d3v3 = dxVx + dyVy + dzVz;
    const float dyVy_dzVz = dyVy + dzVz;
    const float dxVx_dzVz = dxVx + dzVz;
    const float dxVx_dyVy = dxVx + dyVy;

    RCell<NM> R;
    R.template load<NT>(a.R, rn, r);
    float Rxx_n = 0.0f, Ryy_n = 0.0f, Rzz_n = 0.0f;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const float c1 = a.v.c1[m], c2 = a.v.c2[m], d1 = a.v.d1[m];
        const float nxx = c1 * R.v[0 * NM + m] - c2 * (lam2mu * taup1 * d3v3 - mu2 * taus1 * dyVy_dzVz) * dt;
        const float nyy = c1 * R.v[1 * NM + m] - c2 * (lam2mu * taup1 * d3v3 - mu2 * taus1 * dxVx_dzVz) * dt;
        const float nzz = c1 * R.v[2 * NM + m] - c2 * (lam2mu * taup1 * d3v3 - mu2 * taus1 * dxVx_dyVy) * dt;
        R.v[0 * NM + m] = nxx;
        R.v[1 * NM + m] = nyy;
        R.v[2 * NM + m] = nzz;
        Rxx_n = Rxx_n + d1 * nxx;
        Ryy_n = Ryy_n + d1 * nyy;
        Rzz_n = Rzz_n + d1 * nzz;
    }
    const float taup_plus1 = 1.0f + taup1 * (1.0f + d2);
    const float taus_plus1 = 1.0f + taus1 * (1.0f + d2);
    st1<NT>(Sxx + p, ld1<NT>(SxxI + p) + (lam2mu * taup_plus1 * d3v3 - mu2 * taus_plus1 * dyVy_dzVz + Rxx_n) * dt);
    st1<NT>(Syy + p, ld1<NT>(SyyI + p) + (lam2mu * taup_plus1 * d3v3 - mu2 * taus_plus1 * dxVx_dzVz + Ryy_n) * dt);
    st1<NT>(Szz + p, ld1<NT>(SzzI + p) + (lam2mu * taup_plus1 * d3v3 - mu2 * taus_plus1 * dxVx_dyVy + Rzz_n) * dt);

    // shear components, m_kernel.f90:285-333
    const float dxVy_dyVx = (v.vy_ip1 - v.vy_0) * re40x - (v.vy_ip2 - v.vy_im1) * re41x
                          + (v.vx_jp1 - v.vx_0) * re40y - (v.vx_jp2 - v.vx_jm1) * re41y;
    const float dxVz_dzVx = (v.vz_ip1 - v.vz_0) * re40x - (v.vz_ip2 - v.vz_im1) * re41x
                          + (v.vx_p1 - v.vx_0) * re40z - (v.vx_p2 - v.vx_m1) * re41z;
    const float dyVz_dzVy = (v.vz_jp1 - v.vz_0) * re40y - (v.vz_jp2 - v.vz_jm1) * re41y
                          + (v.vy_p1 - v.vy_0) * re40z - (v.vy_p2 - v.vy_m1) * re41z;

    const float muxz = hmean4(mu00, mu00n, mu10, mu10n, EPSL);
    const float muxy = hmean4(mu00, mu10, mu01, mu11, EPSL);
    const float muyz = hmean4(mu00, mu00n, mu01, mu01n, EPSL);

    float Ryz_n = 0.0f, Rxz_n = 0.0f, Rxy_n = 0.0f;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const float c1 = a.v.c1[m], c2 = a.v.c2[m], d1 = a.v.d1[m];
        const float nyz = c1 * R.v[3 * NM + m] - c2 * muyz * taus1 * dyVz_dzVy * dt;
        const float nxz = c1 * R.v[4 * NM + m] - c2 * muxz * taus1 * dxVz_dzVx * dt;
        const float nxy = c1 * R.v[5 * NM + m] - c2 * muxy * taus1 * dxVy_dyVx * dt;
        R.v[3 * NM + m] = nyz;
        R.v[4 * NM + m] = nxz;
        R.v[5 * NM + m] = nxy;
        Ryz_n = Ryz_n + d1 * nyz;
        Rxz_n = Rxz_n + d1 * nxz;
        Rxy_n = Rxy_n + d1 * nxy;
    }
    R.template store<NT>(a.R, rn, r);
    st1<NT>(Syz + p, ld1<NT>(SyzI + p) + (muyz * taus_plus1 * dyVz_dzVy + Ryz_n) * dt);
    st1<NT>(Sxz + p, ld1<NT>(SxzI + p) + (muxz * taus_plus1 * dxVz_dzVx + Rxz_n) * dt);
    st1<NT>(Sxy + p, ld1<NT>(SxyI + p) + (muxy * taus_plus1 * dxVy_dyVx + Rxy_n) * dt);
}

// logical tile (bx, by, bz) of a plain launch: blocks of one XCD take a
// contiguous run of tiles, x fastest
__device__ static inline void plain_tile(const StepArgs &a, int &bx, int &byy, int &bz)
{
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    bx = bid % a.nbx;
    const int t = bid / a.nbx;
    byy = t % a.nby;
    bz = t / a.nby;
}

template <int NM, bool NT, bool OOP>
__device__ __forceinline__ void stress_int_body(const StepArgs &a, int bx, int byy, int bz)
{
    const Grid &g = a.g;
    const int i = 1 + (a.bx0 + bx) * a.bxw + threadIdx.x;
    const int j = g.jb + byy * a.byi + threadIdx.y;
    if (i < g.ib || i > g.ie || j > g.je) return;
    const int k0 = a.zc[bz];
    const int k1 = a.zc[bz + 1] - 1;
    const int ka1 = (a.mask & MASK_INT) ? min(k1, g.ke) : k0 - 1;      // interior planes k0..ka1
    const int kb0 = max(k0, g.ke + 1);
    const int kb1 = (a.has_pml && (a.mask & MASK_PML)) ? k1 : kb0 - 1;  // bottom PML planes kb0..kb1
    if (ka1 < k0 && kb1 < kb0) return;
    const int ks = ka1 >= k0 ? k0 : kb0;

    const long long P = g.plane;
    const int Y = g.pitch;
    const float *__restrict__ Vx = a.f.Vx;
    const float *__restrict__ Vy = a.f.Vy;
    const float *__restrict__ Vz = a.f.Vz;
    const float *__restrict__ mu = a.m.mu;

    long long p = gidx(g, i, j, ks);
    float vx_m1 = Vx[p - P], vx_0 = Vx[p], vx_p1 = Vx[p + P];
    float vy_m1 = Vy[p - P], vy_0 = Vy[p], vy_p1 = Vy[p + P];
    float vz_m2 = Vz[p - 2 * P], vz_m1 = Vz[p - P], vz_0 = Vz[p];
    float mu00 = mu[p], mu10 = mu[p + 1], mu01 = mu[p + Y], mu11 = mu[p + 1 + Y];

    if (ka1 >= k0) {
        const long long c2 = g2idx(g, i, j);
        const int kft = a.m.kfs_top[c2], kfb = a.m.kfs_bot[c2];
        const int kot = a.m.kob_top[c2], kobb = a.m.kob_bot[c2];
        long long r = ((long long)(k0 - g.kb) * g.nyk + (j - g.jb)) * g.rpitch + (i - 1);
        for (int k = k0; k <= ka1; ++k, p += P, r += g.rplane) {
            const float vx_p2 = Vx[p + 2 * P], vy_p2 = Vy[p + 2 * P], vz_p1 = Vz[p + P];
            const float mu00n = mu[p + P], mu10n = mu[p + P + 1], mu01n = mu[p + P + Y];

            VStencil v;
            v.vx_m1 = vx_m1; v.vx_0 = vx_0; v.vx_p1 = vx_p1; v.vx_p2 = vx_p2;
            v.vy_m1 = vy_m1; v.vy_0 = vy_0; v.vy_p1 = vy_p1; v.vy_p2 = vy_p2;
            v.vz_m2 = vz_m2; v.vz_m1 = vz_m1; v.vz_0 = vz_0; v.vz_p1 = vz_p1;
            v.vx_im2 = Vx[p - 2]; v.vx_im1 = Vx[p - 1]; v.vx_ip1 = Vx[p + 1];
            v.vx_jm1 = Vx[p - Y]; v.vx_jp1 = Vx[p + Y]; v.vx_jp2 = Vx[p + 2 * Y];
            v.vy_jm2 = Vy[p - 2 * Y]; v.vy_jm1 = Vy[p - Y]; v.vy_jp1 = Vy[p + Y];
            v.vy_im1 = Vy[p - 1]; v.vy_ip1 = Vy[p + 1]; v.vy_ip2 = Vy[p + 2];
            v.vz_im1 = Vz[p - 1]; v.vz_ip1 = Vz[p + 1]; v.vz_ip2 = Vz[p + 2];
            v.vz_jm1 = Vz[p - Y]; v.vz_jp1 = Vz[p + Y]; v.vz_jp2 = Vz[p + 2 * Y];
            stress_cell<NM, NT, OOP>(a, p, r, in_band(k, kft, kfb, kot, kobb), v, mu00, mu00n, mu10, mu10n, mu01, mu01n,
                                     mu11);

            // rotate the z queues
            vx_m1 = vx_0; vx_0 = vx_p1; vx_p1 = vx_p2;
            vy_m1 = vy_0; vy_0 = vy_p1; vy_p1 = vy_p2;
            vz_m2 = vz_m1; vz_m1 = vz_0; vz_0 = vz_p1;
            mu00 = mu00n; mu10 = mu10n; mu01 = mu01n;
            mu11 = mu[p + P + 1 + Y];
        }
    }
    if (kb1 >= kb0) {
        // bottom PML band of the interior columns (box 4)
        const PmlBox b = a.box[4];
        const long long astride = (long long)b.nxb * b.nyb;
        long long q = b.off + (long long)(j - b.j0) * b.nxb + (i - b.i0) + (long long)(kb0 - b.k0) * astride;
        const float4 gxc = a.gxc[i - 1], gxe = a.gxe[i - 1];
        const float4 gyc = a.gyc[j - 1], gye = a.gye[j - 1];
        for (int k = kb0; k <= kb1; ++k, p += P, q += astride) {
            const float vx_p2 = Vx[p + 2 * P], vy_p2 = Vy[p + 2 * P], vz_p1 = Vz[p + P];
            const float mu00n = mu[p + P], mu10n = mu[p + P + 1], mu01n = mu[p + P + Y];
            pml_stress_cell<NT, OOP>(a, p, q, k, vx_0, vx_p1, vy_0, vy_p1, vz_m1, vz_0, mu00, mu00n, mu10, mu10n, mu01,
                                mu01n, mu11, gxc, gxe, gyc, gye);
            vx_m1 = vx_0; vx_0 = vx_p1; vx_p1 = vx_p2;
            vy_m1 = vy_0; vy_0 = vy_p1; vy_p1 = vy_p2;
            vz_m2 = vz_m1; vz_m1 = vz_0; vz_0 = vz_p1;
            mu00 = mu00n; mu10 = mu10n; mu01 = mu01n;
            mu11 = mu[p + P + 1 + Y];
        }
    }
}

template <int NM, bool NT, bool OOP>
__global__ void __launch_bounds__(256) k_stress_int(StepArgs a)
{
    int bx, byy, bz;
    plain_tile(a, bx, byy, bz);
    stress_int_body<NM, NT, OOP>(a, bx, byy, bz);
}

// batched launch: tile = (member, bx, by, bz), member arguments in HBM
template <int NM, bool NT>
__global__ void __launch_bounds__(256) k_stress_int_b(const StepArgs *__restrict__ mem, const int4 *__restrict__ tiles)
{
    const int4 t = tiles[xcd_remap(blockIdx.x, gridDim.x)];
    stress_int_body<NM, NT, false>(mem[t.x], t.y, t.z, t.w);
}

// Every load of one plane of the interior velocity update: the z-queue
// entries the plane adds, the x/y neighbours, the buoyancy densities and the
// old velocities.  Loading the next plane's set before computing this one's
// keeps two planes of loads in flight per wave.
struct VelLoads {
    float sxz_p1, syz_p1, szz_p2, rho_kp;
    float sxx_im1, sxx_0, sxx_ip1, sxx_ip2;
    float sxy_0, sxy_jm2, sxy_jm1, sxy_jp1, sxy_im2, sxy_im1, sxy_ip1;
    float syy_jm1, syy_0, syy_jp1, syy_jp2;
    float syz_jm2, syz_jm1, syz_jp1;
    float sxz_im2, sxz_im1, sxz_ip1;
    float rho_ip, rho_jp;
    float vx, vy, vz;
};

template <bool OOP>
__device__ __forceinline__ void vel_loads(const StepArgs &a, long long p, VelLoads &L)
{
    const long long P = a.g.plane;
    const int Y = a.g.pitch;
    const float *__restrict__ Sxx = a.f.Sxx;
    const float *__restrict__ Syy = a.f.Syy;
    const float *__restrict__ Szz = a.f.Szz;
    const float *__restrict__ Syz = a.f.Syz;
    const float *__restrict__ Sxz = a.f.Sxz;
    const float *__restrict__ Sxy = a.f.Sxy;
    const float *__restrict__ rho = a.m.rho;
    L.sxz_p1 = Sxz[p + P]; L.syz_p1 = Syz[p + P]; L.szz_p2 = Szz[p + 2 * P]; L.rho_kp = rho[p + P];
    L.sxx_im1 = Sxx[p - 1]; L.sxx_0 = Sxx[p]; L.sxx_ip1 = Sxx[p + 1]; L.sxx_ip2 = Sxx[p + 2];
    L.sxy_0 = Sxy[p];
    L.sxy_jm2 = Sxy[p - 2 * Y]; L.sxy_jm1 = Sxy[p - Y]; L.sxy_jp1 = Sxy[p + Y];
    L.sxy_im2 = Sxy[p - 2]; L.sxy_im1 = Sxy[p - 1]; L.sxy_ip1 = Sxy[p + 1];
    L.syy_jm1 = Syy[p - Y]; L.syy_0 = Syy[p]; L.syy_jp1 = Syy[p + Y]; L.syy_jp2 = Syy[p + 2 * Y];
    L.syz_jm2 = Syz[p - 2 * Y]; L.syz_jm1 = Syz[p - Y]; L.syz_jp1 = Syz[p + Y];
    L.sxz_im2 = Sxz[p - 2]; L.sxz_im1 = Sxz[p - 1]; L.sxz_ip1 = Sxz[p + 1];
    L.rho_ip = rho[p + 1]; L.rho_jp = rho[p + Y];
    L.vx = ld1<true>((OOP ? a.fin.Vx : a.f.Vx) + p);
    L.vy = ld1<true>((OOP ? a.fin.Vy : a.f.Vy) + p);
    L.vz = ld1<true>((OOP ? a.fin.Vz : a.f.Vz) + p);
}

// kernel__update_vel, m_kernel.f90:75-140, then the bottom PML band.
// PF: the interior loop loads plane k+1 while it computes plane k.
template <bool OOP, bool PF>
__device__ __forceinline__ void vel_int_body(const StepArgs &a, int bx, int byy, int bz)
{
    const Grid &g = a.g;
    const int i = 1 + (a.bx0 + bx) * a.bxw + threadIdx.x;
    const int j = g.jb + byy * a.byi + threadIdx.y;
    if (i < g.ib || i > g.ie || j > g.je) return;
    const int k0 = a.zc[bz];
    const int k1 = a.zc[bz + 1] - 1;
    const int ka1 = (a.mask & MASK_INT) ? min(k1, g.ke) : k0 - 1;
    const int kb0 = max(k0, g.ke + 1);
    const int kb1 = (a.has_pml && (a.mask & MASK_PML)) ? k1 : kb0 - 1;
    if (ka1 < k0 && kb1 < kb0) return;
    const int ks = ka1 >= k0 ? k0 : kb0;

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

    long long p = gidx(g, i, j, ks);
    float sxz_m2 = Sxz[p - 2 * P], sxz_m1 = Sxz[p - P], sxz_0 = Sxz[p];
    float syz_m2 = Syz[p - 2 * P], syz_m1 = Syz[p - P], syz_0 = Syz[p];
    float szz_m1 = Szz[p - P], szz_0 = Szz[p], szz_p1 = Szz[p + P];
    float rho0 = rho[p];

    if (PF && ka1 >= k0) {
        const long long c2 = g2idx(g, i, j);
        const int kft = a.m.kfs_top[c2], kfb = a.m.kfs_bot[c2];
        const int kot = a.m.kob_top[c2], kobb = a.m.kob_bot[c2];
        VelLoads L;
        vel_loads<OOP>(a, p, L);
        for (int k = k0; k <= ka1; ++k, p += P) {
            // plane k+1 <= nz + 1: its loads stay within the -2..nz+3 halo
            VelLoads N;
            if (k < ka1) vel_loads<OOP>(a, p + P, N);

            const bool in = in_band(k, kft, kfb, kot, kobb);
            const float re40x = in ? a.c.re40x[1] : a.c.re40x[0];
            const float re41x = in ? a.c.re41x[1] : a.c.re41x[0];
            const float re40y = in ? a.c.re40y[1] : a.c.re40y[0];
            const float re41y = in ? a.c.re41y[1] : a.c.re41y[0];
            const float re40z = in ? a.c.re40z[1] : a.c.re40z[0];
            const float re41z = in ? a.c.re41z[1] : a.c.re41z[0];

            const float d3Sx3 = (L.sxx_ip1 - L.sxx_0) * re40x - (L.sxx_ip2 - L.sxx_im1) * re41x
                              + (L.sxy_0 - L.sxy_jm1) * re40y - (L.sxy_jp1 - L.sxy_jm2) * re41y
                              + (sxz_0 - sxz_m1) * re40z - (L.sxz_p1 - sxz_m2) * re41z;
            const float d3Sy3 = (L.sxy_0 - L.sxy_im1) * re40x - (L.sxy_ip1 - L.sxy_im2) * re41x
                              + (L.syy_jp1 - L.syy_0) * re40y - (L.syy_jp2 - L.syy_jm1) * re41y
                              + (syz_0 - syz_m1) * re40z - (L.syz_p1 - syz_m2) * re41z;
            const float d3Sz3 = (sxz_0 - L.sxz_im1) * re40x - (L.sxz_ip1 - L.sxz_im2) * re41x
                              + (syz_0 - L.syz_jm1) * re40y - (L.syz_jp1 - L.syz_jm2) * re41y
                              + (szz_p1 - szz_0) * re40z - (L.szz_p2 - szz_m1) * re41z;

            st1<true>(Vx + p, L.vx + 2.0f / (rho0 + L.rho_ip) * d3Sx3 * dt);
            st1<true>(Vy + p, L.vy + 2.0f / (rho0 + L.rho_jp) * d3Sy3 * dt);
            st1<true>(Vz + p, L.vz + 2.0f / (rho0 + L.rho_kp) * d3Sz3 * dt);

            sxz_m2 = sxz_m1; sxz_m1 = sxz_0; sxz_0 = L.sxz_p1;
            syz_m2 = syz_m1; syz_m1 = syz_0; syz_0 = L.syz_p1;
            szz_m1 = szz_0; szz_0 = szz_p1; szz_p1 = L.szz_p2;
            rho0 = L.rho_kp;
            L = N;
        }
    } else if (ka1 >= k0) {
        const long long c2 = g2idx(g, i, j);
        const int kft = a.m.kfs_top[c2], kfb = a.m.kfs_bot[c2];
        const int kot = a.m.kob_top[c2], kobb = a.m.kob_bot[c2];
        for (int k = k0; k <= ka1; ++k, p += P) {
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

            st1<true>(Vx + p, ld1<true>((OOP ? a.fin.Vx : Vx) + p) + 2.0f / (rho0 + rho[p + 1]) * d3Sx3 * dt);
            st1<true>(Vy + p, ld1<true>((OOP ? a.fin.Vy : Vy) + p) + 2.0f / (rho0 + rho[p + Y]) * d3Sy3 * dt);
            st1<true>(Vz + p, ld1<true>((OOP ? a.fin.Vz : Vz) + p) + 2.0f / (rho0 + rho_kp) * d3Sz3 * dt);

            sxz_m2 = sxz_m1; sxz_m1 = sxz_0; sxz_0 = sxz_p1;
            syz_m2 = syz_m1; syz_m1 = syz_0; syz_0 = syz_p1;
            szz_m1 = szz_0; szz_0 = szz_p1; szz_p1 = szz_p2;
            rho0 = rho_kp;
        }
    }
    if (kb1 >= kb0) {
        const PmlBox b = a.box[4];
        const long long astride = (long long)b.nxb * b.nyb;
        long long q = b.off + (long long)(j - b.j0) * b.nxb + (i - b.i0) + (long long)(kb0 - b.k0) * astride;
        const float4 gxc = a.gxc[i - 1], gxe = a.gxe[i - 1];
        const float4 gyc = a.gyc[j - 1], gye = a.gye[j - 1];
        for (int k = kb0; k <= kb1; ++k, p += P, q += astride) {
            const float sxz_p1 = Sxz[p + P], syz_p1 = Syz[p + P], szz_p2 = Szz[p + 2 * P];
            const float rho_kp = rho[p + P];
            pml_vel_cell<OOP>(a, p, q, k, szz_0, szz_p1, syz_m1, syz_0, sxz_m1, sxz_0, rho0, rho_kp, gxc, gxe, gyc, gye);
            sxz_m2 = sxz_m1; sxz_m1 = sxz_0; sxz_0 = sxz_p1;
            syz_m2 = syz_m1; syz_m1 = syz_0; syz_0 = syz_p1;
            szz_m1 = szz_0; szz_0 = szz_p1; szz_p1 = szz_p2;
            rho0 = rho_kp;
        }
    }
}

// 4 waves per SIMD: the velocity update fits 128 VGPRs without spilling and
// runs about 6 % faster than at the compiler's default of 3 waves
template <bool OOP, bool PF>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PF ? 3 : 4))) k_vel_int(StepArgs a)
{
    int bx, byy, bz;
    plain_tile(a, bx, byy, bz);
    vel_int_body<OOP, PF>(a, bx, byy, bz);
}

template <bool PF>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PF ? 3 : 4))) k_vel_int_b(const StepArgs *__restrict__ mem, const int4 *__restrict__ tiles)
{
    const int4 t = tiles[xcd_remap(blockIdx.x, gridDim.x)];
    vel_int_body<false, PF>(mem[t.x], t.y, t.z, t.w);
}

// ---------------------------------------------------------------------------
// Velocity and stress in one z-march (variant bit 7).  One launch runs
// kernel__update_vel of step n and kernel__update_stress of step n+1 over the
// interior box, so the stresses cross HBM once for both updates and the new
// velocities reach the stress stencil through LDS instead of HBM.  The old
// fields are read from a.fin and the new ones written to a.f (two buffer sets
// that swap every step), so a tile recomputes the velocities of the cells
// around it from old values that no other tile overwrites.  A tile is FW x TY
// stress columns (TY = 4 SR); its 64 x (TY + 4) velocity columns (two halo
// columns and rows on each side) run two planes ahead of the stress update
// and are kept in a ring of six planes in LDS.  Velocities outside the
// interior box (side slabs, bottom band, outer halo) are read from a.f, where
// the side-column and band launches that precede this one have put them.
// ---------------------------------------------------------------------------
#define FW (BX - 4)

template <int NM, int SR>
__global__ void __launch_bounds__(256) k_fused(StepArgs a)
{
    constexpr int TY = 4 * SR, VR = TY + 4, NS = 6;
    __shared__ float vs[NS][3][VR][BX];
    int bx, byy, bz;
    plain_tile(a, bx, byy, bz);
    const Grid &g = a.g;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i0 = g.ib + bx * FW, j0 = g.jb + byy * TY;
    const int il = min(i0 + FW - 1, g.ie), jl = min(j0 + TY - 1, g.je);  // last stress column and row
    const int k0 = a.zc[bz], k1 = a.zc[bz + 1] - 1;                      // interior planes k0..k1
    const long long P = g.plane;
    const int Y = g.pitch;
    const float dt = a.dt;
    const float *__restrict__ Sxx = a.fin.Sxx;
    const float *__restrict__ Syy = a.fin.Syy;
    const float *__restrict__ Szz = a.fin.Szz;
    const float *__restrict__ Syz = a.fin.Syz;
    const float *__restrict__ Sxz = a.fin.Sxz;
    const float *__restrict__ Sxy = a.fin.Sxy;
    const float *__restrict__ rho = a.m.rho;
    const float *__restrict__ mu = a.m.mu;

    // velocity points of this thread: column iv, rows j0 - 2 + ty + 4 r.
    // vf: 0 not needed by the tile, 1 read from a.f, 2 computed, 3 computed
    // and stored (the tile's own cells)
    const int iv = i0 - 2 + tx;
    const int kq = max(k0 - 2, g.kb);   // first computed plane: the z queues start there
    int vf[SR + 1], vb[SR + 1][4];
    float sxz_m2[SR + 1], sxz_m1[SR + 1], sxz_0[SR + 1], syz_m2[SR + 1], syz_m1[SR + 1], syz_0[SR + 1];
    float szz_m1[SR + 1], szz_0[SR + 1], szz_p1[SR + 1], rho0[SR + 1];
#pragma unroll
    for (int r = 0; r <= SR; ++r) {
        const int jv = j0 - 2 + ty + 4 * r;
        const bool need = iv <= il + 2 && jv <= jl + 2;
        const bool comp = need && iv >= g.ib && iv <= g.ie && jv >= g.jb && jv <= g.je;
        const bool own = comp && iv >= i0 && iv <= il && jv >= j0 && jv <= jl;
        vf[r] = !need ? 0 : !comp ? 1 : own ? 3 : 2;
        vb[r][0] = vb[r][1] = vb[r][2] = vb[r][3] = 0;
        sxz_m2[r] = sxz_m1[r] = sxz_0[r] = syz_m2[r] = syz_m1[r] = syz_0[r] = 0.0f;
        szz_m1[r] = szz_0[r] = szz_p1[r] = rho0[r] = 0.0f;
        if (comp) {
            const long long c2 = g2idx(g, iv, jv);
            vb[r][0] = a.m.kfs_top[c2]; vb[r][1] = a.m.kfs_bot[c2];
            vb[r][2] = a.m.kob_top[c2]; vb[r][3] = a.m.kob_bot[c2];
            const long long p = gidx(g, iv, jv, kq);
            sxz_m2[r] = Sxz[p - 2 * P]; sxz_m1[r] = Sxz[p - P]; sxz_0[r] = Sxz[p];
            syz_m2[r] = Syz[p - 2 * P]; syz_m1[r] = Syz[p - P]; syz_0[r] = Syz[p];
            szz_m1[r] = Szz[p - P]; szz_0[r] = Szz[p]; szz_p1[r] = Szz[p + P];
            rho0[r] = rho[p];
        }
    }

    // stress cells of this thread: column is, rows j0 + ty + 4 s
    const int is = i0 + tx;
    bool sa[SR];
    int sb[SR][4];
    float mu00[SR], mu10[SR], mu01[SR], mu11[SR];
#pragma unroll
    for (int s = 0; s < SR; ++s) {
        const int js = j0 + ty + 4 * s;
        sa[s] = tx < FW && is <= il && js <= jl;
        sb[s][0] = sb[s][1] = sb[s][2] = sb[s][3] = 0;
        mu00[s] = mu10[s] = mu01[s] = mu11[s] = 0.0f;
        if (sa[s]) {
            const long long c2 = g2idx(g, is, js);
            sb[s][0] = a.m.kfs_top[c2]; sb[s][1] = a.m.kfs_bot[c2];
            sb[s][2] = a.m.kob_top[c2]; sb[s][3] = a.m.kob_bot[c2];
            const long long p = gidx(g, is, js, k0);
            mu00[s] = mu[p]; mu10[s] = mu[p + 1]; mu01[s] = mu[p + Y]; mu11[s] = mu[p + 1 + Y];
        }
    }

    for (int kv = k0 - 2; kv <= k1 + 2; ++kv) {
        // velocities of plane kv, kernel__update_vel (m_kernel.f90:75-140)
        const int sl = (kv - k0 + 2) % NS;
#pragma unroll
        for (int r = 0; r <= SR; ++r) {
            if (vf[r] == 0) continue;
            const int row = ty + 4 * r;
            const long long p = gidx(g, iv, j0 - 2 + row, kv);
            float vx, vy, vz;
            if (vf[r] >= 2 && kv >= g.kb && kv <= g.ke) {
                const float sxz_p1 = Sxz[p + P], syz_p1 = Syz[p + P], szz_p2 = Szz[p + 2 * P];
                const float rho_kp = rho[p + P];
                const bool in = in_band(kv, vb[r][0], vb[r][1], vb[r][2], vb[r][3]);
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
                                  + (sxz_0[r] - sxz_m1[r]) * re40z - (sxz_p1 - sxz_m2[r]) * re41z;
                const float d3Sy3 = (sxy_0 - sxy_im1) * re40x - (sxy_ip1 - sxy_im2) * re41x
                                  + (syy_jp1 - syy_0) * re40y - (syy_jp2 - syy_jm1) * re41y
                                  + (syz_0[r] - syz_m1[r]) * re40z - (syz_p1 - syz_m2[r]) * re41z;
                const float d3Sz3 = (sxz_0[r] - sxz_im1) * re40x - (sxz_ip1 - sxz_im2) * re41x
                                  + (syz_0[r] - syz_jm1) * re40y - (syz_jp1 - syz_jm2) * re41y
                                  + (szz_p1[r] - szz_0[r]) * re40z - (szz_p2 - szz_m1[r]) * re41z;

                vx = ld1<true>(a.fin.Vx + p) + 2.0f / (rho0[r] + rho[p + 1]) * d3Sx3 * dt;
                vy = ld1<true>(a.fin.Vy + p) + 2.0f / (rho0[r] + rho[p + Y]) * d3Sy3 * dt;
                vz = ld1<true>(a.fin.Vz + p) + 2.0f / (rho0[r] + rho_kp) * d3Sz3 * dt;
                if (vf[r] == 3 && kv >= k0 && kv <= k1) {
                    st1<true>(a.f.Vx + p, vx);
                    st1<true>(a.f.Vy + p, vy);
                    st1<true>(a.f.Vz + p, vz);
                }
                sxz_m2[r] = sxz_m1[r]; sxz_m1[r] = sxz_0[r]; sxz_0[r] = sxz_p1;
                syz_m2[r] = syz_m1[r]; syz_m1[r] = syz_0[r]; syz_0[r] = syz_p1;
                szz_m1[r] = szz_0[r]; szz_0[r] = szz_p1[r]; szz_p1[r] = szz_p2;
                rho0[r] = rho_kp;
            } else {
                vx = a.f.Vx[p];
                vy = a.f.Vy[p];
                vz = a.f.Vz[p];
            }
            vs[sl][0][row][tx] = vx;
            vs[sl][1][row][tx] = vy;
            vs[sl][2][row][tx] = vz;
        }
        __syncthreads();

        // stresses of plane kv - 2 from the velocities of planes kv - 4 .. kv
        const int k = kv - 2;
        if (k < k0) continue;
        const int s0 = (k - k0 + 2) % NS, sm1 = (k - k0 + 1) % NS, sm2 = (k - k0) % NS;
        const int sp1 = (k - k0 + 3) % NS, sp2 = (k - k0 + 4) % NS;
#pragma unroll
        for (int s = 0; s < SR; ++s) {
            if (!sa[s]) continue;
            const int row = ty + 4 * s + 2, col = tx + 2;
            const int js = j0 + ty + 4 * s;
            const long long p = gidx(g, is, js, k);
            const float mu00n = mu[p + P], mu10n = mu[p + P + 1], mu01n = mu[p + P + Y];
            VStencil v;
            v.vx_m1 = vs[sm1][0][row][col]; v.vx_0 = vs[s0][0][row][col];
            v.vx_p1 = vs[sp1][0][row][col]; v.vx_p2 = vs[sp2][0][row][col];
            v.vy_m1 = vs[sm1][1][row][col]; v.vy_0 = vs[s0][1][row][col];
            v.vy_p1 = vs[sp1][1][row][col]; v.vy_p2 = vs[sp2][1][row][col];
            v.vz_m2 = vs[sm2][2][row][col]; v.vz_m1 = vs[sm1][2][row][col];
            v.vz_0 = vs[s0][2][row][col]; v.vz_p1 = vs[sp1][2][row][col];
            v.vx_im2 = vs[s0][0][row][col - 2]; v.vx_im1 = vs[s0][0][row][col - 1];
            v.vx_ip1 = vs[s0][0][row][col + 1];
            v.vx_jm1 = vs[s0][0][row - 1][col]; v.vx_jp1 = vs[s0][0][row + 1][col];
            v.vx_jp2 = vs[s0][0][row + 2][col];
            v.vy_jm2 = vs[s0][1][row - 2][col]; v.vy_jm1 = vs[s0][1][row - 1][col];
            v.vy_jp1 = vs[s0][1][row + 1][col];
            v.vy_im1 = vs[s0][1][row][col - 1]; v.vy_ip1 = vs[s0][1][row][col + 1];
            v.vy_ip2 = vs[s0][1][row][col + 2];
            v.vz_im1 = vs[s0][2][row][col - 1]; v.vz_ip1 = vs[s0][2][row][col + 1];
            v.vz_ip2 = vs[s0][2][row][col + 2];
            v.vz_jm1 = vs[s0][2][row - 1][col]; v.vz_jp1 = vs[s0][2][row + 1][col];
            v.vz_jp2 = vs[s0][2][row + 2][col];
            const long long rr = ((long long)(k - g.kb) * g.nyk + (js - g.jb)) * g.rpitch + (is - 1);
            stress_cell<NM, true, true>(a, p, rr, in_band(k, sb[s][0], sb[s][1], sb[s][2], sb[s][3]), v, mu00[s],
                                        mu00n, mu10[s], mu10n, mu01[s], mu01n, mu11[s]);
            mu00[s] = mu00n; mu10[s] = mu10n; mu01[s] = mu01n;
            mu11[s] = mu[p + P + 1 + Y];
        }
    }
}

// ---------------------------------------------------------------------------
// Side columns (kbeg_a = 1, m_global.f90:334-343): y slabs as full-row tiles,
// x slabs of interior rows as tiles na columns wide with several rows per
// wave (tiles from a host list).  Each lane marches its column through the
// tile's k chunk.
// ---------------------------------------------------------------------------
struct PmlLane {
    int i, j, kb, ke;
    long long abase, astride;
};

__device__ static inline bool pml_lane(const StepArgs &a, const int4 tl, PmlLane &L)
{
    const Grid &g = a.g;
    const int w = (tl.x >> 20) & 0xFF;
    const bool xslab = (tl.x >> 28) & 1;
    const int lane = threadIdx.y * blockDim.x + threadIdx.x;
    const int rows = (blockDim.x * blockDim.y) / w;
    if (lane >= rows * w) return false;
    L.i = (tl.x & 0xFFFFF) + lane % w;
    L.j = tl.y + lane / w;
    if (L.i > g.nx || L.j > g.ny) return false;
    const int na = g.na;
    int q;
    if (L.j <= na) q = 0;
    else if (L.j > g.ny - na) q = 1;
    else if (L.i <= na) q = 2;
    else if (L.i > g.nx - na) q = 3;
    else return false;  // interior column: its bottom band belongs to the interior kernels
    // x-slab cells of interior rows belong to the slab tiles, y-slab cells to full-row tiles
    if (xslab != (q >= 2)) return false;
    const PmlBox b = pick_box(a, q);
    L.kb = tl.z;
    L.ke = min(tl.w, g.nz);
    L.astride = (long long)b.nxb * b.nyb;
    L.abase = b.off + (long long)(L.j - b.j0) * b.nxb + (L.i - b.i0) - (long long)b.k0 * L.astride;
    return L.kb <= L.ke;
}

// absorb_p__update_stress, m_absorb_p.f90:448-519
template <bool OOP>
__device__ __forceinline__ void stress_pml_body(const StepArgs &a, const int4 tl)
{
    PmlLane L;
    if (!pml_lane(a, tl, L)) return;
    const Grid &g = a.g;
    const int i = L.i, j = L.j;
    const long long P = g.plane;
    const int Y = g.pitch;
    const float *__restrict__ Vx = a.f.Vx;
    const float *__restrict__ Vy = a.f.Vy;
    const float *__restrict__ Vz = a.f.Vz;
    const float *__restrict__ mu = a.m.mu;
    const float4 gxc = a.gxc[i - 1], gxe = a.gxe[i - 1];
    const float4 gyc = a.gyc[j - 1], gye = a.gye[j - 1];

    long long p = gidx(g, i, j, L.kb);
    long long q = L.abase + (long long)L.kb * L.astride;
    float vx_0 = Vx[p], vy_0 = Vy[p], vz_m1 = Vz[p - P];
    float mu00 = mu[p], mu10 = mu[p + 1], mu01 = mu[p + Y], mu11 = mu[p + 1 + Y];
    for (int k = L.kb; k <= L.ke; ++k, p += P, q += L.astride) {
        const float vx_p1 = Vx[p + P], vy_p1 = Vy[p + P], vz_0 = Vz[p];
        const float mu00n = mu[p + P], mu10n = mu[p + P + 1], mu01n = mu[p + P + Y];
        pml_stress_cell<true, OOP>(a, p, q, k, vx_0, vx_p1, vy_0, vy_p1, vz_m1, vz_0, mu00, mu00n, mu10, mu10n, mu01,
                              mu01n, mu11, gxc, gxe, gyc, gye);
        vx_0 = vx_p1; vy_0 = vy_p1; vz_m1 = vz_0;
        mu00 = mu00n; mu10 = mu10n; mu01 = mu01n;
        mu11 = mu[p + P + 1 + Y];
    }
}

// absorb_p__update_vel, m_absorb_p.f90:261-306
template <bool OOP>
__device__ __forceinline__ void vel_pml_body(const StepArgs &a, const int4 tl)
{
    PmlLane L;
    if (!pml_lane(a, tl, L)) return;
    const Grid &g = a.g;
    const int i = L.i, j = L.j;
    const long long P = g.plane;
    const float *__restrict__ Szz = a.f.Szz;
    const float *__restrict__ Syz = a.f.Syz;
    const float *__restrict__ Sxz = a.f.Sxz;
    const float *__restrict__ rho = a.m.rho;
    const float4 gxc = a.gxc[i - 1], gxe = a.gxe[i - 1];
    const float4 gyc = a.gyc[j - 1], gye = a.gye[j - 1];

    long long p = gidx(g, i, j, L.kb);
    long long q = L.abase + (long long)L.kb * L.astride;
    float szz_0 = Szz[p], syz_m1 = Syz[p - P], sxz_m1 = Sxz[p - P], rho0 = rho[p];
    for (int k = L.kb; k <= L.ke; ++k, p += P, q += L.astride) {
        const float szz_p1 = Szz[p + P], syz_0 = Syz[p], sxz_0 = Sxz[p], rho_kp = rho[p + P];
        pml_vel_cell<OOP>(a, p, q, k, szz_0, szz_p1, syz_m1, syz_0, sxz_m1, sxz_0, rho0, rho_kp, gxc, gxe, gyc, gye);
        szz_0 = szz_p1; syz_m1 = syz_0; sxz_m1 = sxz_0; rho0 = rho_kp;
    }
}

template <bool OOP>
__global__ void __launch_bounds__(256) k_stress_pml(StepArgs a)
{
    stress_pml_body<OOP>(a, a.tiles[xcd_remap(blockIdx.x, gridDim.x)]);
}

template <bool OOP>
__global__ void __launch_bounds__(256) k_vel_pml(StepArgs a)
{
    vel_pml_body<OOP>(a, a.tiles[xcd_remap(blockIdx.x, gridDim.x)]);
}

// batched side-column tiles: (i0 | width << 20 | xslab << 28, j0, k0 | k1 << 16, member)
__device__ static inline int4 unpack_bside(const int4 t) { return make_int4(t.x, t.y, t.z & 0xFFFF, t.z >> 16); }

__global__ void __launch_bounds__(256) k_stress_pml_b(const StepArgs *__restrict__ mem, const int4 *__restrict__ tiles)
{
    const int4 t = tiles[xcd_remap(blockIdx.x, gridDim.x)];
    stress_pml_body<false>(mem[t.w], unpack_bside(t));
}

__global__ void __launch_bounds__(256) k_vel_pml_b(const StepArgs *__restrict__ mem, const int4 *__restrict__ tiles)
{
    const int4 t = tiles[xcd_remap(blockIdx.x, gridDim.x)];
    vel_pml_body<false>(mem[t.w], unpack_bside(t));
}

// One launch per phase for the interior columns and the side columns
// together (variant bit 5): a schedule lists interior tiles (bx, by, bz, -1)
// and side tiles (i0 | width << 20 | xslab << 28, j0, k0 | k1 << 16, 0), the
// x-slab tiles of a k chunk next to the interior tile rows they border, so
// the 128-byte lines the two kinds share are fetched once through the L2 of
// one XCD.  The two kinds update disjoint cells from the other phase's
// fields, so their order does not change any result.
template <int NM, bool NT>
__global__ void __launch_bounds__(256) k_stress_all(StepArgs a, const int4 *__restrict__ sched)
{
    const int4 t = sched[xcd_remap(blockIdx.x, gridDim.x)];
    if (t.w < 0) stress_int_body<NM, NT, false>(a, t.x, t.y, t.z);
    else stress_pml_body<false>(a, make_int4(t.x, t.y, t.z & 0xFFFF, t.z >> 16));
}

template <bool PF>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PF ? 3 : 4))) k_vel_all(StepArgs a, const int4 *__restrict__ sched)
{
    const int4 t = sched[xcd_remap(blockIdx.x, gridDim.x)];
    if (t.w < 0) vel_int_body<false, PF>(a, t.x, t.y, t.z);
    else vel_pml_body<false>(a, make_int4(t.x, t.y, t.z & 0xFFFF, t.z >> 16));
}

// batched form of the merged launch: entries (m, bx, by, bz) are interior
// tiles of member m, entries (-(m+1), i0 | width << 20 | xslab << 28, j0,
// k0 | k1 << 16) its side tiles
template <int NM, bool NT>
__global__ void __launch_bounds__(256) k_stress_all_b(const StepArgs *__restrict__ mem, const int4 *__restrict__ tiles)
{
    const int4 t = tiles[xcd_remap(blockIdx.x, gridDim.x)];
    if (t.x >= 0) stress_int_body<NM, NT, false>(mem[t.x], t.y, t.z, t.w);
    else stress_pml_body<false>(mem[-t.x - 1], make_int4(t.y, t.z, t.w & 0xFFFF, t.w >> 16));
}

template <bool PF>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PF ? 3 : 4))) k_vel_all_b(const StepArgs *__restrict__ mem, const int4 *__restrict__ tiles)
{
    const int4 t = tiles[xcd_remap(blockIdx.x, gridDim.x)];
    if (t.x >= 0) vel_int_body<false, PF>(mem[t.x], t.y, t.z, t.w);
    else vel_pml_body<false>(mem[-t.x - 1], make_int4(t.y, t.z, t.w & 0xFFFF, t.w >> 16));
}

void fill_step_args(swpc_ctx *c, StepArgs &a, int mask);
static void step_args(swpc_ctx *c, StepArgs &a, int mask) { fill_step_args(c, a, mask); }

void fill_step_args(swpc_ctx *c, StepArgs &a, int mask)
{
    a.g = c->g;
    a.f.Vx = c->fields[SWPC_VX]; a.f.Vy = c->fields[SWPC_VY]; a.f.Vz = c->fields[SWPC_VZ];
    a.f.Sxx = c->fields[SWPC_SXX]; a.f.Syy = c->fields[SWPC_SYY]; a.f.Szz = c->fields[SWPC_SZZ];
    a.f.Syz = c->fields[SWPC_SYZ]; a.f.Sxz = c->fields[SWPC_SXZ]; a.f.Sxy = c->fields[SWPC_SXY];
    {
        float *const *in = c->fields_in[0] ? c->fields_in : c->fields;
        a.fin.Vx = in[SWPC_VX]; a.fin.Vy = in[SWPC_VY]; a.fin.Vz = in[SWPC_VZ];
        a.fin.Sxx = in[SWPC_SXX]; a.fin.Syy = in[SWPC_SYY]; a.fin.Szz = in[SWPC_SZZ];
        a.fin.Syz = in[SWPC_SYZ]; a.fin.Sxz = in[SWPC_SXZ]; a.fin.Sxy = in[SWPC_SXY];
    }
    a.m.rho = c->rho; a.m.lam = c->lam; a.m.mu = c->mu; a.m.taup = c->taup; a.m.taus = c->taus;
    a.m.kfs_top = c->kfs_top; a.m.kfs_bot = c->kfs_bot; a.m.kob_top = c->kob_top; a.m.kob_bot = c->kob_bot;
    a.m.kob = c->kob;
    a.R = c->R;
    a.v = c->visco;
    a.c = c->coef;
    a.dt = c->desc.dt;
    a.ade = c->ade;
    a.ade_in = c->ade_in ? c->ade_in : c->ade;
    a.npml = c->npml;
    for (int q = 0; q < 5; ++q) a.box[q] = c->box[q];
    a.gxc = c->gxc; a.gxe = c->gxe; a.gyc = c->gyc; a.gye = c->gye; a.gzc = c->gzc; a.gze = c->gze;
    a.r20x = 1.0f / (float)c->desc.dx;
    a.r20y = 1.0f / (float)c->desc.dy;
    a.r20z = 1.0f / (float)c->desc.dz;
    a.tiles = (const int4 *)c->d_pml_tiles;
    a.ntiles = c->npml_tiles;
    a.has_pml = (c->desc.abc == SWPC_ABC_PML && c->npml > 0) ? 1 : 0;
    a.mask = mask;
    a.by = c->by;
    a.zc = c->d_zchunks;
    const Grid &g = c->g;
    // variant bit 2: interior tiles 32 cells wide and twice as many rows, so
    // narrow grids idle fewer lanes at the tile edges
    a.bxw = (c->variant & 4) ? 32 : BX;
    a.byi = a.by * (BX / a.bxw);
    a.bx0 = (g.ib - 1) / a.bxw;
    a.nbx = (g.ie - 1) / a.bxw - a.bx0 + 1;
    a.nby = (g.nyk + a.byi - 1) / a.byi;
    a.nbz = c->nbz;
}

// side-column PML tiles of one phase (y slabs, x slabs)
hipError_t launch_stress_side(swpc_ctx *c, hipStream_t s)
{
    if (c->desc.abc != SWPC_ABC_PML || c->npml == 0 || c->npml_tiles <= 0) return hipSuccess;
    StepArgs a;
    step_args(c, a, MASK_PML);
    if (c->fields_in[0]) hipLaunchKernelGGL(k_stress_pml<true>, dim3(c->npml_tiles), dim3(BX, a.by), 0, s ? s : c->stream, a);
    else hipLaunchKernelGGL(k_stress_pml<false>, dim3(c->npml_tiles), dim3(BX, a.by), 0, s ? s : c->stream, a);
    return hipGetLastError();
}

hipError_t launch_vel_side(swpc_ctx *c, hipStream_t s)
{
    if (c->desc.abc != SWPC_ABC_PML || c->npml == 0 || c->npml_tiles <= 0) return hipSuccess;
    StepArgs a;
    step_args(c, a, MASK_PML);
    if (c->fields_in[0]) hipLaunchKernelGGL(k_vel_pml<true>, dim3(c->npml_tiles), dim3(BX, a.by), 0, s ? s : c->stream, a);
    else hipLaunchKernelGGL(k_vel_pml<false>, dim3(c->npml_tiles), dim3(BX, a.by), 0, s ? s : c->stream, a);
    return hipGetLastError();
}

// interior columns: interior planes (mask & 1) and their bottom PML band (mask & 2)
hipError_t launch_stress_cols(swpc_ctx *c, int mask)
{
    StepArgs a;
    step_args(c, a, mask);
    if (mask == MASK_PML) {   // the bottom band alone: only the chunks that reach below the interior box
        a.zc += c->zband[0];
        a.nbz -= c->zband[0];
    }
    if (c->g.nxk <= 0 || c->g.nyk <= 0 || !((mask & MASK_INT) || a.has_pml)) return hipSuccess;
    dim3 block(a.bxw, a.byi), grid(a.nbx * a.nby * a.nbz);
    const bool nt = (c->variant & 1) != 0;
    const bool oop = c->fields_in[0] != nullptr;   // out of place: non-temporal streams as well
#define SWPC_STRESS_CASE(NMV)                                                                       \
    case NMV:                                                                                       \
        if (oop) hipLaunchKernelGGL((k_stress_int<NMV, true, true>), grid, block, 0, c->stream, a); \
        else if (nt) hipLaunchKernelGGL((k_stress_int<NMV, true, false>), grid, block, 0, c->stream, a); \
        else hipLaunchKernelGGL((k_stress_int<NMV, false, false>), grid, block, 0, c->stream, a); \
        break;
    switch (c->g.nm) {
        SWPC_STRESS_CASE(0) SWPC_STRESS_CASE(1) SWPC_STRESS_CASE(2) SWPC_STRESS_CASE(3) SWPC_STRESS_CASE(4)
        SWPC_STRESS_CASE(5) SWPC_STRESS_CASE(6) SWPC_STRESS_CASE(7) SWPC_STRESS_CASE(8)
    default: return hipErrorInvalidValue;
    }
#undef SWPC_STRESS_CASE
    return hipGetLastError();
}

hipError_t launch_vel_cols(swpc_ctx *c, int mask)
{
    StepArgs a;
    step_args(c, a, mask);
    a.zc = c->d_zchunks + c->nbz + 1;
    if (mask == MASK_PML) {
        a.zc += c->zband[1];
        a.nbz -= c->zband[1];
    }
    if (c->g.nxk <= 0 || c->g.nyk <= 0 || !((mask & MASK_INT) || a.has_pml)) return hipSuccess;
    const dim3 grid(a.nbx * a.nby * a.nbz), block(a.bxw, a.byi);
    const bool pf = (c->variant & 8) != 0;
    if (c->fields_in[0]) {
        if (pf) hipLaunchKernelGGL((k_vel_int<true, true>), grid, block, 0, c->stream, a);
        else hipLaunchKernelGGL((k_vel_int<true, false>), grid, block, 0, c->stream, a);
    } else {
        if (pf) hipLaunchKernelGGL((k_vel_int<false, true>), grid, block, 0, c->stream, a);
        else hipLaunchKernelGGL((k_vel_int<false, false>), grid, block, 0, c->stream, a);
    }
    return hipGetLastError();
}

// the fused interior launch: c->fields holds the new fields, c->fields_in
// the old ones; tiles of FW x 4 sr columns (sr = 2 with variant bit 8)
hipError_t launch_fused(swpc_ctx *c)
{
    StepArgs a;
    step_args(c, a, MASK_INT);
    const Grid &g = c->g;
    if (g.nxk <= 0 || g.nyk <= 0 || c->nfz <= 0 || !c->fields_in[0]) return hipErrorInvalidValue;
    const int sr = (c->variant & 256) ? 2 : 1;
    a.zc = c->d_fzc;
    a.nbx = (g.nxk + FW - 1) / FW;
    a.nby = (g.nyk + 4 * sr - 1) / (4 * sr);
    a.nbz = c->nfz;
    const dim3 grid(a.nbx * a.nby * a.nbz), block(BX, 4);
#define SWPC_FUSED_CASE(NMV)                                                                          \
    case NMV:                                                                                         \
        if (sr == 2) hipLaunchKernelGGL((k_fused<NMV, 2>), grid, block, 0, c->stream, a);              \
        else hipLaunchKernelGGL((k_fused<NMV, 1>), grid, block, 0, c->stream, a);                      \
        break;
    switch (g.nm) {
        SWPC_FUSED_CASE(0) SWPC_FUSED_CASE(1) SWPC_FUSED_CASE(2) SWPC_FUSED_CASE(3) SWPC_FUSED_CASE(4)
        SWPC_FUSED_CASE(5) SWPC_FUSED_CASE(6) SWPC_FUSED_CASE(7) SWPC_FUSED_CASE(8)
    default: return hipErrorInvalidValue;
    }
#undef SWPC_FUSED_CASE
    return hipGetLastError();
}

// schedules of the merged launches, one per phase (see k_stress_all)
int build_schedule(swpc_ctx *c)
{
    for (int ph = 0; ph < 2; ++ph) {
        if (c->d_sched[ph]) (void)hipFree(c->d_sched[ph]);
        c->d_sched[ph] = nullptr;
        c->nsched[ph] = 0;
    }
    const Grid &g = c->g;
    if (c->desc.abc != SWPC_ABC_PML || c->npml == 0 || c->npml_tiles <= 0 || g.nxk <= 0 || g.nyk <= 0 ||
        g.nz > 0x7FFF)
        return SWPC_OK;
    StepArgs a;
    step_args(c, a, MASK_INT | MASK_PML);
    std::vector<int4> side(c->npml_tiles);
    std::vector<int> zc(2 * (c->nbz + 1));
    if (hipMemcpy(side.data(), c->d_pml_tiles, side.size() * sizeof(int4), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(zc.data(), c->d_zchunks, zc.size() * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
        return SWPC_EHIP;
    for (int ph = 0; ph < 2; ++ph) {
        const int *z = zc.data() + ph * (c->nbz + 1);
        std::vector<int4> out;
        std::vector<char> used(side.size(), 0);
        for (int bz = 0; bz < a.nbz; ++bz) {
            const int kb = z[bz], ke = z[bz + 1] - 1;
            std::vector<int> low, high, xs;
            for (size_t n = 0; n < side.size(); ++n) {
                const int4 t = side[n];
                if (used[n] || t.z < kb || t.z > ke) continue;
                used[n] = 1;
                const bool xslab = (t.x >> 28) & 1;
                if (xslab) xs.push_back((int)n);
                else if (t.y <= g.na) low.push_back((int)n);
                else high.push_back((int)n);
            }
            auto emit_side = [&](int n) {
                const int4 t = side[n];
                out.push_back(make_int4(t.x, t.y, t.z | (t.w << 16), 0));
            };
            for (int n : low) emit_side(n);
            size_t q = 0;
            for (int by = 0; by < a.nby; ++by) {
                const int j1 = g.jb + (by + 1) * a.byi - 1;
                while (q < xs.size() && side[xs[q]].y <= j1) emit_side(xs[q++]);
                for (int bx = 0; bx < a.nbx; ++bx) out.push_back(make_int4(bx, by, bz, -1));
            }
            for (; q < xs.size(); ++q) emit_side(xs[q]);
            for (int n : high) emit_side(n);
        }
        for (size_t n = 0; n < side.size(); ++n)
            if (!used[n]) {
                const int4 t = side[n];
                out.push_back(make_int4(t.x, t.y, t.z | (t.w << 16), 0));
            }
        if (hipMalloc((void **)&c->d_sched[ph], out.size() * sizeof(int4)) != hipSuccess) return SWPC_ENOMEM;
        if (hipMemcpy(c->d_sched[ph], out.data(), out.size() * sizeof(int4), hipMemcpyHostToDevice) != hipSuccess)
            return SWPC_EHIP;
        c->nsched[ph] = (int)out.size();
    }
    return SWPC_OK;
}

// both kinds of columns of a phase in one launch, when the schedule exists
// and the step is in place
bool merged_launch_ready(const swpc_ctx *c)
{
    return (c->variant & 32) && c->d_sched[0] && c->d_sched[1] && !c->fields_in[0];
}

hipError_t launch_stress_all(swpc_ctx *c)
{
    StepArgs a;
    step_args(c, a, MASK_INT | MASK_PML);
    const dim3 grid(c->nsched[0]), block(a.bxw, a.byi);
    const bool nt = (c->variant & 1) != 0;
#define SWPC_ALL_CASE(NMV)                                                                                 \
    case NMV:                                                                                              \
        if (nt) hipLaunchKernelGGL((k_stress_all<NMV, true>), grid, block, 0, c->stream, a, c->d_sched[0]); \
        else hipLaunchKernelGGL((k_stress_all<NMV, false>), grid, block, 0, c->stream, a, c->d_sched[0]);   \
        break;
    switch (c->g.nm) {
        SWPC_ALL_CASE(0) SWPC_ALL_CASE(1) SWPC_ALL_CASE(2) SWPC_ALL_CASE(3) SWPC_ALL_CASE(4)
        SWPC_ALL_CASE(5) SWPC_ALL_CASE(6) SWPC_ALL_CASE(7) SWPC_ALL_CASE(8)
    default: return hipErrorInvalidValue;
    }
#undef SWPC_ALL_CASE
    return hipGetLastError();
}

hipError_t launch_vel_all(swpc_ctx *c)
{
    StepArgs a;
    step_args(c, a, MASK_INT | MASK_PML);
    a.zc = c->d_zchunks + c->nbz + 1;
    const dim3 grid(c->nsched[1]), block(a.bxw, a.byi);
    if (c->variant & 8) hipLaunchKernelGGL((k_vel_all<true>), grid, block, 0, c->stream, a, c->d_sched[1]);
    else hipLaunchKernelGGL((k_vel_all<false>), grid, block, 0, c->stream, a, c->d_sched[1]);
    return hipGetLastError();
}

hipError_t launch_stress(swpc_ctx *c, int mask, hipStream_t pml_stream)
{
    hipError_t e;
    if ((mask & MASK_PML) && (e = launch_stress_side(c, pml_stream)) != hipSuccess) return e;
    return launch_stress_cols(c, mask);
}

hipError_t launch_vel(swpc_ctx *c, int mask, hipStream_t pml_stream)
{
    hipError_t e;
    if ((mask & MASK_PML) && (e = launch_vel_side(c, pml_stream)) != hipSuccess) return e;
    return launch_vel_cols(c, mask);
}

hipError_t launch_kernel_stress(swpc_ctx *c) { return launch_stress(c, MASK_INT, nullptr); }
hipError_t launch_pml_stress(swpc_ctx *c) { return launch_stress(c, MASK_PML, nullptr); }
hipError_t launch_kernel_vel(swpc_ctx *c) { return launch_vel(c, MASK_INT, nullptr); }
hipError_t launch_pml_vel(swpc_ctx *c) { return launch_vel(c, MASK_PML, nullptr); }

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

__device__ static inline void cerjan_stress_cell(const CerjanArgs &a, int i, int j, int k)
{
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

__global__ void __launch_bounds__(256) k_cerjan_stress(CerjanArgs a)
{
    cerjan_stress_cell(a, 1 + blockIdx.x * blockDim.x + threadIdx.x, 1 + blockIdx.y, 1 + blockIdx.z);
}

__device__ static inline void cerjan_vel_cell(const CerjanArgs &a, int i, int j, int k)
{
    if (i > a.g.nx || k > a.kend) return;
    const long long p = gidx(a.g, i, j, k);
    const float gxc = a.gx_c[i + 2], gxb = a.gx_b[i + 2];
    const float gyc = a.gy_c[j + 2], gyb = a.gy_b[j + 2];
    const float gzc = a.gz_c[k + 2], gzb = a.gz_b[k + 2];
    a.f.Vx[p] = a.f.Vx[p] * gxb * gyc * gzc;
    a.f.Vy[p] = a.f.Vy[p] * gxc * gyb * gzc;
    a.f.Vz[p] = a.f.Vz[p] * gxc * gyc * gzb;
}

__global__ void __launch_bounds__(256) k_cerjan_vel(CerjanArgs a)
{
    cerjan_vel_cell(a, 1 + blockIdx.x * blockDim.x + threadIdx.x, 1 + blockIdx.y, 1 + blockIdx.z);
}

// batched sponges (swpc_batch): one flat launch over every member's blocks;
// member m owns blocks blk0[m] .. blk0[m+1]-1, its (row chunk, j, k) blocks
// in the order of its own launch
__global__ void __launch_bounds__(256) k_cerjan_b(const CerjanArgs *__restrict__ args, const int *__restrict__ blk0,
                                                  int nmem, int phase)
{
    const int b = blockIdx.x;
    int lo = 0, hi = nmem - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (blk0[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    const CerjanArgs &a = args[lo];
    const int lb = b - blk0[lo];
    const int gx = (a.g.nx + 255) / 256;
    const int t = lb / gx;
    const int i = 1 + (lb % gx) * 256 + threadIdx.x, j = 1 + t % a.g.ny, k = 1 + t / a.g.ny;
    if (phase == 0) cerjan_stress_cell(a, i, j, k);
    else cerjan_vel_cell(a, i, j, k);
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

hipError_t launch_cerjan_stress(swpc_ctx *c, hipStream_t s)
{
    CerjanArgs a;
    cerjan_fill(c, a, c->g.ke);
    dim3 grid((c->g.nx + 255) / 256, c->g.ny, c->g.ke);
    hipLaunchKernelGGL(k_cerjan_stress, grid, dim3(256), 0, (s ? s : c->stream), a);
    return hipGetLastError();
}

hipError_t launch_cerjan_vel(swpc_ctx *c, hipStream_t s)
{
    CerjanArgs a;
    cerjan_fill(c, a, c->g.nz);
    dim3 grid((c->g.nx + 255) / 256, c->g.ny, c->g.nz);
    hipLaunchKernelGGL(k_cerjan_vel, grid, dim3(256), 0, (s ? s : c->stream), a);
    return hipGetLastError();
}

size_t batch_cerjan_size() { return sizeof(CerjanArgs); }

int batch_cerjan_fill(swpc_ctx *c, int phase, void *out)
{
    if (c->desc.abc != SWPC_ABC_CERJAN) return 0;
    CerjanArgs a;
    const int kend = phase == 0 ? c->g.ke : c->g.nz;
    cerjan_fill(c, a, kend);
    *(CerjanArgs *)out = a;
    return (c->g.nx + 255) / 256 * c->g.ny * kend;
}

hipError_t launch_batch_cerjan(int phase, const void *d_args, const int *d_blk0, int nmem, int nblocks, hipStream_t s)
{
    if (nmem <= 0 || nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_cerjan_b, dim3(nblocks), dim3(256), 0, s, (const CerjanArgs *)d_args, d_blk0, nmem, phase);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// batched launches: the arguments of every member live in HBM, the tile
// tables carry the member index (swpc_batch, batch.hip)
// ---------------------------------------------------------------------------
size_t step_args_size() { return sizeof(StepArgs); }

// arguments of one member for the stress (phase 0) or velocity (phase 1)
// launches of a whole step
void fill_batch_args(swpc_ctx *c, void *out, int phase)
{
    StepArgs a;
    step_args(c, a, MASK_INT | MASK_PML);
    if (phase == 1) a.zc = c->d_zchunks + c->nbz + 1;
    *(StepArgs *)out = a;
}

// interior-column grid of one member: tiles (bx, by, bz)
void member_tiles(swpc_ctx *c, int &nbx, int &nby, int &nbz)
{
    StepArgs a;
    step_args(c, a, MASK_INT | MASK_PML);
    nbx = a.nbx;
    nby = a.nby;
    nbz = a.nbz;
}

// one launch per phase for both kinds of tiles (the batch's packing 2)
hipError_t launch_batch_phase_all(int phase, int nm, int variant, int by, int bxw, const void *d_args,
                                  const int4 *d_all, int nall, hipStream_t s)
{
    if (nall <= 0) return hipSuccess;
    const StepArgs *mem = (const StepArgs *)d_args;
    const dim3 block(bxw, by * (BX / bxw));
    const bool nt = (variant & 1) != 0;
    if (phase == 1) {
        if (variant & 8) hipLaunchKernelGGL(k_vel_all_b<true>, dim3(nall), block, 0, s, mem, d_all);
        else hipLaunchKernelGGL(k_vel_all_b<false>, dim3(nall), block, 0, s, mem, d_all);
        return hipGetLastError();
    }
#define SWPC_BALL_CASE(NMV)                                                                           \
    case NMV:                                                                                         \
        if (nt) hipLaunchKernelGGL((k_stress_all_b<NMV, true>), dim3(nall), block, 0, s, mem, d_all); \
        else hipLaunchKernelGGL((k_stress_all_b<NMV, false>), dim3(nall), block, 0, s, mem, d_all);   \
        break;
    switch (nm) {
        SWPC_BALL_CASE(0) SWPC_BALL_CASE(1) SWPC_BALL_CASE(2) SWPC_BALL_CASE(3) SWPC_BALL_CASE(4)
        SWPC_BALL_CASE(5) SWPC_BALL_CASE(6) SWPC_BALL_CASE(7) SWPC_BALL_CASE(8)
    default: return hipErrorInvalidValue;
    }
#undef SWPC_BALL_CASE
    return hipGetLastError();
}

hipError_t launch_batch_phase(int phase, int nm, int variant, int by, int bxw, const void *d_args, const int4 *d_cols,
                              int ncols, const int4 *d_side, int nside, hipStream_t s)
{
    const bool nt = (variant & 1) != 0;
    const StepArgs *mem = (const StepArgs *)d_args;
    const dim3 side(BX, by), block(bxw, by * (BX / bxw));
    if (nside > 0) {
        if (phase == 0) hipLaunchKernelGGL(k_stress_pml_b, dim3(nside), side, 0, s, mem, d_side);
        else hipLaunchKernelGGL(k_vel_pml_b, dim3(nside), side, 0, s, mem, d_side);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (ncols <= 0) return hipSuccess;
    if (phase == 1) {
        if (variant & 8) hipLaunchKernelGGL(k_vel_int_b<true>, dim3(ncols), block, 0, s, mem, d_cols);
        else hipLaunchKernelGGL(k_vel_int_b<false>, dim3(ncols), block, 0, s, mem, d_cols);
        return hipGetLastError();
    }
#define SWPC_BSTRESS_CASE(NMV)                                                                      \
    case NMV:                                                                                       \
        if (nt) hipLaunchKernelGGL((k_stress_int_b<NMV, true>), dim3(ncols), block, 0, s, mem, d_cols); \
        else hipLaunchKernelGGL((k_stress_int_b<NMV, false>), dim3(ncols), block, 0, s, mem, d_cols);   \
        break;
    switch (nm) {
        SWPC_BSTRESS_CASE(0) SWPC_BSTRESS_CASE(1) SWPC_BSTRESS_CASE(2) SWPC_BSTRESS_CASE(3) SWPC_BSTRESS_CASE(4)
        SWPC_BSTRESS_CASE(5) SWPC_BSTRESS_CASE(6) SWPC_BSTRESS_CASE(7) SWPC_BSTRESS_CASE(8)
    default: return hipErrorInvalidValue;
    }
#undef SWPC_BSTRESS_CASE
    return hipGetLastError();
}
