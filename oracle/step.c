/*
 * step.c -- oracle restatement of one time step of OpenSWPC's 3D loop.
 *
 * TEST INFRASTRUCTURE ONLY (see swpc_oracle.h).  Compiled three times
 * (ORC_PREC = 32, 16, 64).  Every routine keeps the reference's loop bounds
 * and the exact evaluation order of each Fortran expression (left to right,
 * with the same kind promotions), so that the FP32 variant rounds exactly as
 * the reference FP32 build compiled without FMA contraction.
 *
 * Parity status (how each function is pinned; the tests named):
 *   orc_kernel_update_vel / _stress, orc_absorb_p_update_*, orc_kernel_vmax,
 *   orc_source_stressglut (normal components)
 *       - the reference's own example/example.out: all 20 printed max-|V| rows
 *         of the 384^3 default example reproduced by the mixed variant
 *         (tests/test_oracle_example.py);
 *       - closed-form full-space solutions of an explosion and of a point
 *         force, x <-> y mirror symmetry with free surface, PML and memory
 *         variables, superposition, zero field, attenuation against the
 *         tau-method Q(f), PML reflection against a larger domain with a
 *         no-damping control (tests/test_oracle_physics.py).
 *   orc_source_bodyforce - the point-force closed form above.
 *   orc_source_stressglut, shear components (the /4 split onto the four
 *       staggered neighbours), orc_green_store, orc_green_source and the
 *       velocity and displacement samples of orc_wav_store - exact discrete
 *       reciprocity in FP64 (tests/test_oracle_green.py): each of the six
 *       moment-tensor components and three forces at a cell, recorded at a
 *       station, equals the Green's function stored at that cell by the
 *       reciprocal run forced at the station, to 1e-11, with a control
 *       that a misplaced shear split is detected; mirror symmetry pins the
 *       shear operand pairing as well.
 *   orc_absorb_c_update_* - the taper profile by its closed form, and
 *       its application by the taper at each field's own staggered point
 *       (tests/test_oracle_setup.py, with a control); the FP32 multiply
 *       order of the application is pinned by reading only.
 *   orc_wav_store stress and strain samples - linear-field closed forms
 *       (tests/test_oracle_wav.py): the strain accumulators of a linear
 *       velocity field equal dt times its symmetric gradient in the order
 *       xx yy zz yz xz xy, the stress samples of linear stress fields equal
 *       the field at the cell centre, decimated samples come at the right
 *       steps; a swapped pairing is detected.  The output convention (no
 *       sign flip for strain and stress, unlike Vz) is pinned by reading.
 *   orc_snap_write - linear-field closed forms, index sampling and
 *       accumulation checks (tests/test_oracle_snap.py).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "swpc_oracle.h"

#define V3(a, k, i, j) (a)[orc_i3(s, (k), (i), (j))]
#define M2(a, i, j) (a)[orc_i2(s, (i), (j))]

/* memory-variable index (m,k,i,j) on the kernel box, m 1-based
   (src/swpc_3d/m_kernel.f90:392-397) */
static inline int64_t ridx(const orc_state *s, int m, int k, int i, int j)
{
    int64_t nzk = s->kend_k - s->kbeg_k + 1, nxk = s->iend_k - s->ibeg_k + 1;
    return (int64_t)(m - 1) +
           (int64_t)s->nm * ((k - s->kbeg_k) + nzk * ((i - s->ibeg_k) + nxk * (int64_t)(j - s->jbeg_k)));
}
#define RV(a, m, k, i, j) (a)[ridx(s, (m), (k), (i), (j))]

/* PML auxiliary index on (kbeg_min:nz, 1:nx, 1:ny) (m_absorb_p.f90:99-116) */
static inline int64_t aidx(const orc_state *s, int k, int i, int j)
{
    int64_t nza = s->nz - s->kbeg_min + 1;
    return (k - s->kbeg_min) + nza * ((int64_t)(i - 1) + (int64_t)s->nx * (j - 1));
}
#define AV(a, k, i, j) (a)[aidx(s, (k), (i), (j))]
#define G(a, c, n) (a)[4 * ((n) - 1) + ((c) - 1)]

/* 4th-order coefficient pairs (m_kernel.f90:43-54) and the per-cell band
   switch re4* = rc4* + isign * rd4* (m_kernel.f90:103-111).  isign is +1
   inside the free-surface / ocean-bottom band, -1 outside. */
typedef struct {
    mp_t rc40x, rc41x, rc40y, rc41y, rc40z, rc41z;
    mp_t rd40x, rd41x, rd40y, rd41y, rd40z, rd41z;
} coef4;

static coef4 make_coef4(const orc_state *s)
{
    coef4 c;
    mp_t dx = (mp_t)s->dx, dy = (mp_t)s->dy, dz = (mp_t)s->dz;
    c.rc40x = (mp_t)17.0 / (mp_t)16.0 / dx;
    c.rc40y = (mp_t)17.0 / (mp_t)16.0 / dy;
    c.rc40z = (mp_t)17.0 / (mp_t)16.0 / dz;
    c.rc41x = (mp_t)1.0 / (mp_t)48.0 / dx;
    c.rc41y = (mp_t)1.0 / (mp_t)48.0 / dy;
    c.rc41z = (mp_t)1.0 / (mp_t)48.0 / dz;
    c.rd40x = (mp_t)-1.0 / (mp_t)16.0 / dx;
    c.rd40y = (mp_t)-1.0 / (mp_t)16.0 / dy;
    c.rd40z = (mp_t)-1.0 / (mp_t)16.0 / dz;
    c.rd41x = (mp_t)-1.0 / (mp_t)48.0 / dx;
    c.rd41y = (mp_t)-1.0 / (mp_t)48.0 / dy;
    c.rd41z = (mp_t)-1.0 / (mp_t)48.0 / dz;
    return c;
}

static inline int band_sign(const orc_state *s, int k, int i, int j)
{
    /* isign = sign(1, max((k-kfs_top)*(kfs_bot-k), (k-kob_top)*(kob_bot-k))) */
    int a = (k - M2(s->kfs_top, i, j)) * (M2(s->kfs_bot, i, j) - k);
    int b = (k - M2(s->kob_top, i, j)) * (M2(s->kob_bot, i, j) - k);
    int m = a > b ? a : b;
    return m >= 0 ? 1 : -1;
}

/* kernel__update_vel, src/swpc_3d/m_kernel.f90:75-140 */
void orc_kernel_update_vel(orc_state *s)
{
    const coef4 c = make_coef4(s);
    const mp_t dt = (mp_t)s->dt;
#pragma omp parallel for schedule(static)
    for (int j = s->jbeg_k; j <= s->jend_k; j++) {
        for (int i = s->ibeg_k; i <= s->iend_k; i++) {
            for (int k = s->kbeg_k; k <= s->kend_k; k++) {
                const int isign = band_sign(s, k, i, j);
                const mp_t re40x = c.rc40x + (mp_t)isign * c.rd40x;
                const mp_t re41x = c.rc41x + (mp_t)isign * c.rd41x;
                const mp_t re40y = c.rc40y + (mp_t)isign * c.rd40y;
                const mp_t re41y = c.rc41y + (mp_t)isign * c.rd41y;
                const mp_t re40z = c.rc40z + (mp_t)isign * c.rd40z;
                const mp_t re41z = c.rc41z + (mp_t)isign * c.rd41z;

                const mp_t d3Sx3 = (V3(s->Sxx, k, i + 1, j) - V3(s->Sxx, k, i, j)) * re40x
                                 - (V3(s->Sxx, k, i + 2, j) - V3(s->Sxx, k, i - 1, j)) * re41x
                                 + (V3(s->Sxy, k, i, j) - V3(s->Sxy, k, i, j - 1)) * re40y
                                 - (V3(s->Sxy, k, i, j + 1) - V3(s->Sxy, k, i, j - 2)) * re41y
                                 + (V3(s->Sxz, k, i, j) - V3(s->Sxz, k - 1, i, j)) * re40z
                                 - (V3(s->Sxz, k + 1, i, j) - V3(s->Sxz, k - 2, i, j)) * re41z;
                const mp_t d3Sy3 = (V3(s->Sxy, k, i, j) - V3(s->Sxy, k, i - 1, j)) * re40x
                                 - (V3(s->Sxy, k, i + 1, j) - V3(s->Sxy, k, i - 2, j)) * re41x
                                 + (V3(s->Syy, k, i, j + 1) - V3(s->Syy, k, i, j)) * re40y
                                 - (V3(s->Syy, k, i, j + 2) - V3(s->Syy, k, i, j - 1)) * re41y
                                 + (V3(s->Syz, k, i, j) - V3(s->Syz, k - 1, i, j)) * re40z
                                 - (V3(s->Syz, k + 1, i, j) - V3(s->Syz, k - 2, i, j)) * re41z;
                const mp_t d3Sz3 = (V3(s->Sxz, k, i, j) - V3(s->Sxz, k, i - 1, j)) * re40x
                                 - (V3(s->Sxz, k, i + 1, j) - V3(s->Sxz, k, i - 2, j)) * re41x
                                 + (V3(s->Syz, k, i, j) - V3(s->Syz, k, i, j - 1)) * re40y
                                 - (V3(s->Syz, k, i, j + 1) - V3(s->Syz, k, i, j - 2)) * re41y
                                 + (V3(s->Szz, k + 1, i, j) - V3(s->Szz, k, i, j)) * re40z
                                 - (V3(s->Szz, k + 2, i, j) - V3(s->Szz, k - 1, i, j)) * re41z;

                const sp_t r0 = V3(s->rho, k, i, j);
                V3(s->Vx, k, i, j) = V3(s->Vx, k, i, j) + (sp_t)2.0 / (r0 + V3(s->rho, k, i + 1, j)) * d3Sx3 * dt;
                V3(s->Vy, k, i, j) = V3(s->Vy, k, i, j) + (sp_t)2.0 / (r0 + V3(s->rho, k, i, j + 1)) * d3Sy3 * dt;
                V3(s->Vz, k, i, j) = V3(s->Vz, k, i, j) + (sp_t)2.0 / (r0 + V3(s->rho, k + 1, i, j)) * d3Sz3 * dt;
            }
        }
    }
}

/* harmonic average of mu over four cells (m_kernel.f90:293-309) */
static inline sp_t hmean4(sp_t a, sp_t b, sp_t c, sp_t d, sp_t epsl)
{
    return (sp_t)4 * a * b * c * d / (a * b * c + a * b * d + a * c * d + b * c * d + epsl);
}

/* kernel__update_stress, src/swpc_3d/m_kernel.f90:142-348 */
void orc_kernel_update_stress(orc_state *s)
{
    const coef4 c = make_coef4(s);
    const sp_t dt = (sp_t)s->dt;
    const mp_t dtm = (mp_t)s->dt;
    const sp_t epsl = (sp_t)1.1920928955078125e-07; /* epsilon(1.0), default real */
    const int nm = s->nm;
    const sp_t d2 = s->d2;

    /* normal components (m_kernel.f90:179-242) */
#pragma omp parallel for schedule(static)
    for (int j = s->jbeg_k; j <= s->jend_k; j++) {
        for (int i = s->ibeg_k; i <= s->iend_k; i++) {
            for (int k = s->kbeg_k; k <= s->kend_k; k++) {
                const int isign = band_sign(s, k, i, j);
                const mp_t re40x = c.rc40x + (mp_t)isign * c.rd40x;
                const mp_t re41x = c.rc41x + (mp_t)isign * c.rd41x;
                const mp_t re40y = c.rc40y + (mp_t)isign * c.rd40y;
                const mp_t re41y = c.rc41y + (mp_t)isign * c.rd41y;
                const mp_t re40z = c.rc40z + (mp_t)isign * c.rd40z;
                const mp_t re41z = c.rc41z + (mp_t)isign * c.rd41z;

                const mp_t dxVx = (V3(s->Vx, k, i, j) - V3(s->Vx, k, i - 1, j)) * re40x
                                - (V3(s->Vx, k, i + 1, j) - V3(s->Vx, k, i - 2, j)) * re41x;
                const mp_t dyVy = (V3(s->Vy, k, i, j) - V3(s->Vy, k, i, j - 1)) * re40y
                                - (V3(s->Vy, k, i, j + 1) - V3(s->Vy, k, i, j - 2)) * re41y;
                const mp_t dzVz = (V3(s->Vz, k, i, j) - V3(s->Vz, k - 1, i, j)) * re40z
                                - (V3(s->Vz, k + 1, i, j) - V3(s->Vz, k - 2, i, j)) * re41z;

                const sp_t mu2 = (sp_t)2 * V3(s->mu, k, i, j);
                const sp_t lam2mu = V3(s->lam, k, i, j) + mu2;
                const sp_t taup1 = V3(s->taup, k, i, j);
                const sp_t taus1 = V3(s->taus, k, i, j);

                const sp_t d3v3 = (sp_t)(dxVx + dyVy + dzVz);
                const sp_t dyVy_dzVz = (sp_t)(dyVy + dzVz);
                const sp_t dxVx_dzVz = (sp_t)(dxVx + dzVz);
                const sp_t dxVx_dyVy = (sp_t)(dxVx + dyVy);

                sp_t Rxx_n = 0, Ryy_n = 0, Rzz_n = 0;
                for (int m = 1; m <= nm; m++) {
                    const sp_t c1 = s->c1[m - 1], c2 = s->c2[m - 1], d1 = s->d1[m - 1];
                    sp_t *rxx = &RV(s->Rxx, m, k, i, j), *ryy = &RV(s->Ryy, m, k, i, j), *rzz = &RV(s->Rzz, m, k, i, j);
                    *rxx = c1 * *rxx - c2 * (lam2mu * taup1 * d3v3 - mu2 * taus1 * dyVy_dzVz) * dt;
                    *ryy = c1 * *ryy - c2 * (lam2mu * taup1 * d3v3 - mu2 * taus1 * dxVx_dzVz) * dt;
                    *rzz = c1 * *rzz - c2 * (lam2mu * taup1 * d3v3 - mu2 * taus1 * dxVx_dyVy) * dt;
                    Rxx_n = Rxx_n + d1 * *rxx;
                    Ryy_n = Ryy_n + d1 * *ryy;
                    Rzz_n = Rzz_n + d1 * *rzz;
                }

                const sp_t taup_plus1 = (sp_t)1 + taup1 * ((sp_t)1 + d2);
                const sp_t taus_plus1 = (sp_t)1 + taus1 * ((sp_t)1 + d2);

                V3(s->Sxx, k, i, j) = V3(s->Sxx, k, i, j) + (lam2mu * taup_plus1 * d3v3 - mu2 * taus_plus1 * dyVy_dzVz + Rxx_n) * dt;
                V3(s->Syy, k, i, j) = V3(s->Syy, k, i, j) + (lam2mu * taup_plus1 * d3v3 - mu2 * taus_plus1 * dxVx_dzVz + Ryy_n) * dt;
                V3(s->Szz, k, i, j) = V3(s->Szz, k, i, j) + (lam2mu * taup_plus1 * d3v3 - mu2 * taus_plus1 * dxVx_dyVy + Rzz_n) * dt;
            }
        }
    }

    /* shear components (m_kernel.f90:266-337) */
#pragma omp parallel for schedule(static)
    for (int j = s->jbeg_k; j <= s->jend_k; j++) {
        for (int i = s->ibeg_k; i <= s->iend_k; i++) {
            for (int k = s->kbeg_k; k <= s->kend_k; k++) {
                const int isign = band_sign(s, k, i, j);
                const mp_t re40x = c.rc40x + (mp_t)isign * c.rd40x;
                const mp_t re41x = c.rc41x + (mp_t)isign * c.rd41x;
                const mp_t re40y = c.rc40y + (mp_t)isign * c.rd40y;
                const mp_t re41y = c.rc41y + (mp_t)isign * c.rd41y;
                const mp_t re40z = c.rc40z + (mp_t)isign * c.rd40z;
                const mp_t re41z = c.rc41z + (mp_t)isign * c.rd41z;

                const mp_t dxVy_dyVx = (V3(s->Vy, k, i + 1, j) - V3(s->Vy, k, i, j)) * re40x
                                     - (V3(s->Vy, k, i + 2, j) - V3(s->Vy, k, i - 1, j)) * re41x
                                     + (V3(s->Vx, k, i, j + 1) - V3(s->Vx, k, i, j)) * re40y
                                     - (V3(s->Vx, k, i, j + 2) - V3(s->Vx, k, i, j - 1)) * re41y;
                const mp_t dxVz_dzVx = (V3(s->Vz, k, i + 1, j) - V3(s->Vz, k, i, j)) * re40x
                                     - (V3(s->Vz, k, i + 2, j) - V3(s->Vz, k, i - 1, j)) * re41x
                                     + (V3(s->Vx, k + 1, i, j) - V3(s->Vx, k, i, j)) * re40z
                                     - (V3(s->Vx, k + 2, i, j) - V3(s->Vx, k - 1, i, j)) * re41z;
                const mp_t dyVz_dzVy = (V3(s->Vz, k, i, j + 1) - V3(s->Vz, k, i, j)) * re40y
                                     - (V3(s->Vz, k, i, j + 2) - V3(s->Vz, k, i, j - 1)) * re41y
                                     + (V3(s->Vy, k + 1, i, j) - V3(s->Vy, k, i, j)) * re40z
                                     - (V3(s->Vy, k + 2, i, j) - V3(s->Vy, k - 1, i, j)) * re41z;

                const sp_t muxz = hmean4(V3(s->mu, k, i, j), V3(s->mu, k + 1, i, j),
                                         V3(s->mu, k, i + 1, j), V3(s->mu, k + 1, i + 1, j), epsl);
                const sp_t muxy = hmean4(V3(s->mu, k, i, j), V3(s->mu, k, i + 1, j),
                                         V3(s->mu, k, i, j + 1), V3(s->mu, k, i + 1, j + 1), epsl);
                const sp_t muyz = hmean4(V3(s->mu, k, i, j), V3(s->mu, k + 1, i, j),
                                         V3(s->mu, k, i, j + 1), V3(s->mu, k + 1, i, j + 1), epsl);

                const sp_t taus1 = V3(s->taus, k, i, j);
                sp_t Ryz_n = 0, Rxz_n = 0, Rxy_n = 0;
                for (int m = 1; m <= nm; m++) {
                    const sp_t c1 = s->c1[m - 1], c2 = s->c2[m - 1], d1 = s->d1[m - 1];
                    sp_t *ryz = &RV(s->Ryz, m, k, i, j), *rxz = &RV(s->Rxz, m, k, i, j), *rxy = &RV(s->Rxy, m, k, i, j);
                    *ryz = (sp_t)(c1 * *ryz - c2 * muyz * taus1 * dyVz_dzVy * dtm);
                    *rxz = (sp_t)(c1 * *rxz - c2 * muxz * taus1 * dxVz_dzVx * dtm);
                    *rxy = (sp_t)(c1 * *rxy - c2 * muxy * taus1 * dxVy_dyVx * dtm);
                    Ryz_n = Ryz_n + d1 * *ryz;
                    Rxz_n = Rxz_n + d1 * *rxz;
                    Rxy_n = Rxy_n + d1 * *rxy;
                }
                const sp_t taus_plus1 = (sp_t)1 + taus1 * ((sp_t)1 + d2);
                V3(s->Syz, k, i, j) = V3(s->Syz, k, i, j) + (muyz * taus_plus1 * dyVz_dzVy + Ryz_n) * dtm;
                V3(s->Sxz, k, i, j) = V3(s->Sxz, k, i, j) + (muxz * taus_plus1 * dxVz_dzVx + Rxz_n) * dtm;
                V3(s->Sxy, k, i, j) = V3(s->Sxy, k, i, j) + (muxy * taus_plus1 * dxVy_dyVx + Rxy_n) * dtm;
            }
        }
    }
}

/* absorb_p__update_vel, src/swpc_3d/m_absorb_p.f90:246-317 (pw_mode branch excluded) */
void orc_absorb_p_update_vel(orc_state *s)
{
    const mp_t r20x = (mp_t)1.0 / (mp_t)s->dx;
    const mp_t r20y = (mp_t)1.0 / (mp_t)s->dy;
    const mp_t r20z = (mp_t)1.0 / (mp_t)s->dz;
    const sp_t dt = (sp_t)s->dt;
#pragma omp parallel for schedule(dynamic)
    for (int j = 1; j <= s->ny; j++) {
        for (int i = 1; i <= s->nx; i++) {
            for (int k = M2(s->kbeg_a, i, j); k <= s->nz; k++) {
                const mp_t dxSxx = (V3(s->Sxx, k, i + 1, j) - V3(s->Sxx, k, i, j)) * r20x;
                const mp_t dySyy = (V3(s->Syy, k, i, j + 1) - V3(s->Syy, k, i, j)) * r20y;
                const mp_t dzSzz = (V3(s->Szz, k + 1, i, j) - V3(s->Szz, k, i, j)) * r20z;
                const mp_t dySyz = (V3(s->Syz, k, i, j) - V3(s->Syz, k, i, j - 1)) * r20y;
                const mp_t dzSyz = (V3(s->Syz, k, i, j) - V3(s->Syz, k - 1, i, j)) * r20z;
                const mp_t dxSxz = (V3(s->Sxz, k, i, j) - V3(s->Sxz, k, i - 1, j)) * r20x;
                const mp_t dzSxz = (V3(s->Sxz, k, i, j) - V3(s->Sxz, k - 1, i, j)) * r20z;
                const mp_t dxSxy = (V3(s->Sxy, k, i, j) - V3(s->Sxy, k, i - 1, j)) * r20x;
                const mp_t dySxy = (V3(s->Sxy, k, i, j) - V3(s->Sxy, k, i, j - 1)) * r20y;

                const sp_t r0 = V3(s->rho, k, i, j);
                const sp_t bx = (sp_t)2.0 / (r0 + V3(s->rho, k, i + 1, j));
                const sp_t by = (sp_t)2.0 / (r0 + V3(s->rho, k, i, j + 1));
                const sp_t bz = (sp_t)2.0 / (r0 + V3(s->rho, k + 1, i, j));

                sp_t *axSxx = &AV(s->axSxx, k, i, j), *aySxy = &AV(s->aySxy, k, i, j), *azSxz = &AV(s->azSxz, k, i, j);
                sp_t *axSxy = &AV(s->axSxy, k, i, j), *aySyy = &AV(s->aySyy, k, i, j), *azSyz = &AV(s->azSyz, k, i, j);
                sp_t *axSxz = &AV(s->axSxz, k, i, j), *aySyz = &AV(s->aySyz, k, i, j), *azSzz = &AV(s->azSzz, k, i, j);

                V3(s->Vx, k, i, j) = V3(s->Vx, k, i, j) + bx
                    * (dr_t)(G(s->gxe, 1, i) * dxSxx + G(s->gyc, 1, j) * dySxy + G(s->gzc, 1, k) * dzSxz
                           + G(s->gxe, 2, i) * *axSxx + G(s->gyc, 2, j) * *aySxy + G(s->gzc, 2, k) * *azSxz) * dt;
                V3(s->Vy, k, i, j) = V3(s->Vy, k, i, j) + by
                    * (dr_t)(G(s->gxc, 1, i) * dxSxy + G(s->gye, 1, j) * dySyy + G(s->gzc, 1, k) * dzSyz
                           + G(s->gxc, 2, i) * *axSxy + G(s->gye, 2, j) * *aySyy + G(s->gzc, 2, k) * *azSyz) * dt;
                V3(s->Vz, k, i, j) = V3(s->Vz, k, i, j) + bz
                    * (dr_t)(G(s->gxc, 1, i) * dxSxz + G(s->gyc, 1, j) * dySyz + G(s->gze, 1, k) * dzSzz
                           + G(s->gxc, 2, i) * *axSxz + G(s->gyc, 2, j) * *aySyz + G(s->gze, 2, k) * *azSzz) * dt;

                *axSxx = G(s->gxe, 3, i) * *axSxx + G(s->gxe, 4, i) * (dr_t)dxSxx * dt;
                *aySxy = G(s->gyc, 3, j) * *aySxy + G(s->gyc, 4, j) * (dr_t)dySxy * dt;
                *azSxz = G(s->gzc, 3, k) * *azSxz + G(s->gzc, 4, k) * (dr_t)dzSxz * dt;
                *axSxy = G(s->gxc, 3, i) * *axSxy + G(s->gxc, 4, i) * (dr_t)dxSxy * dt;
                *aySyy = G(s->gye, 3, j) * *aySyy + G(s->gye, 4, j) * (dr_t)dySyy * dt;
                *azSyz = G(s->gzc, 3, k) * *azSyz + G(s->gzc, 4, k) * (dr_t)dzSyz * dt;
                *axSxz = G(s->gxc, 3, i) * *axSxz + G(s->gxc, 4, i) * (dr_t)dxSxz * dt;
                *aySyz = G(s->gyc, 3, j) * *aySyz + G(s->gyc, 4, j) * (dr_t)dySyz * dt;
                *azSzz = G(s->gze, 3, k) * *azSzz + G(s->gze, 4, k) * (dr_t)dzSzz * dt;
            }
        }
    }
}

/* absorb_p__update_stress, src/swpc_3d/m_absorb_p.f90:319-531 (pw_mode branch excluded) */
void orc_absorb_p_update_stress(orc_state *s)
{
    const mp_t r20x = (mp_t)1.0 / (mp_t)s->dx;
    const mp_t r20y = (mp_t)1.0 / (mp_t)s->dy;
    const mp_t r20z = (mp_t)1.0 / (mp_t)s->dz;
    const sp_t dt = (sp_t)s->dt;
    const sp_t epsl = (sp_t)1.1920928955078125e-07;
#pragma omp parallel for schedule(dynamic)
    for (int j = 1; j <= s->ny; j++) {
        for (int i = 1; i <= s->nx; i++) {
            const int kb = M2(s->kbeg_a, i, j);
            for (int k = kb; k <= s->nz; k++) {
                const mp_t dxVx = (V3(s->Vx, k, i, j) - V3(s->Vx, k, i - 1, j)) * r20x;
                const mp_t dyVy = (V3(s->Vy, k, i, j) - V3(s->Vy, k, i, j - 1)) * r20y;
                const mp_t dzVz = (V3(s->Vz, k, i, j) - V3(s->Vz, k - 1, i, j)) * r20z;
                const sp_t lam2mu_R = (V3(s->lam, k, i, j) + (sp_t)2 * V3(s->mu, k, i, j));
                const sp_t lam_R = lam2mu_R - (sp_t)2 * V3(s->mu, k, i, j);
                sp_t *axVx = &AV(s->axVx, k, i, j), *ayVy = &AV(s->ayVy, k, i, j), *azVz = &AV(s->azVz, k, i, j);
                const sp_t dxVx_ade = G(s->gxc, 1, i) * (dr_t)dxVx + G(s->gxc, 2, i) * *axVx;
                const sp_t dyVy_ade = G(s->gyc, 1, j) * (dr_t)dyVy + G(s->gyc, 2, j) * *ayVy;
                const sp_t dzVz_ade = G(s->gzc, 1, k) * (dr_t)dzVz + G(s->gzc, 2, k) * *azVz;

                V3(s->Sxx, k, i, j) = V3(s->Sxx, k, i, j) + (lam2mu_R * dxVx_ade + lam_R * (dyVy_ade + dzVz_ade)) * dt;
                V3(s->Syy, k, i, j) = V3(s->Syy, k, i, j) + (lam2mu_R * dyVy_ade + lam_R * (dxVx_ade + dzVz_ade)) * dt;
                V3(s->Szz, k, i, j) = V3(s->Szz, k, i, j) + (lam2mu_R * dzVz_ade + lam_R * (dxVx_ade + dyVy_ade)) * dt;

                *axVx = G(s->gxc, 3, i) * *axVx + G(s->gxc, 4, i) * (dr_t)dxVx * dt;
                *ayVy = G(s->gyc, 3, j) * *ayVy + G(s->gyc, 4, j) * (dr_t)dyVy * dt;
                *azVz = G(s->gzc, 3, k) * *azVz + G(s->gzc, 4, k) * (dr_t)dzVz * dt;
            }
            for (int k = kb; k <= s->nz; k++) {
                const mp_t dxVy = (V3(s->Vy, k, i + 1, j) - V3(s->Vy, k, i, j)) * r20x;
                const mp_t dxVz = (V3(s->Vz, k, i + 1, j) - V3(s->Vz, k, i, j)) * r20x;
                const mp_t dyVx = (V3(s->Vx, k, i, j + 1) - V3(s->Vx, k, i, j)) * r20y;
                const mp_t dyVz = (V3(s->Vz, k, i, j + 1) - V3(s->Vz, k, i, j)) * r20y;
                const mp_t dzVx = (V3(s->Vx, k + 1, i, j) - V3(s->Vx, k, i, j)) * r20z;
                const mp_t dzVy = (V3(s->Vy, k + 1, i, j) - V3(s->Vy, k, i, j)) * r20z;

                const sp_t muxz = hmean4(V3(s->mu, k, i, j), V3(s->mu, k + 1, i, j),
                                         V3(s->mu, k, i + 1, j), V3(s->mu, k + 1, i + 1, j), epsl);
                const sp_t muxy = hmean4(V3(s->mu, k, i, j), V3(s->mu, k, i + 1, j),
                                         V3(s->mu, k, i, j + 1), V3(s->mu, k, i + 1, j + 1), epsl);
                const sp_t muyz = hmean4(V3(s->mu, k, i, j), V3(s->mu, k + 1, i, j),
                                         V3(s->mu, k, i, j + 1), V3(s->mu, k + 1, i, j + 1), epsl);

                sp_t *ayVx = &AV(s->ayVx, k, i, j), *azVx = &AV(s->azVx, k, i, j);
                sp_t *axVy = &AV(s->axVy, k, i, j), *azVy = &AV(s->azVy, k, i, j);
                sp_t *axVz = &AV(s->axVz, k, i, j), *ayVz = &AV(s->ayVz, k, i, j);

                V3(s->Syz, k, i, j) = V3(s->Syz, k, i, j) + muyz * (G(s->gye, 1, j) * dyVz + G(s->gze, 1, k) * dzVy
                                      + G(s->gye, 2, j) * *ayVz + G(s->gze, 2, k) * *azVy) * dt;
                V3(s->Sxz, k, i, j) = V3(s->Sxz, k, i, j) + muxz * (G(s->gxe, 1, i) * dxVz + G(s->gze, 1, k) * dzVx
                                      + G(s->gxe, 2, i) * *axVz + G(s->gze, 2, k) * *azVx) * dt;
                V3(s->Sxy, k, i, j) = V3(s->Sxy, k, i, j) + muxy * (G(s->gxe, 1, i) * dxVy + G(s->gye, 1, j) * dyVx
                                      + G(s->gxe, 2, i) * *axVy + G(s->gye, 2, j) * *ayVx) * dt;

                *ayVx = G(s->gye, 3, j) * *ayVx + G(s->gye, 4, j) * (dr_t)dyVx * dt;
                *azVx = G(s->gze, 3, k) * *azVx + G(s->gze, 4, k) * (dr_t)dzVx * dt;
                *axVy = G(s->gxe, 3, i) * *axVy + G(s->gxe, 4, i) * (dr_t)dxVy * dt;
                *azVy = G(s->gze, 3, k) * *azVy + G(s->gze, 4, k) * (dr_t)dzVy * dt;
                *axVz = G(s->gxe, 3, i) * *axVz + G(s->gxe, 4, i) * (dr_t)dxVz * dt;
                *ayVz = G(s->gye, 3, j) * *ayVz + G(s->gye, 4, j) * (dr_t)dyVz * dt;
            }
        }
    }
}

/* absorb_c__update_stress, src/swpc_3d/m_absorb_c.f90:113-168 */
void orc_absorb_c_update_stress(orc_state *s)
{
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= s->ny; j++) {
        for (int i = 1; i <= s->nx; i++) {
            for (int k = 1; k <= s->kend_k; k++) {
                const sp_t gcc = s->cgx_c[i + ORC_OFF] * s->cgy_c[j + ORC_OFF] * s->cgz_c[k + ORC_OFF];
                V3(s->Sxx, k, i, j) = V3(s->Sxx, k, i, j) * gcc;
                V3(s->Syy, k, i, j) = V3(s->Syy, k, i, j) * gcc;
                V3(s->Szz, k, i, j) = V3(s->Szz, k, i, j) * gcc;
            }
        }
    }
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= s->ny; j++) {
        for (int i = 1; i <= s->nx; i++) {
            for (int k = 1; k <= s->kend_k; k++) {
                V3(s->Syz, k, i, j) = V3(s->Syz, k, i, j) * s->cgx_c[i + ORC_OFF] * s->cgy_b[j + ORC_OFF] * s->cgz_b[k + ORC_OFF];
                V3(s->Sxz, k, i, j) = V3(s->Sxz, k, i, j) * s->cgx_b[i + ORC_OFF] * s->cgy_c[j + ORC_OFF] * s->cgz_b[k + ORC_OFF];
                V3(s->Sxy, k, i, j) = V3(s->Sxy, k, i, j) * s->cgx_b[i + ORC_OFF] * s->cgy_b[j + ORC_OFF] * s->cgz_c[k + ORC_OFF];
            }
        }
    }
}

/* absorb_c__update_vel, src/swpc_3d/m_absorb_c.f90:170-198 */
void orc_absorb_c_update_vel(orc_state *s)
{
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= s->ny; j++) {
        for (int i = 1; i <= s->nx; i++) {
            for (int k = 1; k <= s->nz; k++) {
                V3(s->Vx, k, i, j) = V3(s->Vx, k, i, j) * s->cgx_b[i + ORC_OFF] * s->cgy_c[j + ORC_OFF] * s->cgz_c[k + ORC_OFF];
                V3(s->Vy, k, i, j) = V3(s->Vy, k, i, j) * s->cgx_c[i + ORC_OFF] * s->cgy_b[j + ORC_OFF] * s->cgz_c[k + ORC_OFF];
                V3(s->Vz, k, i, j) = V3(s->Vz, k, i, j) * s->cgx_c[i + ORC_OFF] * s->cgy_c[j + ORC_OFF] * s->cgz_b[k + ORC_OFF];
            }
        }
    }
}

/* absorb__update_stress / absorb__update_vel dispatch, src/swpc_3d/m_absorb.f90:43-75 */
void orc_absorb_update_stress(orc_state *s)
{
    if (s->abc == 0) orc_absorb_p_update_stress(s);
    else orc_absorb_c_update_stress(s);
}
void orc_absorb_update_vel(orc_state *s)
{
    if (s->abc == 0) orc_absorb_p_update_vel(s);
    else orc_absorb_c_update_vel(s);
}

/*
 * source__stressglut, src/swpc_3d/m_source.f90:776-848.
 * stime is the moment-rate value of each source at t = tbeg + (it-0.5)dt,
 * evaluated on the host (momentrate, src/shared/m_fdtool.f90:472-497).
 * mo_g, M0 and the m_ij are the single-precision set-up values; the
 * normalisation mo = mo_g / M0 (m_source.f90:302) and dt_dxyz
 * (m_source.f90:306) are formed here in MP as the reference does.
 */
void orc_source_stressglut(orc_state *s, int nsrc, const int32_t *isrc, const int32_t *jsrc, const int32_t *ksrc,
                           const float *mo_g, float M0, const float *mxx, const float *myy, const float *mzz,
                           const float *myz, const float *mxz, const float *mxy, const float *stime)
{
    const mp_t dt_dxyz = (mp_t)s->dt / ((mp_t)s->dx * (mp_t)s->dy * (mp_t)s->dz);
    for (int n = 0; n < nsrc; n++) {
        const mp_t mo = (mp_t)mo_g[n] / (mp_t)M0;
        const mp_t sdrop = mo * (sp_t)stime[n] * dt_dxyz;
        const int ii = isrc[n], jj = jsrc[n], kk = ksrc[n];
        V3(s->Sxx, kk, ii, jj) = V3(s->Sxx, kk, ii, jj) - (mp_t)mxx[n] * sdrop;
        V3(s->Syy, kk, ii, jj) = V3(s->Syy, kk, ii, jj) - (mp_t)myy[n] * sdrop;
        V3(s->Szz, kk, ii, jj) = V3(s->Szz, kk, ii, jj) - (mp_t)mzz[n] * sdrop;

        V3(s->Sxy, kk, ii, jj) = V3(s->Sxy, kk, ii, jj) - (mp_t)mxy[n] * sdrop / 4;
        V3(s->Sxy, kk, ii, jj - 1) = V3(s->Sxy, kk, ii, jj - 1) - (mp_t)mxy[n] * sdrop / 4;
        V3(s->Sxy, kk, ii - 1, jj) = V3(s->Sxy, kk, ii - 1, jj) - (mp_t)mxy[n] * sdrop / 4;
        V3(s->Sxy, kk, ii - 1, jj - 1) = V3(s->Sxy, kk, ii - 1, jj - 1) - (mp_t)mxy[n] * sdrop / 4;

        V3(s->Sxz, kk, ii, jj) = V3(s->Sxz, kk, ii, jj) - (mp_t)mxz[n] * sdrop / 4;
        V3(s->Sxz, kk - 1, ii, jj) = V3(s->Sxz, kk - 1, ii, jj) - (mp_t)mxz[n] * sdrop / 4;
        V3(s->Sxz, kk, ii - 1, jj) = V3(s->Sxz, kk, ii - 1, jj) - (mp_t)mxz[n] * sdrop / 4;
        V3(s->Sxz, kk - 1, ii - 1, jj) = V3(s->Sxz, kk - 1, ii - 1, jj) - (mp_t)mxz[n] * sdrop / 4;

        V3(s->Syz, kk, ii, jj) = V3(s->Syz, kk, ii, jj) - (mp_t)myz[n] * sdrop / 4;
        V3(s->Syz, kk - 1, ii, jj) = V3(s->Syz, kk - 1, ii, jj) - (mp_t)myz[n] * sdrop / 4;
        V3(s->Syz, kk, ii, jj - 1) = V3(s->Syz, kk, ii, jj - 1) - (mp_t)myz[n] * sdrop / 4;
        V3(s->Syz, kk - 1, ii, jj - 1) = V3(s->Syz, kk - 1, ii, jj - 1) - (mp_t)myz[n] * sdrop / 4;
    }
}

/* source__bodyforce, src/swpc_3d/m_source.f90:850-898; stime at t = tbeg + it*dt,
   fx_g / M0 normalisation of m_source.f90:298-300 */
void orc_source_bodyforce(orc_state *s, int nsrc, const int32_t *isrc, const int32_t *jsrc, const int32_t *ksrc,
                          const float *fx_g, const float *fy_g, const float *fz_g, float M0, const float *stime)
{
    const mp_t dt_dxyz = (mp_t)s->dt / ((mp_t)s->dx * (mp_t)s->dy * (mp_t)s->dz);
    for (int n = 0; n < nsrc; n++) {
        const mp_t fx = (mp_t)fx_g[n] / (mp_t)M0, fy = (mp_t)fy_g[n] / (mp_t)M0, fz = (mp_t)fz_g[n] / (mp_t)M0;
        const sp_t st = (sp_t)stime[n];
        const int ii = isrc[n], jj = jsrc[n], kk = ksrc[n];
        const sp_t r0 = V3(s->rho, kk, ii, jj);
        V3(s->Vx, kk, ii, jj) = V3(s->Vx, kk, ii, jj) + ((sp_t)2.0 / (r0 + V3(s->rho, kk, ii + 1, jj))) * fx * st * dt_dxyz / 2;
        V3(s->Vx, kk, ii - 1, jj) = V3(s->Vx, kk, ii - 1, jj) + ((sp_t)2.0 / (r0 + V3(s->rho, kk, ii - 1, jj))) * fx * st * dt_dxyz / 2;
        V3(s->Vy, kk, ii, jj) = V3(s->Vy, kk, ii, jj) + ((sp_t)2.0 / (r0 + V3(s->rho, kk, ii, jj + 1))) * fy * st * dt_dxyz / 2;
        V3(s->Vy, kk, ii, jj - 1) = V3(s->Vy, kk, ii, jj - 1) + ((sp_t)2.0 / (r0 + V3(s->rho, kk, ii, jj - 1))) * fy * st * dt_dxyz / 2;
        V3(s->Vz, kk, ii, jj) = V3(s->Vz, kk, ii, jj) + ((sp_t)2.0 / (r0 + V3(s->rho, kk + 1, ii, jj))) * fz * st * dt_dxyz / 2;
        V3(s->Vz, kk - 1, ii, jj) = V3(s->Vz, kk - 1, ii, jj) + ((sp_t)2.0 / (r0 + V3(s->rho, kk - 1, ii, jj))) * fz * st * dt_dxyz / 2;
    }
}

/*
 * wav__store, src/swpc_3d/m_wav.f90:397-625.  Accumulators u(3,nst) and
 * e(6,nst) persist between calls (zero before it == 1).  Outputs are
 * (ntw, ncmp, nst) in Fortran order; flags bit0 = V, bit1 = U, bit2 = stress,
 * bit3 = strain.  M0 and UC are the single-precision output scales.
 */
void orc_wav_store(orc_state *s, int it, int nst, const int32_t *ist, const int32_t *jst, const int32_t *kst,
                   int flags, int ntdec_w, int ntw, float M0, float UC, sp_t *uacc, sp_t *eacc,
                   sp_t *wav_vel, sp_t *wav_disp, sp_t *wav_stress, sp_t *wav_strain)
{
    const sp_t dt = (sp_t)s->dt;
    const sp_t m0 = (sp_t)M0, uc = (sp_t)UC;
    if (nst <= 0) return;
    if (flags & 2) {
        for (int n = 0; n < nst; n++) {
            const int i = ist[n], j = jst[n], k = kst[n];
            uacc[3 * n + 0] = uacc[3 * n + 0] + (dr_t)(V3(s->Vx, k, i, j) + V3(s->Vx, k, i - 1, j)) * (sp_t)0.5 * dt;
            uacc[3 * n + 1] = uacc[3 * n + 1] + (dr_t)(V3(s->Vy, k, i, j) + V3(s->Vy, k, i, j - 1)) * (sp_t)0.5 * dt;
            uacc[3 * n + 2] = uacc[3 * n + 2] - (dr_t)(V3(s->Vz, k, i, j) + V3(s->Vz, k - 1, i, j)) * (sp_t)0.5 * dt;
        }
    }
    if (flags & 8) {
        const mp_t r40x = (mp_t)9.0 / (mp_t)8.0 / (mp_t)s->dx, r40y = (mp_t)9.0 / (mp_t)8.0 / (mp_t)s->dy,
                   r40z = (mp_t)9.0 / (mp_t)8.0 / (mp_t)s->dz;
        const mp_t r41x = (mp_t)1.0 / (mp_t)24.0 / (mp_t)s->dx, r41y = (mp_t)1.0 / (mp_t)24.0 / (mp_t)s->dy,
                   r41z = (mp_t)1.0 / (mp_t)24.0 / (mp_t)s->dz;
        for (int n = 0; n < nst; n++) {
            const int i = ist[n], j = jst[n], k = kst[n];
#define VX(kk, ii, jj) V3(s->Vx, kk, ii, jj)
#define VY(kk, ii, jj) V3(s->Vy, kk, ii, jj)
#define VZ(kk, ii, jj) V3(s->Vz, kk, ii, jj)
            const mp_t dxVx = (VX(k, i, j) - VX(k, i - 1, j)) * r40x - (VX(k, i + 1, j) - VX(k, i - 2, j)) * r41x;
            const mp_t dyVy = (VY(k, i, j) - VY(k, i, j - 1)) * r40y - (VY(k, i, j + 1) - VY(k, i, j - 2)) * r41y;
            const mp_t dzVz = (VZ(k, i, j) - VZ(k - 1, i, j)) * r40z - (VZ(k + 1, i, j) - VZ(k - 2, i, j)) * r41z;
            const mp_t dxVy = ((VY(k, i + 1, j) - VY(k, i, j)) * r40x - (VY(k, i + 2, j) - VY(k, i - 1, j)) * r41x
                             + (VY(k, i + 1, j - 1) - VY(k, i, j - 1)) * r40x - (VY(k, i + 2, j - 1) - VY(k, i - 1, j - 1)) * r41x
                             + (VY(k, i, j) - VY(k, i - 1, j)) * r40x - (VY(k, i + 1, j) - VY(k, i - 2, j)) * r41x
                             + (VY(k, i, j - 1) - VY(k, i - 1, j - 1)) * r40x - (VY(k, i + 1, j - 1) - VY(k, i - 2, j - 1)) * r41x) / (mp_t)4.0;
            const mp_t dxVz = ((VZ(k, i + 1, j) - VZ(k, i, j)) * r40x - (VZ(k, i + 2, j) - VZ(k, i - 1, j)) * r41x
                             + (VZ(k - 1, i + 1, j) - VZ(k - 1, i, j)) * r40x - (VZ(k - 1, i + 2, j) - VZ(k - 1, i - 1, j)) * r41x
                             + (VZ(k, i, j) - VZ(k, i - 1, j)) * r40x - (VZ(k, i + 1, j) - VZ(k, i - 2, j)) * r41x
                             + (VZ(k - 1, i, j) - VZ(k - 1, i - 1, j)) * r40x - (VZ(k - 1, i + 1, j) - VZ(k - 1, i - 2, j)) * r41x) / (mp_t)4.0;
            const mp_t dyVx = ((VX(k, i, j + 1) - VX(k, i, j)) * r40y - (VX(k, i, j + 2) - VX(k, i, j - 1)) * r41y
                             + (VX(k, i - 1, j + 1) - VX(k, i - 1, j)) * r40y - (VX(k, i - 1, j + 2) - VX(k, i - 1, j - 1)) * r41y
                             + (VX(k, i, j) - VX(k, i, j - 1)) * r40y - (VX(k, i, j + 1) - VX(k, i, j - 2)) * r41y
                             + (VX(k, i - 1, j) - VX(k, i - 1, j - 1)) * r40y - (VX(k, i - 1, j + 1) - VX(k, i - 1, j - 2)) * r41y) / (mp_t)4.0;
            const mp_t dyVz = ((VZ(k, i, j + 1) - VZ(k, i, j)) * r40y - (VZ(k, i, j + 2) - VZ(k, i, j - 1)) * r41y
                             + (VZ(k - 1, i, j + 1) - VZ(k - 1, i, j)) * r40y - (VZ(k - 1, i, j + 2) - VZ(k - 1, i, j - 1)) * r41y
                             + (VZ(k, i, j) - VZ(k, i, j - 1)) * r40y - (VZ(k, i, j + 1) - VZ(k, i, j - 2)) * r41y
                             + (VZ(k - 1, i, j) - VZ(k - 1, i, j - 1)) * r40y - (VZ(k - 1, i, j + 1) - VZ(k - 1, i, j - 2)) * r41y) / (mp_t)4.0;
            const mp_t dzVx = ((VX(k + 1, i, j) - VX(k, i, j)) * r40z - (VX(k + 2, i, j) - VX(k - 1, i, j)) * r41z
                             + (VX(k + 1, i - 1, j) - VX(k, i - 1, j)) * r40z - (VX(k + 2, i - 1, j) - VX(k - 1, i - 1, j)) * r41z
                             + (VX(k, i, j) - VX(k - 1, i, j)) * r40z - (VX(k + 1, i, j) - VX(k - 2, i, j)) * r41z
                             + (VX(k, i - 1, j) - VX(k - 1, i - 1, j)) * r40z - (VX(k + 1, i - 1, j) - VX(k - 2, i - 1, j)) * r41z) / (mp_t)4.0;
            const mp_t dzVy = ((VY(k + 1, i, j) - VY(k, i, j)) * r40z - (VY(k + 2, i, j) - VY(k - 1, i, j)) * r41z
                             + (VY(k + 1, i, j - 1) - VY(k, i, j - 1)) * r40z - (VY(k + 2, i, j - 1) - VY(k - 1, i, j - 1)) * r41z
                             + (VY(k, i, j) - VY(k - 1, i, j)) * r40z - (VY(k + 1, i, j) - VY(k - 2, i, j)) * r41z
                             + (VY(k, i, j - 1) - VY(k - 1, i, j - 1)) * r40z - (VY(k + 1, i, j - 1) - VY(k - 2, i, j - 1)) * r41z) / (mp_t)4.0;
#undef VX
#undef VY
#undef VZ
            sp_t *e = &eacc[6 * n];
            e[0] = e[0] + (dr_t)dxVx * dt;
            e[1] = e[1] + (dr_t)dyVy * dt;
            e[2] = e[2] + (dr_t)dzVz * dt;
            e[3] = e[3] + (dr_t)(dyVz + dzVy) / (sp_t)2.0 * dt;
            e[4] = e[4] + (dr_t)(dxVz + dzVx) / (sp_t)2.0 * dt;
            e[5] = e[5] + (dr_t)(dxVy + dyVx) / (sp_t)2.0 * dt;
        }
    }
    if ((it - 1) % ntdec_w != 0) return;
    const int itw = (it - 1) / ntdec_w; /* 0-based */
    if (itw >= ntw) return;
#define OUT(a, nc, c, n) (a)[((int64_t)(n) * (nc) + (c)) * ntw + itw]
    if (flags & 1) {
        for (int n = 0; n < nst; n++) {
            const int i = ist[n], j = jst[n], k = kst[n];
            OUT(wav_vel, 3, 0, n) = (dr_t)(V3(s->Vx, k, i, j) + V3(s->Vx, k, i - 1, j)) / (sp_t)2.0 * m0 * uc * (sp_t)1e9;
            OUT(wav_vel, 3, 1, n) = (dr_t)(V3(s->Vy, k, i, j) + V3(s->Vy, k, i, j - 1)) / (sp_t)2.0 * m0 * uc * (sp_t)1e9;
            OUT(wav_vel, 3, 2, n) = -((dr_t)(V3(s->Vz, k, i, j) + V3(s->Vz, k - 1, i, j)) / (sp_t)2.0 * m0 * uc * (sp_t)1e9);
        }
    }
    if (flags & 2) {
        for (int n = 0; n < nst; n++) {
            for (int c = 0; c < 3; c++)
                OUT(wav_disp, 3, c, n) = (dr_t)uacc[3 * n + c] * m0 * uc * (sp_t)1e9;
        }
    }
    if (flags & 4) {
        for (int n = 0; n < nst; n++) {
            const int i = ist[n], j = jst[n], k = kst[n];
            OUT(wav_stress, 6, 0, n) = (dr_t)V3(s->Sxx, k, i, j) * m0 * uc * (sp_t)1e6;
            OUT(wav_stress, 6, 1, n) = (dr_t)V3(s->Syy, k, i, j) * m0 * uc * (sp_t)1e6;
            OUT(wav_stress, 6, 2, n) = (dr_t)V3(s->Szz, k, i, j) * m0 * uc * (sp_t)1e6;
            OUT(wav_stress, 6, 3, n) = (dr_t)(V3(s->Syz, k, i, j) + V3(s->Syz, k, i, j - 1) + V3(s->Syz, k - 1, i, j)
                                              + V3(s->Syz, k - 1, i, j - 1)) / (sp_t)4.0 * m0 * uc * (sp_t)1e6;
            OUT(wav_stress, 6, 4, n) = (dr_t)(V3(s->Sxz, k, i, j) + V3(s->Sxz, k, i - 1, j) + V3(s->Sxz, k - 1, i, j)
                                              + V3(s->Sxz, k - 1, i - 1, j)) / (sp_t)4.0 * m0 * uc * (sp_t)1e6;
            OUT(wav_stress, 6, 5, n) = (dr_t)(V3(s->Sxy, k, i, j) + V3(s->Sxy, k, i, j - 1) + V3(s->Sxy, k, i - 1, j)
                                              + V3(s->Sxy, k, i - 1, j - 1)) / (sp_t)4.0 * m0 * uc * (sp_t)1e6;
        }
    }
    if (flags & 8) {
        for (int n = 0; n < nst; n++)
            for (int c = 0; c < 6; c++)
                OUT(wav_strain, 6, c, n) = eacc[6 * n + c] * m0 * uc * (sp_t)1e-3f;
    }
#undef OUT
}

/* kernel__vmax, src/swpc_3d/m_kernel.f90:350-374 (max |V| one cell below kob,
   five cells inside the absorber) */
void orc_kernel_vmax(const orc_state *s, float *out3)
{
    const int margin = 5;
    float xm = 0, ym = 0, zm = 0;
    int j0 = s->na + margin + 1 > s->jbeg_k ? s->na + margin + 1 : s->jbeg_k;
    int j1 = s->ny - s->na - margin < s->jend_k ? s->ny - s->na - margin : s->jend_k;
    int i0 = s->na + margin + 1 > s->ibeg_k ? s->na + margin + 1 : s->ibeg_k;
    int i1 = s->nx - s->na - margin < s->iend_k ? s->nx - s->na - margin : s->iend_k;
    for (int j = j0; j <= j1; j++)
        for (int i = i0; i <= i1; i++) {
            const int k = M2(s->kob, i, j) + 1;
            float ax = (float)fabs((double)V3(s->Vx, k, i, j));
            float ay = (float)fabs((double)V3(s->Vy, k, i, j));
            float az = (float)fabs((double)V3(s->Vz, k, i, j));
            if (ax > xm) xm = ax;
            if (ay > ym) ym = ay;
            if (az > zm) zm = az;
        }
    out3[0] = xm;
    out3[1] = ym;
    out3[2] = zm;
}

/* size in elements of the arrays the caller allocates */
int64_t orc_n3(const orc_state *s)
{
    return (int64_t)(s->nz + 2 * ORC_HALO) * (s->nx + 2 * ORC_HALO) * (s->ny + 2 * ORC_HALO);
}

int orc_prec(void) { return ORC_PREC; }
int orc_sizeof_mp(void) { return (int)sizeof(mp_t); }
int orc_sizeof_sp(void) { return (int)sizeof(sp_t); }

/*
 * snap__write, src/swpc_3d/m_snap.f90:919-947, with the section buffers of
 * wbuf_{xy,yz,xz,fs,ob}_{ps,v,u} (:981-2189), native-format semantics: a
 * frame is taken when mod(it-1, ntdec_s) == 0.
 *
 * Section s = plane*3 + q, plane 0..4 = xy yz xz fs ob, q 0..2 = ps v u;
 * sw[s] != 0 enables it.  Sizes: nxs = (nx + idec/2)/idec (likewise nys,
 * nzs, :130-132); xy/fs/ob sections are (nxs, nys), yz (nys, nzs), xz
 * (nxs, nzs), section point ii -> grid i = ii*idec - idec/2 (:137).
 * frame[s] holds (ns1, ns2, ncomp, nframe) in Fortran order (ncomp 4 for
 * ps, 3 for v and u).  uacc[plane] (ns1, ns2, 3) accumulates V*dt every call
 * for u sections; vmax[plane], umax[plane] (nxs, nys, 3) are the running
 * maxima of the fs and ob v/u sections (:1697-1707 and likewise).  All
 * persistent buffers start at zero.  kfs is the (i,j) free-surface map.
 */
static inline int snap_n(int n, int dec) { return (n + dec / 2) / dec; }

void orc_snap_write(orc_state *s, int it, const int32_t *kfs, const int32_t *sw, int k0_xy, int i0_yz, int j0_xz,
                    int idec, int jdec, int kdec, int ntdec_s, float M0, float UC, sp_t **frame, sp_t **uacc,
                    sp_t **vmax, sp_t **umax)
{
    const int nxs = snap_n(s->nx, idec), nys = snap_n(s->ny, jdec), nzs = snap_n(s->nz, kdec);
    const mp_t r20x = (mp_t)1.0 / (mp_t)s->dx, r20y = (mp_t)1.0 / (mp_t)s->dy, r20z = (mp_t)1.0 / (mp_t)s->dz;
    const sp_t epsl = (sp_t)1.1920928955078125e-07f;
    const sp_t uc = (sp_t)UC, m0 = (sp_t)M0, dt = (sp_t)s->dt;
    const int out = ((it - 1) % ntdec_s) == 0;
    const int f = (it - 1) / ntdec_s;
    for (int plane = 0; plane < 5; plane++) {
        int ns1, ns2;
        if (plane == 1) { ns1 = nys; ns2 = nzs; }
        else if (plane == 2) { ns1 = nxs; ns2 = nzs; }
        else { ns1 = nxs; ns2 = nys; }
        const int64_t npt = (int64_t)ns1 * ns2;
        for (int q = 0; q < 3; q++) {
            const int sec = plane * 3 + q;
            if (!sw[sec]) continue;
            /* ps and v of xy/yz/xz are evaluated only on output steps; fs/ob
               v every step (max maps); u every step */
            const int every = (q == 2) || (q == 1 && plane >= 3);
            if (!out && !every) continue;
            const int ncomp = q == 0 ? 4 : 3;
            sp_t *fr = out ? frame[sec] + (int64_t)f * npt * ncomp : NULL;
            for (int b = 1; b <= ns2; b++)
                for (int a = 1; a <= ns1; a++) {
                    int i, j, k;
                    switch (plane) {
                    case 0: i = a * idec - idec / 2; j = b * jdec - jdec / 2; k = k0_xy; break;
                    case 1: i = i0_yz; j = a * jdec - jdec / 2; k = b * kdec - kdec / 2; break;
                    case 2: i = a * idec - idec / 2; j = j0_xz; k = b * kdec - kdec / 2; break;
                    case 3: i = a * idec - idec / 2; j = b * jdec - jdec / 2; k = M2(kfs, i, j) + 1; break;
                    default: i = a * idec - idec / 2; j = b * jdec - jdec / 2; k = M2(s->kob, i, j) + 1; break;
                    }
                    const int64_t p = (int64_t)(b - 1) * ns1 + (a - 1);
                    if (q == 0) {
                        /* m_snap.f90:1016-1036 */
                        sp_t div = (dr_t)((V3(s->Vx, k, i, j) - V3(s->Vx, k, i - 1, j)) * r20x
                                        + (V3(s->Vy, k, i, j) - V3(s->Vy, k, i, j - 1)) * r20y
                                        + (V3(s->Vz, k, i, j) - V3(s->Vz, k - 1, i, j)) * r20z);
                        sp_t rot_x = (dr_t)((V3(s->Vz, k, i, j + 1) - V3(s->Vz, k, i, j)) * r20y
                                          - (V3(s->Vy, k + 1, i, j) - V3(s->Vy, k, i, j)) * r20z);
                        sp_t rot_y = (dr_t)((V3(s->Vx, k + 1, i, j) - V3(s->Vx, k, i, j)) * r20z
                                          - (V3(s->Vz, k, i + 1, j) - V3(s->Vz, k, i, j)) * r20x);
                        sp_t rot_z = (dr_t)((V3(s->Vy, k, i + 1, j) - V3(s->Vy, k, i, j)) * r20x
                                          - (V3(s->Vx, k, i, j + 1) - V3(s->Vx, k, i, j)) * r20y);
                        const sp_t lam = V3(s->lam, k, i, j);
                        div = div * lam / (sp_t)fabs((double)(lam + epsl));
                        const mp_t syz = V3(s->Syz, k, i, j), sxz = V3(s->Sxz, k, i, j), sxy = V3(s->Sxy, k, i, j);
                        if (plane == 1) {
                            /* wbuf_yz_ps masks with abs(S + eps) (m_snap.f90:1028-1030); every
                               other plane with abs(S) + eps (:1121-1123) */
                            rot_x = (sp_t)((mp_t)rot_x * (mp_t)fabs((double)syz) / (mp_t)fabs((double)(syz + (mp_t)epsl)));
                            rot_y = (sp_t)((mp_t)rot_y * (mp_t)fabs((double)sxz) / (mp_t)fabs((double)(sxz + (mp_t)epsl)));
                            rot_z = (sp_t)((mp_t)rot_z * (mp_t)fabs((double)sxy) / (mp_t)fabs((double)(sxy + (mp_t)epsl)));
                        } else {
                            rot_x = (sp_t)((mp_t)rot_x * (mp_t)fabs((double)syz) / ((mp_t)fabs((double)syz) + (mp_t)epsl));
                            rot_y = (sp_t)((mp_t)rot_y * (mp_t)fabs((double)sxz) / ((mp_t)fabs((double)sxz) + (mp_t)epsl));
                            rot_z = (sp_t)((mp_t)rot_z * (mp_t)fabs((double)sxy) / ((mp_t)fabs((double)sxy) + (mp_t)epsl));
                        }
                        fr[p + 0 * npt] = div * uc * m0 * (sp_t)1e-3f;
                        fr[p + 1 * npt] = rot_x * uc * m0 * (sp_t)1e-3f;
                        fr[p + 2 * npt] = rot_y * uc * m0 * (sp_t)1e-3f;
                        fr[p + 3 * npt] = rot_z * uc * m0 * (sp_t)1e-3f;
                    } else if (q == 1) {
                        /* m_snap.f90:1678-1683 */
                        const sp_t b1 = (sp_t)(V3(s->Vx, k, i, j) * (mp_t)uc * (mp_t)m0);
                        const sp_t b2 = (sp_t)(V3(s->Vy, k, i, j) * (mp_t)uc * (mp_t)m0);
                        const sp_t b3 = (sp_t)(V3(s->Vz, k, i, j) * (mp_t)uc * (mp_t)m0);
                        if (plane >= 3) {
                            /* m_snap.f90:1697-1707 */
                            sp_t *mx = vmax[plane];
                            const sp_t h = (sp_t)sqrt((double)(b1 * b1 + b2 * b2));
                            const sp_t az = (sp_t)fabs((double)b3);
                            mx[p] = mx[p] > az ? mx[p] : az;
                            mx[p + npt] = mx[p + npt] > h ? mx[p + npt] : h;
                            mx[p + 2 * npt] = (sp_t)sqrt((double)(mx[p] * mx[p] + mx[p + npt] * mx[p + npt]));
                        }
                        if (out) {
                            fr[p] = b1;
                            fr[p + npt] = b2;
                            fr[p + 2 * npt] = b3;
                        }
                    } else {
                        /* m_snap.f90:1857-1859 */
                        sp_t *u = uacc[plane];
                        u[p] = (sp_t)((mp_t)u[p] + V3(s->Vx, k, i, j) * (mp_t)uc * (mp_t)m0 * (mp_t)dt);
                        u[p + npt] = (sp_t)((mp_t)u[p + npt] + V3(s->Vy, k, i, j) * (mp_t)uc * (mp_t)m0 * (mp_t)dt);
                        u[p + 2 * npt] = (sp_t)((mp_t)u[p + 2 * npt] + V3(s->Vz, k, i, j) * (mp_t)uc * (mp_t)m0 * (mp_t)dt);
                        if (plane >= 3) {
                            sp_t *mx = umax[plane];
                            const sp_t h = (sp_t)sqrt((double)(u[p] * u[p] + u[p + npt] * u[p + npt]));
                            const sp_t az = (sp_t)fabs((double)u[p + 2 * npt]);
                            mx[p] = mx[p] > az ? mx[p] : az;
                            mx[p + npt] = mx[p + npt] > h ? mx[p + npt] : h;
                            mx[p + 2 * npt] = (sp_t)sqrt((double)(mx[p] * mx[p] + mx[p + npt] * mx[p + npt]));
                        }
                        if (out) {
                            fr[p] = u[p];
                            fr[p + npt] = u[p + npt];
                            fr[p + 2 * npt] = u[p + 2 * npt];
                        }
                    }
                }
        }
    }
}

/* ---- reciprocity Green's-function mode (src/swpc_3d/m_green.f90) ---- */

/* UC_DERIV = 10.0**(-15) and UC_BF = 10.0**(-12), single-precision constants
   as the Fortran compiler folds them (m_green.f90:30-31) */
#define ORC_UC_DERIV 0x1.203afcp-50f
#define ORC_UC_BF 0x1.197998p-40f

/*
 * green__store, src/swpc_3d/m_green.f90:406-545.  acc holds the twelve
 * accumulators component-major, (12, ng): dxUx dxUy dxUz dyUx dyUy dyUz
 * dzUx dzUy dzUz Ux Uy Uz (all zero before it == 1, m_green.f90:424-441).
 * Each shear derivative is the mean of the four staggered 4th-order
 * derivatives around the cell (m_green.f90:458-490); r40 = C40/d and
 * r41 = C41/d with C40 = 9/8, C41 = 1/24 (m_green.f90:43-45,110-115).
 * gf is (ntw, ncmp*ng) in Fortran order with ncmp = 9 when bforce is set,
 * else 6; column (n*ncmp + c) holds mxx myy mzz myz mxz mxy [fx fy fz]
 * (m_green.f90:511-541).
 */
void orc_green_store(orc_state *s, int it, int ng, const int32_t *ig, const int32_t *jg, const int32_t *kg,
                     int bforce, int ntdec_w, int ntw, sp_t *acc, sp_t *gf)
{
    const mp_t C40 = (mp_t)9.0 / (mp_t)8.0, C41 = (mp_t)1.0 / (mp_t)24.0;
    const mp_t r40x = C40 / (mp_t)s->dx, r40y = C40 / (mp_t)s->dy, r40z = C40 / (mp_t)s->dz;
    const mp_t r41x = C41 / (mp_t)s->dx, r41y = C41 / (mp_t)s->dy, r41z = C41 / (mp_t)s->dz;
    const sp_t dt = (sp_t)s->dt;
    const int ncmp = bforce ? 9 : 6;
    for (int n = 0; n < ng; n++) {
        const int ii = ig[n], jj = jg[n], kk = kg[n];
#define VX(k_, i_, j_) V3(s->Vx, k_, i_, j_)
#define VY(k_, i_, j_) V3(s->Vy, k_, i_, j_)
#define VZ(k_, i_, j_) V3(s->Vz, k_, i_, j_)
        const mp_t dxVx = (VX(kk, ii, jj) - VX(kk, ii - 1, jj)) * r40x - (VX(kk, ii + 1, jj) - VX(kk, ii - 2, jj)) * r41x;
        const mp_t dyVy = (VY(kk, ii, jj) - VY(kk, ii, jj - 1)) * r40y - (VY(kk, ii, jj + 1) - VY(kk, ii, jj - 2)) * r41y;
        const mp_t dzVz = (VZ(kk, ii, jj) - VZ(kk - 1, ii, jj)) * r40z - (VZ(kk + 1, ii, jj) - VZ(kk - 2, ii, jj)) * r41z;

        const mp_t dxVy1 = (VY(kk, ii + 1, jj) - VY(kk, ii, jj)) * r40x - (VY(kk, ii + 2, jj) - VY(kk, ii - 1, jj)) * r41x;
        const mp_t dxVy2 = (VY(kk, ii + 1, jj - 1) - VY(kk, ii, jj - 1)) * r40x - (VY(kk, ii + 2, jj - 1) - VY(kk, ii - 1, jj - 1)) * r41x;
        const mp_t dxVy3 = (VY(kk, ii, jj) - VY(kk, ii - 1, jj)) * r40x - (VY(kk, ii + 1, jj) - VY(kk, ii - 2, jj)) * r41x;
        const mp_t dxVy4 = (VY(kk, ii, jj - 1) - VY(kk, ii - 1, jj - 1)) * r40x - (VY(kk, ii + 1, jj - 1) - VY(kk, ii - 2, jj - 1)) * r41x;

        const mp_t dxVz1 = (VZ(kk, ii + 1, jj) - VZ(kk, ii, jj)) * r40x - (VZ(kk, ii + 2, jj) - VZ(kk, ii - 1, jj)) * r41x;
        const mp_t dxVz2 = (VZ(kk - 1, ii + 1, jj) - VZ(kk - 1, ii, jj)) * r40x - (VZ(kk - 1, ii + 2, jj) - VZ(kk - 1, ii - 1, jj)) * r41x;
        const mp_t dxVz3 = (VZ(kk, ii, jj) - VZ(kk, ii - 1, jj)) * r40x - (VZ(kk, ii + 1, jj) - VZ(kk, ii - 2, jj)) * r41x;
        const mp_t dxVz4 = (VZ(kk - 1, ii, jj) - VZ(kk - 1, ii - 1, jj)) * r40x - (VZ(kk - 1, ii + 1, jj) - VZ(kk - 1, ii - 2, jj)) * r41x;

        const mp_t dyVx1 = (VX(kk, ii, jj + 1) - VX(kk, ii, jj)) * r40y - (VX(kk, ii, jj + 2) - VX(kk, ii, jj - 1)) * r41y;
        const mp_t dyVx2 = (VX(kk, ii - 1, jj + 1) - VX(kk, ii - 1, jj)) * r40y - (VX(kk, ii - 1, jj + 2) - VX(kk, ii - 1, jj - 1)) * r41y;
        const mp_t dyVx3 = (VX(kk, ii, jj) - VX(kk, ii, jj - 1)) * r40y - (VX(kk, ii, jj + 1) - VX(kk, ii, jj - 2)) * r41y;
        const mp_t dyVx4 = (VX(kk, ii - 1, jj) - VX(kk, ii - 1, jj - 1)) * r40y - (VX(kk, ii - 1, jj + 1) - VX(kk, ii - 1, jj - 2)) * r41y;

        const mp_t dyVz1 = (VZ(kk, ii, jj + 1) - VZ(kk, ii, jj)) * r40y - (VZ(kk, ii, jj + 2) - VZ(kk, ii, jj - 1)) * r41y;
        const mp_t dyVz2 = (VZ(kk - 1, ii, jj + 1) - VZ(kk - 1, ii, jj)) * r40y - (VZ(kk - 1, ii, jj + 2) - VZ(kk - 1, ii, jj - 1)) * r41y;
        const mp_t dyVz3 = (VZ(kk, ii, jj) - VZ(kk, ii, jj - 1)) * r40y - (VZ(kk, ii, jj + 1) - VZ(kk, ii, jj - 2)) * r41y;
        const mp_t dyVz4 = (VZ(kk - 1, ii, jj) - VZ(kk - 1, ii, jj - 1)) * r40y - (VZ(kk - 1, ii, jj + 1) - VZ(kk - 1, ii, jj - 2)) * r41y;

        const mp_t dzVx1 = (VX(kk + 1, ii, jj) - VX(kk, ii, jj)) * r40z - (VX(kk + 2, ii, jj) - VX(kk - 1, ii, jj)) * r41z;
        const mp_t dzVx2 = (VX(kk + 1, ii - 1, jj) - VX(kk, ii - 1, jj)) * r40z - (VX(kk + 2, ii - 1, jj) - VX(kk - 1, ii - 1, jj)) * r41z;
        const mp_t dzVx3 = (VX(kk, ii, jj) - VX(kk - 1, ii, jj)) * r40z - (VX(kk + 1, ii, jj) - VX(kk - 2, ii, jj)) * r41z;
        const mp_t dzVx4 = (VX(kk, ii - 1, jj) - VX(kk - 1, ii - 1, jj)) * r40z - (VX(kk + 1, ii - 1, jj) - VX(kk - 2, ii - 1, jj)) * r41z;

        const mp_t dzVy1 = (VY(kk + 1, ii, jj) - VY(kk, ii, jj)) * r40z - (VY(kk + 2, ii, jj) - VY(kk - 1, ii, jj)) * r41z;
        const mp_t dzVy2 = (VY(kk + 1, ii, jj - 1) - VY(kk, ii, jj - 1)) * r40z - (VY(kk + 2, ii, jj - 1) - VY(kk - 1, ii, jj - 1)) * r41z;
        const mp_t dzVy3 = (VY(kk, ii, jj) - VY(kk - 1, ii, jj)) * r40z - (VY(kk + 1, ii, jj) - VY(kk - 2, ii, jj)) * r41z;
        const mp_t dzVy4 = (VY(kk, ii, jj - 1) - VY(kk - 1, ii, jj - 1)) * r40z - (VY(kk + 1, ii, jj - 1) - VY(kk - 2, ii, jj - 1)) * r41z;

        const mp_t quarter = (mp_t)(sp_t)0.25;
        const mp_t dxVy = (dxVy1 + dxVy2 + dxVy3 + dxVy4) * quarter;
        const mp_t dxVz = (dxVz1 + dxVz2 + dxVz3 + dxVz4) * quarter;
        const mp_t dyVx = (dyVx1 + dyVx2 + dyVx3 + dyVx4) * quarter;
        const mp_t dyVz = (dyVz1 + dyVz2 + dyVz3 + dyVz4) * quarter;
        const mp_t dzVx = (dzVx1 + dzVx2 + dzVx3 + dzVx4) * quarter;
        const mp_t dzVy = (dzVy1 + dzVy2 + dzVy3 + dzVy4) * quarter;

        const mp_t d[9] = {dxVx, dxVy, dxVz, dyVx, dyVy, dyVz, dzVx, dzVy, dzVz};
        for (int q = 0; q < 9; q++)
            acc[(int64_t)q * ng + n] = acc[(int64_t)q * ng + n] + (sp_t)(dr_t)(d[q] * (mp_t)dt);

        if (bforce) {
            sp_t *u = acc + (int64_t)9 * ng;
            u[n] = u[n] + (sp_t)0.5 * (sp_t)(dr_t)(VX(kk, ii, jj) + VX(kk, ii - 1, jj)) * dt;
            u[(int64_t)ng + n] = u[(int64_t)ng + n] + (sp_t)0.5 * (sp_t)(dr_t)(VY(kk, ii, jj) + VY(kk, ii, jj - 1)) * dt;
            u[(int64_t)2 * ng + n] = u[(int64_t)2 * ng + n] + (sp_t)0.5 * (sp_t)(dr_t)(VZ(kk, ii, jj) + VZ(kk - 1, ii, jj)) * dt;
        }
#undef VX
#undef VY
#undef VZ
    }

    if ((it - 1) % ntdec_w != 0) return;
    const int itw = (it - 1) / ntdec_w;
    if (itw < 0 || itw >= ntw) return;
    const sp_t ucd = (sp_t)ORC_UC_DERIV, ucb = (sp_t)ORC_UC_BF, e9 = (sp_t)1e9f;
    for (int n = 0; n < ng; n++) {
#define A(q) acc[(int64_t)(q) * ng + n]
#define GF(c) gf[((int64_t)n * ncmp + (c)) * ntw + itw]
        GF(0) = A(0) * ucd * e9;
        GF(1) = A(4) * ucd * e9;
        GF(2) = A(8) * ucd * e9;
        GF(3) = (A(5) + A(7)) * ucd * e9;
        GF(4) = (A(2) + A(6)) * ucd * e9;
        GF(5) = (A(1) + A(3)) * ucd * e9;
        if (bforce) {
            GF(6) = A(9) * ucb * e9;
            GF(7) = A(10) * ucb * e9;
            GF(8) = A(11) * ucb * e9;
        }
#undef A
#undef GF
    }
}

/*
 * green__source, src/swpc_3d/m_green.f90:614-649: the unit force
 * (fx1, fy1, fz1) times dt_dxyz and the moment rate stf (taken at
 * t = tbeg + it*dt, m_green.f90:632) split onto the two staggered velocity
 * nodes around the station cell, each with the averaged buoyancy of its two
 * cells.  dt_dxyz = real(dt/(dx*dy*dz)) is single precision
 * (m_green.f90:33,149).  All six updates are made, zero ones included.
 */
void orc_green_source(orc_state *s, int isrc, int jsrc, int ksrc, float fx1, float fy1, float fz1, float stf)
{
    const sp_t dt_dxyz = (sp_t)(dr_t)((mp_t)s->dt / ((mp_t)s->dx * (mp_t)s->dy * (mp_t)s->dz));
    const sp_t st = (sp_t)stf;
    const sp_t fx = (sp_t)fx1 * dt_dxyz * st;
    const sp_t fy = (sp_t)fy1 * dt_dxyz * st;
    const sp_t fz = (sp_t)fz1 * dt_dxyz * st;
    const int ii = isrc, jj = jsrc, kk = ksrc;
    const sp_t r0 = V3(s->rho, kk, ii, jj);
    V3(s->Vx, kk, ii, jj) = V3(s->Vx, kk, ii, jj) + ((sp_t)2.0 / (r0 + V3(s->rho, kk, ii + 1, jj))) * fx / 2;
    V3(s->Vx, kk, ii - 1, jj) = V3(s->Vx, kk, ii - 1, jj) + ((sp_t)2.0 / (r0 + V3(s->rho, kk, ii - 1, jj))) * fx / 2;
    V3(s->Vy, kk, ii, jj) = V3(s->Vy, kk, ii, jj) + ((sp_t)2.0 / (r0 + V3(s->rho, kk, ii, jj + 1))) * fy / 2;
    V3(s->Vy, kk, ii, jj - 1) = V3(s->Vy, kk, ii, jj - 1) + ((sp_t)2.0 / (r0 + V3(s->rho, kk, ii, jj - 1))) * fy / 2;
    V3(s->Vz, kk, ii, jj) = V3(s->Vz, kk, ii, jj) + ((sp_t)2.0 / (r0 + V3(s->rho, kk + 1, ii, jj))) * fz / 2;
    V3(s->Vz, kk - 1, ii, jj) = V3(s->Vz, kk - 1, ii, jj) + ((sp_t)2.0 / (r0 + V3(s->rho, kk - 1, ii, jj))) * fz / 2;
}
