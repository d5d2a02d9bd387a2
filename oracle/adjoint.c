/*
 * adjoint.c -- oracle restatement of the discrete adjoint of one time step
 * (generalized Zener memory variables, ADE-CFS PML, free-surface band).
 * TEST INFRASTRUCTURE ONLY (see swpc_oracle.h).
 *
 * The forward step (src/swpc_3d/main.f90:124-133, m_kernel.f90, m_absorb_p.f90)
 * is linear in the state for a fixed medium:
 *     S' = S + Ks(lam, mu)[V, aV],  aV' = g3 aV + g4 dt D V      (stress)
 *     V' = V + Kv(rho)[S', aS],     aS' = g3 aS + g4 dt D S'     (velocity)
 * With NM mechanisms the interior stress update also advances the memory
 * variables first (m_kernel.f90:222-238,319-333):
 *     R'_c[m] = c1[m] R_c[m] - c2[m] E_c dt,   S'_c = S_c + (F_c + sum_m d1[m] R'_c[m]) dt
 * with E_c, F_c linear in the derivatives (taup/taus and 1 + tau(1+d2)
 * weights).  Its transpose: rho_c[m] = aR'_c[m] + d1[m] dt aS'_c, the new
 * aR_c[m] = c1[m] rho_c[m], and E_c receives -dt sum_m c2[m] rho_c[m].
 * Its adjoint applies the transposes in reverse order.  Each forward
 * update reads spatial derivatives of the other field; the adjoint first
 * forms, per cell, the adjoint of every derivative the cell used (a
 * "density" D_*), then every field point gathers the densities of the cells
 * whose stencils touch it, with the coefficients of those cells: the 4th-order
 * band-switched pair (re40, re41) for interior cells, 1/h for PML cells.
 *
 * Gradients with respect to lam, mu, rho accumulate in double from the
 * forward state of the same step (V_n, aV_n for the stress update; S_{n+1},
 * aS_n for the velocity update), including the harmonic mean of mu at the
 * shear points (m_kernel.f90:293-309) and the two-cell buoyancy
 * 2/(rho + rho').
 *
 * Parity status: pinned by the dot-product test (the adjoint is the
 * transpose of the forward step to rounding, elastic and with three memory
 * variables, with controls that dropped auxiliary or memory-variable
 * adjoints are detected) and by central finite differences of the misfit,
 * elastic and NM = 3 (tests/test_oracle_adjoint.py).
 */
#include <math.h>
#include <string.h>
#include "swpc_oracle.h"

#define V3(a, k, i, j) (a)[orc_i3(s, (k), (i), (j))]
#define M2(a, i, j) (a)[orc_i2(s, (i), (j))]

typedef struct orc_adj {
    mp_t *Vx, *Vy, *Vz, *Sxx, *Syy, *Szz, *Syz, *Sxz, *Sxy; /* adjoint fields, (k,i,j) halo */
    sp_t *aV[9];  /* adjoints of axVx ayVx azVx axVy ayVy azVy axVz ayVz azVz */
    sp_t *aS[9];  /* adjoints of axSxx aySxy azSxz axSxy aySyy azSyz axSxz aySyz azSzz */
    mp_t *D[9];   /* per-cell derivative adjoints (scratch), (k,i,j) halo */
    double *g_lam, *g_mu, *g_rho; /* (k,i,j) halo */
    sp_t *aR[6];  /* adjoints of Rxx Ryy Rzz Ryz Rxz Rxy, (m,k,i,j) on the kernel box */
} orc_adj;

/* memory-variable index (m,k,i,j) on the kernel box, m 1-based (m_kernel.f90:392-397) */
static inline int64_t ridx(const orc_state *s, int m, int k, int i, int j)
{
    int64_t nzk = s->kend_k - s->kbeg_k + 1, nxk = s->iend_k - s->ibeg_k + 1;
    return (int64_t)(m - 1) +
           (int64_t)s->nm * ((k - s->kbeg_k) + nzk * ((i - s->ibeg_k) + nxk * (int64_t)(j - s->jbeg_k)));
}

static inline int64_t aidx(const orc_state *s, int k, int i, int j)
{
    int64_t nza = s->nz - s->kbeg_min + 1;
    return (k - s->kbeg_min) + nza * ((int64_t)(i - 1) + (int64_t)s->nx * (j - 1));
}
#define G(a, c, n) (double)(a)[4 * ((n) - 1) + ((c) - 1)]

static inline int is_int(const orc_state *s, int k, int i, int j)
{
    return i >= s->ibeg_k && i <= s->iend_k && j >= s->jbeg_k && j <= s->jend_k && k >= s->kbeg_k && k <= s->kend_k;
}
static inline int is_pml(const orc_state *s, int k, int i, int j)
{
    if (s->abc != 0 || i < 1 || i > s->nx || j < 1 || j > s->ny || k > s->nz) return 0;
    return !is_int(s, k, i, j) && k >= M2(s->kbeg_a, i, j);
}

/* 4th-order pair (re40, re41) along an axis at a cell (m_kernel.f90:43-54,103-111) */
static void coef_at(const orc_state *s, int axis, int k, int i, int j, double *a, double *b)
{
    const double h = axis == 0 ? (double)(mp_t)s->dx : axis == 1 ? (double)(mp_t)s->dy : (double)(mp_t)s->dz;
    const int ka = (k - M2(s->kfs_top, i, j)) * (M2(s->kfs_bot, i, j) - k);
    const int kb = (k - M2(s->kob_top, i, j)) * (M2(s->kob_bot, i, j) - k);
    const double isign = (ka > kb ? ka : kb) >= 0 ? 1.0 : -1.0;
    *a = 17.0 / 16.0 / h + isign * (-1.0 / 16.0 / h);
    *b = 1.0 / 48.0 / h + isign * (-1.0 / 48.0 / h);
}

static inline double hinv(const orc_state *s, int axis)
{
    return 1.0 / (axis == 0 ? (double)(mp_t)s->dx : axis == 1 ? (double)(mp_t)s->dy : (double)(mp_t)s->dz);
}

/* value of D at the cell p + o along axis, 0 outside the updated cells */
static inline double dens(const orc_state *s, const mp_t *D, int axis, int o, int k, int i, int j, int *kind)
{
    if (axis == 0) i += o; else if (axis == 1) j += o; else k += o;
    if (is_int(s, k, i, j)) *kind = 1;
    else if (is_pml(s, k, i, j)) *kind = 2;
    else { *kind = 0; return 0.0; }
    return (double)V3(D, k, i, j);
}

/* gather of a backward-type derivative (cell c reads f(c), f(c-1) and, at
   4th order, f(c+1), f(c-2)) onto the point p = (k,i,j) */
static double gather_bwd(const orc_state *s, const mp_t *D, int axis, int k, int i, int j)
{
    static const int off[4] = {0, 1, -1, 2};
    double acc = 0.0;
    for (int t = 0; t < 4; t++) {
        int kind;
        const int o = off[t];
        const double d = dens(s, D, axis, o, k, i, j, &kind);
        if (!kind) continue;
        int kk = k, ii = i, jj = j;
        if (axis == 0) ii += o; else if (axis == 1) jj += o; else kk += o;
        if (kind == 1) {
            double a, b;
            coef_at(s, axis, kk, ii, jj, &a, &b);
            const double w = o == 0 ? a : o == 1 ? -a : o == -1 ? -b : b;
            acc += w * d;
        } else {
            const double h = hinv(s, axis);
            if (o == 0) acc += h * d;
            else if (o == 1) acc -= h * d;
        }
    }
    return acc;
}

/* gather of a forward-type derivative (cell c reads f(c+1), f(c) and, at
   4th order, f(c+2), f(c-1)) onto the point p */
static double gather_fwd(const orc_state *s, const mp_t *D, int axis, int k, int i, int j)
{
    static const int off[4] = {-1, 0, -2, 1};
    double acc = 0.0;
    for (int t = 0; t < 4; t++) {
        int kind;
        const int o = off[t];
        const double d = dens(s, D, axis, o, k, i, j, &kind);
        if (!kind) continue;
        int kk = k, ii = i, jj = j;
        if (axis == 0) ii += o; else if (axis == 1) jj += o; else kk += o;
        if (kind == 1) {
            double a, b;
            coef_at(s, axis, kk, ii, jj, &a, &b);
            const double w = o == -1 ? a : o == 0 ? -a : o == -2 ? -b : b;
            acc += w * d;
        } else {
            const double h = hinv(s, axis);
            if (o == -1) acc += h * d;
            else if (o == 0) acc -= h * d;
        }
    }
    return acc;
}

/* 4th-order derivatives of the forward kernel at an interior cell */
#define BWD4(F, dk, di, dj, a, b) \
    (((double)V3(F, k, i, j) - (double)V3(F, k - (dk), i - (di), j - (dj))) * (a) \
     - ((double)V3(F, k + (dk), i + (di), j + (dj)) - (double)V3(F, k - 2 * (dk), i - 2 * (di), j - 2 * (dj))) * (b))
#define FWD4(F, dk, di, dj, a, b) \
    (((double)V3(F, k + (dk), i + (di), j + (dj)) - (double)V3(F, k, i, j)) * (a) \
     - ((double)V3(F, k + 2 * (dk), i + 2 * (di), j + 2 * (dj)) - (double)V3(F, k - (dk), i - (di), j - (dj))) * (b))
#define BWD2(F, dk, di, dj, h) (((double)V3(F, k, i, j) - (double)V3(F, k - (dk), i - (di), j - (dj))) * (h))
#define FWD2(F, dk, di, dj, h) (((double)V3(F, k + (dk), i + (di), j + (dj)) - (double)V3(F, k, i, j)) * (h))

static void hmean4_grad(double a, double b, double c, double d, double eps, double *h, double g[4])
{
    const double N = 4 * a * b * c * d, Dn = a * b * c + a * b * d + a * c * d + b * c * d + eps;
    *h = N / Dn;
    g[0] = (4 * b * c * d * Dn - N * (b * c + b * d + c * d)) / (Dn * Dn);
    g[1] = (4 * a * c * d * Dn - N * (a * c + a * d + c * d)) / (Dn * Dn);
    g[2] = (4 * a * b * d * Dn - N * (a * b + a * d + b * d)) / (Dn * Dn);
    g[3] = (4 * a * b * c * Dn - N * (a * b + a * c + b * c)) / (Dn * Dn);
}

/*
 * Adjoint of the velocity update of one step.  s holds the forward S_{n+1}
 * and the S-auxiliaries aS_n (before the update); a->V* hold the adjoint of
 * V_{n+1} (unchanged: V_n's adjoint receives it through the identity), a->aS
 * the adjoint of aS_{n+1}, replaced by that of aS_n; a->S* receive the
 * contributions of this update.  g_rho accumulates.
 */
void orc_adj_vel(const orc_state *s, orc_adj *a)
{
    const double dt = (double)s->dt;
    const double hx = hinv(s, 0), hy = hinv(s, 1), hz = hinv(s, 2);
    for (int q = 0; q < 9; q++) memset(a->D[q], 0, sizeof(mp_t) * orc_n3(s));
    /* per-cell densities: D[0..8] = d/d(dxSxx dySxy dzSxz dxSxy dySyy dzSyz dxSxz dySyz dzSzz).
       A cell also adds to the rho gradient of row j + 1: two passes over the
       rows of one parity keep the rows each thread writes disjoint. */
    for (int par = 0; par < 2; par++)
#pragma omp parallel for schedule(static)
    for (int j = 1 + par; j <= s->ny; j += 2)
        for (int i = 1; i <= s->nx; i++)
            for (int k = 1; k <= s->nz; k++) {
                const int in = is_int(s, k, i, j), pm = is_pml(s, k, i, j);
                if (!in && !pm) continue;
                const double r0 = V3(s->rho, k, i, j);
                const double bx = 2.0 / (r0 + V3(s->rho, k, i + 1, j));
                const double by = 2.0 / (r0 + V3(s->rho, k, i, j + 1));
                const double bz = 2.0 / (r0 + V3(s->rho, k + 1, i, j));
                const double wx = bx * dt * (double)V3(a->Vx, k, i, j);
                const double wy = by * dt * (double)V3(a->Vy, k, i, j);
                const double wz = bz * dt * (double)V3(a->Vz, k, i, j);
                double fx, fy, fz; /* forward force terms, for the rho gradient */
                if (in) {
                    double ax, bxx, ay, byy, az, bzz;
                    coef_at(s, 0, k, i, j, &ax, &bxx);
                    coef_at(s, 1, k, i, j, &ay, &byy);
                    coef_at(s, 2, k, i, j, &az, &bzz);
                    V3(a->D[0], k, i, j) = wx; V3(a->D[1], k, i, j) = wx; V3(a->D[2], k, i, j) = wx;
                    V3(a->D[3], k, i, j) = wy; V3(a->D[4], k, i, j) = wy; V3(a->D[5], k, i, j) = wy;
                    V3(a->D[6], k, i, j) = wz; V3(a->D[7], k, i, j) = wz; V3(a->D[8], k, i, j) = wz;
                    fx = FWD4(s->Sxx, 0, 1, 0, ax, bxx) + BWD4(s->Sxy, 0, 0, 1, ay, byy) + BWD4(s->Sxz, 1, 0, 0, az, bzz);
                    fy = BWD4(s->Sxy, 0, 1, 0, ax, bxx) + FWD4(s->Syy, 0, 0, 1, ay, byy) + BWD4(s->Syz, 1, 0, 0, az, bzz);
                    fz = BWD4(s->Sxz, 0, 1, 0, ax, bxx) + BWD4(s->Syz, 0, 0, 1, ay, byy) + FWD4(s->Szz, 1, 0, 0, az, bzz);
                } else {
                    const int64_t p = aidx(s, k, i, j);
                    /* profile of each derivative: (array, index) in forward order */
                    const sp_t *gp[9] = {s->gxe, s->gyc, s->gzc, s->gxc, s->gye, s->gzc, s->gxc, s->gyc, s->gze};
                    const int gn[9] = {i, j, k, i, j, k, i, j, k};
                    const double w[9] = {wx, wx, wx, wy, wy, wy, wz, wz, wz};
                    const sp_t *aux[9] = {s->axSxx, s->aySxy, s->azSxz, s->axSxy, s->aySyy, s->azSyz, s->axSxz, s->aySyz, s->azSzz};
                    const double der[9] = {FWD2(s->Sxx, 0, 1, 0, hx), BWD2(s->Sxy, 0, 0, 1, hy), BWD2(s->Sxz, 1, 0, 0, hz),
                                           BWD2(s->Sxy, 0, 1, 0, hx), FWD2(s->Syy, 0, 0, 1, hy), BWD2(s->Syz, 1, 0, 0, hz),
                                           BWD2(s->Sxz, 0, 1, 0, hx), BWD2(s->Syz, 0, 0, 1, hy), FWD2(s->Szz, 1, 0, 0, hz)};
                    double f[3] = {0, 0, 0};
                    for (int q = 0; q < 9; q++) {
                        const double g1 = G(gp[q], 1, gn[q]), g2 = G(gp[q], 2, gn[q]);
                        const double g3 = G(gp[q], 3, gn[q]), g4 = G(gp[q], 4, gn[q]);
                        const double ahat = (double)a->aS[q][p];
                        V3(a->D[q], k, i, j) = g1 * w[q] + g4 * dt * ahat;
                        a->aS[q][p] = (sp_t)(g2 * w[q] + g3 * ahat);
                        f[q / 3] += g1 * der[q] + g2 * (double)aux[q][p];
                    }
                    fx = f[0]; fy = f[1]; fz = f[2];
                }
                /* rho gradient: both densities of each staggered buoyancy */
                const double qx = -2.0 / ((r0 + V3(s->rho, k, i + 1, j)) * (r0 + V3(s->rho, k, i + 1, j))) * dt * fx * (double)V3(a->Vx, k, i, j);
                const double qy = -2.0 / ((r0 + V3(s->rho, k, i, j + 1)) * (r0 + V3(s->rho, k, i, j + 1))) * dt * fy * (double)V3(a->Vy, k, i, j);
                const double qz = -2.0 / ((r0 + V3(s->rho, k + 1, i, j)) * (r0 + V3(s->rho, k + 1, i, j))) * dt * fz * (double)V3(a->Vz, k, i, j);
                V3(a->g_rho, k, i, j) += qx + qy + qz;
                V3(a->g_rho, k, i + 1, j) += qx;
                V3(a->g_rho, k, i, j + 1) += qy;
                V3(a->g_rho, k + 1, i, j) += qz;
            }
    /* gather onto the stress points */
#pragma omp parallel for schedule(static)
    for (int j = -1; j <= s->ny + 2; j++)
        for (int i = -1; i <= s->nx + 2; i++)
            for (int k = -1; k <= s->nz + 2; k++) {
                V3(a->Sxx, k, i, j) += (mp_t)gather_fwd(s, a->D[0], 0, k, i, j);
                V3(a->Syy, k, i, j) += (mp_t)gather_fwd(s, a->D[4], 1, k, i, j);
                V3(a->Szz, k, i, j) += (mp_t)gather_fwd(s, a->D[8], 2, k, i, j);
                V3(a->Sxy, k, i, j) += (mp_t)(gather_bwd(s, a->D[1], 1, k, i, j) + gather_bwd(s, a->D[3], 0, k, i, j));
                V3(a->Sxz, k, i, j) += (mp_t)(gather_bwd(s, a->D[2], 2, k, i, j) + gather_bwd(s, a->D[6], 0, k, i, j));
                V3(a->Syz, k, i, j) += (mp_t)(gather_bwd(s, a->D[5], 2, k, i, j) + gather_bwd(s, a->D[7], 1, k, i, j));
            }
}

/*
 * Adjoint of the stress update of one step.  s holds the forward V_n and
 * the V-auxiliaries aV_n; a->S* hold the adjoint of S_{n+1} (unchanged: the
 * identity carries it to S_n), a->aV that of aV_{n+1}, replaced by that of
 * aV_n; a->V* receive this update's contributions.  g_lam and g_mu
 * accumulate.
 */
void orc_adj_stress(const orc_state *s, orc_adj *a)
{
    const double dt = (double)s->dt;
    const double hx = hinv(s, 0), hy = hinv(s, 1), hz = hinv(s, 2);
    const double eps = 1.1920928955078125e-07;
    for (int q = 0; q < 9; q++) memset(a->D[q], 0, sizeof(mp_t) * orc_n3(s));
    /* D[0..8] = d/d(dxVx dyVy dzVz dyVz dzVy dxVz dzVx dxVy dyVx).  A cell
       also adds to the mu gradient of row j + 1: rows of one parity per pass. */
    for (int par = 0; par < 2; par++)
#pragma omp parallel for schedule(static)
    for (int j = 1 + par; j <= s->ny; j += 2)
        for (int i = 1; i <= s->nx; i++)
            for (int k = 1; k <= s->nz; k++) {
                const int in = is_int(s, k, i, j), pm = is_pml(s, k, i, j);
                if (!in && !pm) continue;
                const double lam = V3(s->lam, k, i, j), mu = V3(s->mu, k, i, j);
                const double sxx = V3(a->Sxx, k, i, j), syy = V3(a->Syy, k, i, j), szz = V3(a->Szz, k, i, j);
                const double X = dt * ((lam + 2 * mu) * sxx + lam * (syy + szz));
                const double Y = dt * ((lam + 2 * mu) * syy + lam * (sxx + szz));
                const double Z = dt * ((lam + 2 * mu) * szz + lam * (sxx + syy));
                double hyz, hxz, hxy, gyz[4], gxz[4], gxy[4];
                hmean4_grad(V3(s->mu, k, i, j), V3(s->mu, k + 1, i, j), V3(s->mu, k, i, j + 1), V3(s->mu, k + 1, i, j + 1), eps, &hyz, gyz);
                hmean4_grad(V3(s->mu, k, i, j), V3(s->mu, k + 1, i, j), V3(s->mu, k, i + 1, j), V3(s->mu, k + 1, i + 1, j), eps, &hxz, gxz);
                hmean4_grad(V3(s->mu, k, i, j), V3(s->mu, k, i + 1, j), V3(s->mu, k, i, j + 1), V3(s->mu, k, i + 1, j + 1), eps, &hxy, gxy);
                hyz = (double)(sp_t)hyz; hxz = (double)(sp_t)hxz; hxy = (double)(sp_t)hxy;
                const double Wyz = dt * hyz * (double)V3(a->Syz, k, i, j);
                const double Wxz = dt * hxz * (double)V3(a->Sxz, k, i, j);
                const double Wxy = dt * hxy * (double)V3(a->Sxy, k, i, j);
                double ex, ey, ez, tyz, txz, txy; /* forward derivative combinations, for the gradients */
                if (in) {
                    double ax, bxx, ay, byy, az, bzz;
                    coef_at(s, 0, k, i, j, &ax, &bxx);
                    coef_at(s, 1, k, i, j, &ay, &byy);
                    coef_at(s, 2, k, i, j, &az, &bzz);
                    ex = BWD4(s->Vx, 0, 1, 0, ax, bxx);
                    ey = BWD4(s->Vy, 0, 0, 1, ay, byy);
                    ez = BWD4(s->Vz, 1, 0, 0, az, bzz);
                    tyz = FWD4(s->Vz, 0, 0, 1, ay, byy) + FWD4(s->Vy, 1, 0, 0, az, bzz);
                    txz = FWD4(s->Vz, 0, 1, 0, ax, bxx) + FWD4(s->Vx, 1, 0, 0, az, bzz);
                    txy = FWD4(s->Vy, 0, 1, 0, ax, bxx) + FWD4(s->Vx, 0, 0, 1, ay, byy);
                    if (s->nm > 0) {
                        /* transpose of the memory variables and of the weighted
                           stress sums (m_kernel.f90:222-238,319-333) */
                        const double taup = V3(s->taup, k, i, j), taus = V3(s->taus, k, i, j);
                        const double d2 = (double)s->d2;
                        const double tp1 = 1.0 + taup * (1.0 + d2), ts1 = 1.0 + taus * (1.0 + d2);
                        const double as[6] = {sxx, syy, szz, (double)V3(a->Syz, k, i, j),
                                              (double)V3(a->Sxz, k, i, j), (double)V3(a->Sxy, k, i, j)};
                        double sig[6] = {0, 0, 0, 0, 0, 0};
                        for (int m = 1; m <= s->nm; m++) {
                            const double c1 = s->c1[m - 1], c2 = s->c2[m - 1], d1 = s->d1[m - 1];
                            const int64_t r = ridx(s, m, k, i, j);
                            for (int c = 0; c < 6; c++) {
                                const double rh = (double)a->aR[c][r] + d1 * dt * as[c];
                                a->aR[c][r] = (sp_t)(c1 * rh);
                                sig[c] -= dt * c2 * rh;
                            }
                        }
                        double P[3], Q[6];
                        for (int c = 0; c < 6; c++) {
                            if (c < 3) P[c] = dt * as[c] * tp1 + sig[c] * taup;
                            Q[c] = dt * as[c] * ts1 + sig[c] * taus;
                        }
                        const double sP = P[0] + P[1] + P[2], l2m = lam + 2 * mu, d3 = ex + ey + ez;
                        V3(a->D[0], k, i, j) = l2m * sP - 2 * mu * (Q[1] + Q[2]);
                        V3(a->D[1], k, i, j) = l2m * sP - 2 * mu * (Q[0] + Q[2]);
                        V3(a->D[2], k, i, j) = l2m * sP - 2 * mu * (Q[0] + Q[1]);
                        V3(a->D[3], k, i, j) = hyz * Q[3]; V3(a->D[4], k, i, j) = hyz * Q[3];
                        V3(a->D[5], k, i, j) = hxz * Q[4]; V3(a->D[6], k, i, j) = hxz * Q[4];
                        V3(a->D[7], k, i, j) = hxy * Q[5]; V3(a->D[8], k, i, j) = hxy * Q[5];
                        V3(a->g_lam, k, i, j) += d3 * sP;
                        V3(a->g_mu, k, i, j) += 2 * d3 * sP - 2 * ((d3 - ex) * Q[0] + (d3 - ey) * Q[1] + (d3 - ez) * Q[2]);
                        const double qyz = tyz * Q[3], qxz = txz * Q[4], qxy = txy * Q[5];
                        V3(a->g_mu, k, i, j) += qyz * gyz[0] + qxz * gxz[0] + qxy * gxy[0];
                        V3(a->g_mu, k + 1, i, j) += qyz * gyz[1] + qxz * gxz[1];
                        V3(a->g_mu, k, i, j + 1) += qyz * gyz[2] + qxy * gxy[2];
                        V3(a->g_mu, k + 1, i, j + 1) += qyz * gyz[3];
                        V3(a->g_mu, k, i + 1, j) += qxz * gxz[2] + qxy * gxy[1];
                        V3(a->g_mu, k + 1, i + 1, j) += qxz * gxz[3];
                        V3(a->g_mu, k, i + 1, j + 1) += qxy * gxy[3];
                        continue;
                    }
                    V3(a->D[0], k, i, j) = X; V3(a->D[1], k, i, j) = Y; V3(a->D[2], k, i, j) = Z;
                    V3(a->D[3], k, i, j) = Wyz; V3(a->D[4], k, i, j) = Wyz;
                    V3(a->D[5], k, i, j) = Wxz; V3(a->D[6], k, i, j) = Wxz;
                    V3(a->D[7], k, i, j) = Wxy; V3(a->D[8], k, i, j) = Wxy;
                } else {
                    const int64_t p = aidx(s, k, i, j);
                    const sp_t *gp[9] = {s->gxc, s->gyc, s->gzc, s->gye, s->gze, s->gxe, s->gze, s->gxe, s->gye};
                    const int gn[9] = {i, j, k, j, k, i, k, i, j};
                    const double w[9] = {X, Y, Z, Wyz, Wyz, Wxz, Wxz, Wxy, Wxy};
                    /* auxiliaries in derivative order: axVx ayVy azVz ayVz azVy axVz azVx axVy ayVx */
                    const int ai[9] = {0, 4, 8, 7, 5, 6, 2, 3, 1};
                    const sp_t *aux[9] = {s->axVx, s->ayVx, s->azVx, s->axVy, s->ayVy, s->azVy, s->axVz, s->ayVz, s->azVz};
                    const double der[9] = {BWD2(s->Vx, 0, 1, 0, hx), BWD2(s->Vy, 0, 0, 1, hy), BWD2(s->Vz, 1, 0, 0, hz),
                                           FWD2(s->Vz, 0, 0, 1, hy), FWD2(s->Vy, 1, 0, 0, hz), FWD2(s->Vz, 0, 1, 0, hx),
                                           FWD2(s->Vx, 1, 0, 0, hz), FWD2(s->Vy, 0, 1, 0, hx), FWD2(s->Vx, 0, 0, 1, hy)};
                    double t[9];
                    for (int q = 0; q < 9; q++) {
                        const double g1 = G(gp[q], 1, gn[q]), g2 = G(gp[q], 2, gn[q]);
                        const double g3 = G(gp[q], 3, gn[q]), g4 = G(gp[q], 4, gn[q]);
                        const double ahat = (double)a->aV[ai[q]][p];
                        V3(a->D[q], k, i, j) = g1 * w[q] + g4 * dt * ahat;
                        a->aV[ai[q]][p] = (sp_t)(g2 * w[q] + g3 * ahat);
                        t[q] = g1 * der[q] + g2 * (double)aux[ai[q]][p];
                    }
                    ex = t[0]; ey = t[1]; ez = t[2];
                    tyz = t[3] + t[4]; txz = t[5] + t[6]; txy = t[7] + t[8];
                }
                V3(a->g_lam, k, i, j) += dt * (ex + ey + ez) * (sxx + syy + szz);
                V3(a->g_mu, k, i, j) += 2 * dt * (ex * sxx + ey * syy + ez * szz);
                const double qyz = dt * tyz * (double)V3(a->Syz, k, i, j);
                const double qxz = dt * txz * (double)V3(a->Sxz, k, i, j);
                const double qxy = dt * txy * (double)V3(a->Sxy, k, i, j);
                V3(a->g_mu, k, i, j) += qyz * gyz[0] + qxz * gxz[0] + qxy * gxy[0];
                V3(a->g_mu, k + 1, i, j) += qyz * gyz[1] + qxz * gxz[1];
                V3(a->g_mu, k, i, j + 1) += qyz * gyz[2] + qxy * gxy[2];
                V3(a->g_mu, k + 1, i, j + 1) += qyz * gyz[3];
                V3(a->g_mu, k, i + 1, j) += qxz * gxz[2] + qxy * gxy[1];
                V3(a->g_mu, k + 1, i + 1, j) += qxz * gxz[3];
                V3(a->g_mu, k, i + 1, j + 1) += qxy * gxy[3];
            }
    /* gather onto the velocity points */
#pragma omp parallel for schedule(static)
    for (int j = -1; j <= s->ny + 2; j++)
        for (int i = -1; i <= s->nx + 2; i++)
            for (int k = -1; k <= s->nz + 2; k++) {
                V3(a->Vx, k, i, j) += (mp_t)(gather_bwd(s, a->D[0], 0, k, i, j) + gather_fwd(s, a->D[8], 1, k, i, j)
                                             + gather_fwd(s, a->D[6], 2, k, i, j));
                V3(a->Vy, k, i, j) += (mp_t)(gather_bwd(s, a->D[1], 1, k, i, j) + gather_fwd(s, a->D[7], 0, k, i, j)
                                             + gather_fwd(s, a->D[4], 2, k, i, j));
                V3(a->Vz, k, i, j) += (mp_t)(gather_bwd(s, a->D[2], 2, k, i, j) + gather_fwd(s, a->D[5], 0, k, i, j)
                                             + gather_fwd(s, a->D[3], 1, k, i, j));
            }
}

/* adjoint of the velocity traces of wav__store (m_wav.f90:517-531) at one
   recorded step: res(3, nst) residuals d - d_obs in the output units */
void orc_adj_record(const orc_state *s, orc_adj *a, int nst, const int32_t *ist, const int32_t *jst,
                    const int32_t *kst, float M0, float UC, const double *res)
{
    const double sc = 0.5 * (double)M0 * (double)UC * 1e9;
    for (int n = 0; n < nst; n++) {
        const int i = ist[n], j = jst[n], k = kst[n];
        V3(a->Vx, k, i, j) += (mp_t)(sc * res[3 * n + 0]);
        V3(a->Vx, k, i - 1, j) += (mp_t)(sc * res[3 * n + 0]);
        V3(a->Vy, k, i, j) += (mp_t)(sc * res[3 * n + 1]);
        V3(a->Vy, k, i, j - 1) += (mp_t)(sc * res[3 * n + 1]);
        V3(a->Vz, k, i, j) -= (mp_t)(sc * res[3 * n + 2]);
        V3(a->Vz, k - 1, i, j) -= (mp_t)(sc * res[3 * n + 2]);
    }
}
