/*
 * swpc_gpu.h -- C ABI of libswpc_gpu, the MI355X (gfx950) implementation of
 * OpenSWPC's 3D velocity-stress time loop.
 *
 * The reference exposes this path as argument-less Fortran module routines
 * sharing module globals.  Each entry point below replaces one of them and
 * says which (file:line in the reference tree).  A Fortran caller binds them
 * with iso_c_binding interfaces.
 *
 * Conventions
 *   - Every function returns 0 on success and a negative SWPC_E* code on
 *     failure; swpc_last_error() gives the text.  The library never aborts.
 *   - Host arrays passed in use the reference layout: 3D arrays are (k,i,j)
 *     with k fastest and a 3-cell halo on every side (indices -2..n+3,
 *     src/swpc_3d/m_global.f90:295-300); (i,j) maps carry the same halo.
 *     The library keeps its own x-fastest layout on the device.
 *   - Arithmetic is IEEE FP32 in the reference's expression order (the FP32
 *     build, MP = SP at src/swpc_3d/m_global.f90:30).
 *   - A context is used by one host thread at a time.  Calls are ordered on
 *     the context's HIP stream; swpc_get_* calls synchronise it.
 */
#ifndef SWPC_GPU_H
#define SWPC_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SWPC_OK 0
#define SWPC_EINVAL -1  /* bad argument                      */
#define SWPC_ESTATE -2  /* call out of order / missing setup  */
#define SWPC_EHIP -3    /* HIP runtime error                  */
#define SWPC_ENOMEM -4  /* device allocation failed           */

#define SWPC_ABC_PML 0    /* abc_type = 'pml'    */
#define SWPC_ABC_CERJAN 1 /* abc_type = 'cerjan' */

/* field ids for swpc_get_field / swpc_set_field */
enum swpc_field {
    SWPC_VX = 0, SWPC_VY, SWPC_VZ,
    SWPC_SXX, SWPC_SYY, SWPC_SZZ, SWPC_SYZ, SWPC_SXZ, SWPC_SXY,
    SWPC_NFIELD
};

/* receiver output kinds (bit flags, m_wav.f90:29-33 sw_wav_*) */
#define SWPC_WAV_V 1
#define SWPC_WAV_U 2
#define SWPC_WAV_STRESS 4
#define SWPC_WAV_STRAIN 8

typedef struct swpc_ctx swpc_ctx;

typedef struct swpc_desc {
    int nx, ny, nz;     /* cells (m_global.f90:52)                    */
    int na;             /* absorber thickness (m_global.f90:82)       */
    int nm;             /* Zener mechanisms, 0..8 (m_global.f90:31)   */
    int abc;            /* SWPC_ABC_PML or SWPC_ABC_CERJAN            */
    double dx, dy, dz;  /* grid widths [km], rounded to float         */
    float dt;           /* time step [s]                              */
    int device;         /* HIP device ordinal                         */
} swpc_desc;

/* Context: allocates all device state (fields, memory variables, PML
   auxiliaries) zero-initialised.  Replaces the allocations of
   kernel__setup/memory_allocate (src/swpc_3d/m_kernel.f90:30-73,376-400)
   and absorb_p__setup (src/swpc_3d/m_absorb_p.f90:95-116). */
int swpc_create(const swpc_desc *desc, swpc_ctx **out);
int swpc_destroy(swpc_ctx *ctx);
const char *swpc_last_error(const swpc_ctx *ctx);

/* Medium and free-surface bands: the arrays medium__setup leaves in module
   m_global (src/swpc_3d/m_medium.f90:36-229): rho, relaxed lam/mu,
   tau-method taup/taus, and the 2nd-order band limits kfs_top..kob_bot
   plus kob (used by kernel__vmax). */
int swpc_set_medium(swpc_ctx *ctx, const float *rho, const float *lam, const float *mu, const float *taup,
                    const float *taus, const int32_t *kfs_top, const int32_t *kfs_bot, const int32_t *kob_top,
                    const int32_t *kob_bot, const int32_t *kob);

/* Zener recursion coefficients c1, c2, d1 (nm each) and d2
   (src/swpc_3d/m_kernel.f90:58-67). */
int swpc_set_visco(swpc_ctx *ctx, const float *c1, const float *c2, const float *d1, float d2);

/* ADE-CFS PML profiles g(4,n) at cell centre and edge, n = nx, ny, nz
   (src/swpc_3d/m_absorb_p.f90:74-93). */
int swpc_set_pml(swpc_ctx *ctx, const float *gxc, const float *gxe, const float *gyc, const float *gye,
                 const float *gzc, const float *gze);

/* Cerjan sponge profiles with the 3-cell halo (n+6 each)
   (src/swpc_3d/m_absorb_c.f90:27-101). */
int swpc_set_cerjan(swpc_ctx *ctx, const float *gx_c, const float *gx_b, const float *gy_c, const float *gy_b,
                    const float *gz_c, const float *gz_b);

/* Sources (source__setup, src/swpc_3d/m_source.f90:41-314, after the
   reference's own host set-up).  bf_mode = 0: moment tensors, comp is
   (6, nsrc) = mxx myy mzz myz mxz mxy per source, mo the seismic moment of
   each source; bf_mode = 1: body forces, comp is (3, nsrc) = fx fy fz and mo
   is ignored.  M0 is the normalisation (total moment or force norm).  stf is
   (nt, nsrc): the moment rate of each source at t = tbeg + (it - 1/2) dt for
   moment tensors (m_source.f90:800) or t = tbeg + it dt for body forces
   (m_source.f90:873).  Sources must lie inside [1,nx]x[1,ny]x[1,nz]. */
int swpc_set_sources(swpc_ctx *ctx, int nsrc, int bf_mode, const int32_t *isrc, const int32_t *jsrc,
                     const int32_t *ksrc, const float *mo, const float *comp, float M0, int nt, const float *stf);

/* Receivers (wav__setup, src/swpc_3d/m_wav.f90:46-137 and set_stinfo
   :141-271, after station mapping on the host).  what = SWPC_WAV_* flags,
   ntw = samples per trace.  M0 and UC are the output scales of
   m_wav.f90:517-607. */
int swpc_set_receivers(swpc_ctx *ctx, int nst, const int32_t *ist, const int32_t *jst, const int32_t *kst,
                       int what, int ntdec_w, int ntw, float M0, float UC);

/* ---- one time step, fine grained (reference routine for routine) ---- */
/* kernel__update_stress (src/swpc_3d/m_kernel.f90:142-348) */
int swpc_kernel_update_stress(swpc_ctx *ctx);
/* absorb__update_stress (src/swpc_3d/m_absorb.f90:60-75) */
int swpc_absorb_update_stress(swpc_ctx *ctx);
/* source__stressglut (src/swpc_3d/m_source.f90:776-848) */
int swpc_source_stressglut(swpc_ctx *ctx, int it);
/* kernel__update_vel (src/swpc_3d/m_kernel.f90:75-140) */
int swpc_kernel_update_vel(swpc_ctx *ctx);
/* absorb__update_vel (src/swpc_3d/m_absorb.f90:43-58) */
int swpc_absorb_update_vel(swpc_ctx *ctx);
/* source__bodyforce (src/swpc_3d/m_source.f90:850-898) */
int swpc_source_bodyforce(swpc_ctx *ctx, int it);
/* wav__store (src/swpc_3d/m_wav.f90:397-625) */
int swpc_wav_store(swpc_ctx *ctx, int it);

/* Both updates of one phase as the time loop issues them: kernel__update_stress
   followed by absorb__update_stress (src/swpc_3d/main.f90:126-127), and
   kernel__update_vel followed by absorb__update_vel (main.f90:132-133).  With
   the PML each is one fused launch over the whole grid. */
int swpc_update_stress(swpc_ctx *ctx);
int swpc_update_vel(swpc_ctx *ctx);

/* nsteps whole steps it0 .. it0+nsteps-1 in the order of the main loop
   (src/swpc_3d/main.f90:119-139): green__store, wav__store, snap__write,
   stress, glut, vel, body force, green__source.  With use_graph != 0 the steps run as replays of a captured
   HIP graph of `block` steps (falls back to direct launches for the tail). */
int swpc_step(swpc_ctx *ctx, int it0, int nsteps, int use_graph, int block);

/* ---- wavefield sections (snap__setup / snap__write,
   src/swpc_3d/m_snap.f90:95-322, 919-2189) ---- */
enum swpc_snap_plane { SWPC_SNP_XY = 0, SWPC_SNP_YZ, SWPC_SNP_XZ, SWPC_SNP_FS, SWPC_SNP_OB };
enum swpc_snap_qty { SWPC_SNP_PS = 0, SWPC_SNP_V, SWPC_SNP_U };
#define SWPC_SNP_NSECTION 15
/* section id of plane p and quantity q: xy_ps = 0, xy_v = 1, ... ob_u = 14 */
#define SWPC_SNP_SECTION(p, q) ((p) * 3 + (q))

typedef struct swpc_snap_desc {
    int sw[SWPC_SNP_NSECTION]; /* nonzero: section on (xy_ps%sw .. ob_u%sw, m_snap.f90:105-119) */
    int k0_xy, i0_yz, j0_xz;   /* section planes as grid indices (z2k(z0_xy) etc., m_snap.f90:158-160) */
    int idec, jdec, kdec;      /* spatial decimation: section point n samples grid n*dec - dec/2 */
    int ntdec_s;               /* a frame at every step it with mod(it-1, ntdec_s) == 0 */
    int nframe;                /* frames kept on the device per section: frame f = (it-1)/ntdec_s */
    float M0, UC;              /* output scales, as for the receivers */
} swpc_snap_desc;

/* Enable sections.  kfs is the free-surface map (i,j) with the halo
   (medium__setup, m_medium.f90:339-394), used by the fs sections.  The
   sections are (nxs, nys) for xy/fs/ob, (nys, nzs) for yz and (nxs, nzs)
   for xz, with nxs = (nx + idec/2)/idec etc. (m_snap.f90:130-132).  Frames,
   u accumulators and maximum maps start at zero.  Replaces snap__setup. */
int swpc_set_snapshots(swpc_ctx *ctx, const swpc_snap_desc *desc, const int32_t *kfs);

/* snap__write (src/swpc_3d/m_snap.f90:919-947): takes the frames of step it
   (ps and v at output steps; u accumulation and the fs/ob maxima every
   step).  swpc_step calls it after wav__store, as main.f90:123-124 does.
   Fails if a frame index would reach nframe. */
int swpc_snap_write(swpc_ctx *ctx, int it);

/* Frame f of a section as (ns1, ns2, ncomp) in Fortran order, ncomp = 4
   for ps (div, rot_x, rot_y, rot_z) and 3 for v and u: the buffer
   the reference writes per frame (m_snap.f90:1054-1057). */
int swpc_get_snapshot(swpc_ctx *ctx, int section, int frame, float *out);

/* Maximum-amplitude map (nxs, nys, 3) = (max |z|, max horizontal, their
   norm) of an fs or ob, v or u section (max_fs_v etc., m_snap.f90:85,
   written by output__put_maxval :2295-2348). */
int swpc_get_snap_max(swpc_ctx *ctx, int section, float *out);

/* dims = (ns1, ns2, ncomp) of a section. */
int swpc_snap_dims(const swpc_ctx *ctx, int section, int dims[3]);

/* Launch slots timed by swpc_step_timed, in step order. */
enum swpc_timer {
    SWPC_T_WAV = 0,       /* green__store + wav__store                           */
    SWPC_T_SNAP,          /* snap__write                                         */
    SWPC_T_STRESS_SIDE,   /* absorb__update_stress, side columns (x and y slabs)  */
    SWPC_T_STRESS,        /* kernel__update_stress + its bottom PML band         */
    SWPC_T_STRESS_ABC,    /* absorb__update_stress with the Cerjan sponge        */
    SWPC_T_GLUT,          /* source__stressglut                                  */
    SWPC_T_VEL_SIDE,      /* absorb__update_vel, side columns                    */
    SWPC_T_VEL,           /* kernel__update_vel + its bottom PML band            */
    SWPC_T_VEL_ABC,       /* absorb__update_vel with the Cerjan sponge           */
    SWPC_T_BODYFORCE,     /* source__bodyforce + green__source                   */
    SWPC_NTIMER
};

/* swpc_step with direct launches and HIP events between them: ms[t] receives
   the milliseconds of slot t summed over the nsteps steps (the per-routine
   breakdown of the reference's stopwatch_mode, src/shared/m_pwatch.f90:35-195,
   src/swpc_3d/main.f90:66).  Synchronises after every step. */
int swpc_step_timed(swpc_ctx *ctx, int it0, int nsteps, float *ms);

/* kernel__vmax (src/swpc_3d/m_kernel.f90:350-374): max |V| one cell below
   the ocean bottom, 5 cells inside the absorber. Synchronises. */
int swpc_kernel_vmax(swpc_ctx *ctx, float out[3]);

/* swpc_kernel_vmax without allocation.  No counterpart in the reference: it
   is for callers that cannot allocate device memory (code under HIP graph
   capture, tight serving loops).  Writes the three maxima to d_out, device
   memory of the context's GPU; stream-ordered on the context's stream, does
   not synchronise.  Its partial maxima live in a fixed per-GPU scratch array
   of the library (16384 blocks of 256 surface cells, 192 KiB); a larger
   sampled surface returns SWPC_EINVAL.  Calls that overlap on one GPU share
   that array and race. */
int swpc_kernel_vmax_noalloc(swpc_ctx *ctx, float *d_out);

/* Field transfer in the reference (k,i,j) halo-3 layout, FP32. */
int swpc_get_field(swpc_ctx *ctx, int field, float *out);
int swpc_set_field(swpc_ctx *ctx, int field, const float *in);

/* Reads of the medium the kernels use (the module arrays rho, lam, mu, taup,
   taus and the band maps of src/swpc_3d/m_medium.f90:31-44,339-394):
   which = SWPC_MED_RHO..SWPC_MED_TAUS give (k,i,j) halo-3 float arrays,
   SWPC_MED_KFS_TOP..SWPC_MED_KOB give (i,j) halo-3 int32 maps. */
enum { SWPC_MED_RHO = 0, SWPC_MED_LAM, SWPC_MED_MU, SWPC_MED_TAUP, SWPC_MED_TAUS,
       SWPC_MED_KFS_TOP, SWPC_MED_KFS_BOT, SWPC_MED_KOB_TOP, SWPC_MED_KOB_BOT, SWPC_MED_KOB, SWPC_NMEDIUM };
int swpc_get_medium(swpc_ctx *ctx, int which, void *out);

/* medium__setup of a layered model, built on the device
   (src/swpc_3d/m_medium.f90:124-427 after vmodel_lhm / vmodel_lgm,
   src/swpc_3d/m_vmodel_lhm.f90:20-158, m_vmodel_lgm.f90:20-174, with the
   per-cell random perturbation of m_vmodel_lgm_rmed.f90:195-203): the same
   kernel inputs swpc_set_medium takes, without the 3D arrays crossing from
   the host.  prof: (5, nz+6) floats, rows rho, lam, mu, Qp, Qs of each plane
   k = -2..nz+3 of the layered model.  xi: the relative perturbation on the
   (k,i,j) halo grid, in host or device memory, or NULL.  clip: vmin, vmax,
   rhomin of the perturbation (vcheck, src/shared/m_fdtool.f90:15-48).
   tau: zeta and the real and imaginary parts of the mean
   cc = <i w ts / (1 - i w ts)> at the reference frequency.  Absorber cells
   copy the nearest interior cell; the free surface, ocean bottom and their
   2nd-order bands are detected on the device.  Returns the kfs and kob maps
   ((i,j) halo-3 int32, for stations and sections) and vrange = vmin (of
   vs >= epsilon), vmax (of vp) below the free surface.  Synchronises. */
int swpc_medium_layered(swpc_ctx *ctx, const float *prof, const float *xi, const float clip[3], const float tau[3],
                        int32_t *kfs, int32_t *kob, float vrange[2]);

/* Seismograms (ntw, ncmp, nst) like wav_vel/wav_disp/wav_stress/wav_strain
   (src/swpc_3d/m_wav.f90:37-38).  kind = one SWPC_WAV_* flag. */
int swpc_get_seismograms(swpc_ctx *ctx, int kind, float *out);

/* ---- reciprocity Green's-function mode (src/swpc_3d/m_green.f90) ----
   The device part of green__setup (m_green.f90:66-404): ng GF cells
   (ig, jg, kg; 1-based indices the host has already filtered to the solid
   part, m_green.f90:225-296), the virtual source cell (isrc, jsrc, ksrc) of
   the station named by green_stnm, the force direction cmp (0 = 'x',
   1 = 'y', 2 = 'z'), body-force Green's functions on (bforce != 0: ncmp = 9)
   or off (ncmp = 6), output decimation ntdec_w and length ntw, and the
   moment rate stf[it-1] at t = tbeg + it*dt for it = 1..nt (m_green.f90:632).
   Accumulators and outputs start at zero, as at it == 1 of green__store.
   cmp < 0 switches the mode off.  swpc_step runs green__store first and
   green__source last in every step (main.f90:122,137). */
int swpc_set_green(swpc_ctx *ctx, int ng, const int32_t *ig, const int32_t *jg, const int32_t *kg,
                   int isrc, int jsrc, int ksrc, int cmp, int bforce, int ntdec_w, int ntw, int nt,
                   const float *stf);
/* green__store(it) (m_green.f90:406-545) and green__source(it)
   (m_green.f90:614-649) */
int swpc_green_store(swpc_ctx *ctx, int it);
int swpc_green_source(swpc_ctx *ctx, int it);
/* gf(ntw, ncmp*ng) as green__store leaves it (m_green.f90:56), Fortran order:
   column n*ncmp + c holds mxx myy mzz myz mxz mxy [fx fy fz] of GF cell n.
   green__export's sign flip for cmp = 'z' (m_green.f90:563-566) is the
   caller's.  out holds ntw*ncmp*ng floats.  Synchronises. */
int swpc_get_green(swpc_ctx *ctx, float *out);

/* ---- random media (src/shared/m_rmedia.f90:40-173, rmedia__3dgen_ani) ----
   A realisation of a 3D random medium on an n1 x n2 x n3 grid, axis 1
   fastest in out (n1*n2*n3 floats in host memory, or in device memory of
   `device`, which is written directly): uniform noise, forward
   FFT, whitening, filtering by the square root of the PSDF of type ptype
   (1 Gaussian, 2 exponential, 3 von Karman with Hurst number kappa;
   correlation lengths a1..a3, spacings d1..d3), zero mean, rms scaled to
   epsil.  The reference's media(nx, ny, nz) is (n1, n2, n3) = (nx, ny, nz);
   a (k,i,j) array is (nz, nx, ny).  The noise is a counter-based stream of
   seed, so equal arguments give bit-identical fields on any launch shape;
   any sizes (the reference's FFT needs powers of two).  Runs on device
   `device` and synchronises; swpc_rmedia_last_error() gives the text of a
   failure of the calling thread. */
int swpc_rmedia_3d(int device, int ptype, int n1, int n2, int n3, float d1, float d2, float d3, float a1,
                   float a2, float a3, float epsil, float kappa, uint64_t seed, float *out);
const char *swpc_rmedia_last_error(void);

/* ---- batches: many independent simulations stepped together ----
   The reference runs one simulation per process; a batch of Green's
   functions (one run per source component and station, the workload of
   src/swpc_3d/m_green.f90) there means many separate runs.  A batch steps
   member contexts (created and set up as usual, same device and same nm;
   grids, absorbers, sources and receivers may differ) with one launch per
   phase for all members.  Members keep their state and outputs: read them
   with swpc_get_seismograms etc.  Do not step a member on its own while its
   batch steps (the batch uses its own stream); tuning changes made to
   members are picked up by the next swpc_batch_step. */
typedef struct swpc_batch swpc_batch;
int swpc_batch_create(swpc_ctx *const *members, int n, swpc_batch **out);
int swpc_batch_destroy(swpc_batch *batch);
const char *swpc_batch_last_error(const swpc_batch *batch);
/* nsteps whole steps it0 .. it0+nsteps-1 of every member (main.f90:119-139
   order), as replays of a captured graph of `block` steps when use_graph. */
int swpc_batch_step(swpc_batch *batch, int it0, int nsteps, int use_graph, int block);
int swpc_batch_synchronize(swpc_batch *batch);
/* Autotuner: times candidate k-chunk depths, tile rows, interior tile
   widths (64 or 32 cells) and member packings on trial_steps dry steps
   (step index 0: no source, no recording) and keeps the fastest.  Call it
   after set-up and before the first step, while every field is zero, so the
   trials leave the state untouched.  Outputs may be NULL; the packing chosen
   is read with swpc_batch_get_packing. */
int swpc_batch_tune(swpc_batch *batch, int trial_steps, int *kchunk, int *by, int *width, float *ms_per_step);
/* Order of the members' tiles in the batched launches: 0 member-major,
   members largest first (the default); 1 chunk-major, every member's first
   k chunk, then every member's second, and so on; 2 member-major with each
   member's side-column PML tiles in the interior launch, next to the tile
   rows they border (one launch per phase).  Results do not depend on it. */
int swpc_batch_set_packing(swpc_batch *batch, int packing);
int swpc_batch_get_packing(const swpc_batch *batch);
void *swpc_batch_stream(swpc_batch *batch);

/* ---- discrete adjoint (ADE-CFS PML; elastic or with Zener memory variables) ----
   Misfit J = 1/2 sum over recorded samples, stations and components of
   (d - d_obs)^2, d the velocity traces of swpc_get_seismograms(SWPC_WAV_V),
   and its gradient with respect to the medium arrays lam, mu, rho of
   swpc_set_medium, by the discrete adjoint of the time step.  None of these
   routines has a counterpart in the reference (which has no adjoint); the
   forward they differentiate is the one of src/swpc_3d/main.f90:119-139.

   Checkpointing, two levels: a first forward pass of nt steps from rest with
   the context's sources records the traces and keeps the state (fields, PML
   auxiliaries, memory variables) every ckp_outer steps; walking backward,
   each outer segment is recomputed once to keep inner checkpoints every
   ckp_inner steps, and each inner segment once more into per-step records
   (V_n, S_{n+1} and their auxiliaries) that the adjoint steps read.  Device
   memory: ceil(nt/ckp_outer) + ceil(ckp_outer/ckp_inner) - 1 states and
   ckp_inner step records, besides the adjoint fields.  ckp_inner = 0 or
   >= ckp_outer is the single-level scheme (ckp_outer states, ckp_outer
   records, two forward passes).  Every scheme gives the same bits.

   Gradients of several sources sum on the device: swpc_adj_reset, then per
   source swpc_set_sources / swpc_set_receivers and swpc_adj_accumulate, then
   swpc_adj_get_gradient.  The context keeps the adjoint's device memory
   between these calls until swpc_adj_release or swpc_destroy.  The context's
   fields are left in an intermediate state. */

/* The scheme with the fewest forward steps whose device memory fits budget
   bytes (budget <= 0: the device's free memory, plus what the adjoint
   already holds, less 1 GiB).  Outputs may be NULL: outer and inner spacing,
   the device bytes it needs, and the forward steps it runs. */
int swpc_adj_plan(swpc_ctx *ctx, int nt, int64_t budget, int *ckp_outer, int *ckp_inner, int64_t *bytes,
                  int64_t *fwd_steps);
/* Zero the device gradient sums (allocating the adjoint fields). */
int swpc_adj_reset(swpc_ctx *ctx);
/* Misfit of the current sources and receivers against obs ((ntw, 3, nst)
   like the traces) and its gradient, added to the device sums. */
int swpc_adj_accumulate(swpc_ctx *ctx, int nt, const float *obs, int ckp_outer, int ckp_inner, double *misfit);
/* The sums in the reference (k,i,j) halo layout for cells 1..n (halo
   entries zero); any output pointer may be NULL. */
int swpc_adj_get_gradient(swpc_ctx *ctx, float *g_lam, float *g_mu, float *g_rho);
/* Free the adjoint's device memory (sums included). */
int swpc_adj_release(swpc_ctx *ctx);
/* One source in one call: reset, accumulate with single-level checkpoints
   every ckp steps, get, release. */
int swpc_adj_gradient(swpc_ctx *ctx, int nt, const float *obs, int ckp, double *misfit, float *g_lam, float *g_mu,
                      float *g_rho);

/* Wait for all work queued on the context. */
int swpc_synchronize(swpc_ctx *ctx);

/* Device bytes held by the context (for reporting). */
int64_t swpc_device_bytes(const swpc_ctx *ctx);

/* Device memory striped over the HBM regions.  No counterpart in the
   reference (whose arrays are plain Fortran allocations on the host): the
   placement of the per-cell arrays in HBM.  An MI355X's 288 GiB fall into
   three 96 GiB regions (nregions = 3); the stencil updates stream 10 to 18
   arrays at once and run 12-14 % faster with those arrays spread over the
   regions than with all of them in one, where hipMalloc puts them on an idle
   device.  swpc_stripe_reserve holds every free 256 MiB granule of the
   device for a moment, keeps bytes_per_region from the middle of each
   region's share of them and releases the rest.  Contexts created on the
   device afterwards take V, S, the medium, the memory variables and the PML
   auxiliaries from this arena, each array (or stream segment of a packed
   array) in a fixed region; they fall back to hipMalloc when it is full.
   Results do not depend on the placement.  swpc_stripe_release returns
   SWPC_ESTATE while any context still holds arena memory;
   swpc_stripe_free_bytes reports what is left in a region (0 without an
   arena).  Without a reserved arena, swpc_create gives a context whose
   per-cell arrays come to 4 GiB or more a private arena sized for it (only
   that context draws from it; it goes at swpc_destroy); SWPC_STRIPE=0 in the
   environment turns that off. */
int swpc_stripe_reserve(int device, int64_t bytes_per_region, int nregions);
int swpc_stripe_release(int device);
int64_t swpc_stripe_free_bytes(int device, int region);

/* Launch-shape tuning knobs.  kchunk: k-planes marched per thread block
   (0 = keep, -1 = the default: 16, 32 when nm = 0, halved on small grids
   until the interior launch has two blocks per CU); by: rows per 64-wide tile, 1, 2
   or 4 (0 = keep; default 4); variant: kernel variant flags (-1 = keep;
   default 9): bit 0 = non-temporal loads and
   stores for the single-use streams of the stress update, bit 1 = side-column
   PML kernels on a second stream alongside the interior ones, bit 2 =
   interior tiles 32 cells wide with 2 by rows (narrow grids idle fewer
   lanes), bit 3 = the interior velocity update loads plane k+1 while it
   computes plane k, bit 4 = the adjoint copies its step records instead of
   stepping out of place into them, bit 5 = the interior and side columns of
   a phase in one launch, x-slab tiles next to the interior rows they
   border, bit 7 = the velocity update of a step and the stress update of
   the next in one interior launch (PML runs without body forces or the
   reciprocity mode; the context then holds a second set of the 9 fields,
   taken from the striped arena when it has room; kchunk sets the launch's
   chunk depth), bit 8 = that launch with 8-row tiles instead of 4.  All
   variants give bit-identical results.  Fails with SWPC_ENOMEM when the
   second field set does not fit. */
int swpc_set_tuning(swpc_ctx *ctx, int kchunk, int by, int variant);

/* The HIP stream the context launches on (hipStream_t as void*), so a
   caller can time or order work against it. */
void *swpc_stream(swpc_ctx *ctx);

/* The context's staging buffer: device memory of (nz+6)(nx+6)(ny+6) floats
   (*nfloats), the size of one (k,i,j) halo array, which the library uses for
   layout transposes.  A caller may fill it between calls, e.g. with a random
   medium from swpc_rmedia_3d, and pass it as the xi of swpc_medium_layered;
   the next transfer call overwrites it.  NULL for a NULL context. */
void *swpc_staging_buffer(swpc_ctx *ctx, int64_t *nfloats);

#ifdef __cplusplus
}
#endif

#endif /* SWPC_GPU_H */
