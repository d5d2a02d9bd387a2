// swpc_internal.h -- device layout and context of libswpc_gpu (gfx950).
//
// Device layout (all FP32, x fastest):
//   3D halo arrays (fields V3/S6, medium rho/lam/mu/taup/taus):
//     idx(i,j,k) = (k+2)*plane + (j+2)*pitch + (i-1+ox),  i,j,k in -2..n+3
//     ox = 32 puts i = 1 on a 128-byte line; pitch is a multiple of 32 floats.
//   Memory variables: 6*NM values per cell packed into float4 streams (and a
//     float2 tail) over the kernel rows j in [jb,je], planes
//     k in [kb,ke], with full-width rows: ridx(i,j,k) = ((k-kb)*nyk + (j-jb))*rpitch + (i-1).
//   PML auxiliaries: 18 arrays over the PML cells only, indexed by a compact
//     cell number running over five boxes (two y slabs, two x slabs, bottom);
//     the bottom box keeps full aligned rows (i from 1, rpitch wide).
//   (i,j) maps: (j+2)*pitch + (i-1+ox).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/swpc_gpu.h"

#define SWPC_MAXNM 8

struct Grid {
    int nx, ny, nz, na, nm;
    int ox, pitch, nyp, nzp;  // nyp = ny+6 rows, nzp = nz+6 planes
    long long plane, n;       // floats per plane / per array
    // kernel (interior) box, m_global.f90:345-376
    int ib, ie, jb, je, kb, ke;
    int nxk, nyk, nzk, rpitch;
    long long rplane, rn;
};

__host__ __device__ static inline long long gidx(const Grid &g, int i, int j, int k)
{
    return (long long)(k + 2) * g.plane + (long long)(j + 2) * g.pitch + (i - 1 + g.ox);
}
__host__ __device__ static inline long long g2idx(const Grid &g, int i, int j)
{
    return (long long)(j + 2) * g.pitch + (i - 1 + g.ox);
}

// Blocks are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8).
// Give each XCD a contiguous run of logical tiles so tiles that share halo
// rows share one L2.
__device__ static inline int xcd_remap(int b, int n)
{
    const int x = b & 7, t = b >> 3;
    const int q = n >> 3, r = n & 7;
    return x < r ? x * (q + 1) + t : r * (q + 1) + (x - r) * q + t;
}

// one of the five PML boxes (m_global.f90:334-343 kbeg_a partition)
struct PmlBox {
    int i0, j0, k0;       // first cell
    int nxb, nyb, nzb;    // extent
    long long off;        // first compact cell number
};

// 4th-order coefficients for the two band states:
// [0] outside the band (isign = -1), [1] inside (isign = +1)
struct Coef4 {
    float re40x[2], re41x[2], re40y[2], re41y[2], re40z[2], re41z[2];
};

struct FieldPtrs {
    float *Vx, *Vy, *Vz;
    float *Sxx, *Syy, *Szz, *Syz, *Sxz, *Sxy;
};

struct MediumPtrs {
    const float *rho, *lam, *mu, *taup, *taus;
    const int *kfs_top, *kfs_bot, *kob_top, *kob_bot, *kob;
};

struct ViscoCoef {
    float c1[SWPC_MAXNM], c2[SWPC_MAXNM], d1[SWPC_MAXNM];
    float d2;
};

// ADE auxiliaries, order of m_absorb_p.f90:47-52
enum {
    A_axVx = 0, A_ayVx, A_azVx, A_axVy, A_ayVy, A_azVy, A_axVz, A_ayVz, A_azVz,
    A_axSxx, A_aySxy, A_azSxz, A_axSxy, A_aySyy, A_azSyz, A_axSxz, A_aySyz, A_azSzz,
    A_N
};


struct SrcTarget {
    long long idx;   // device index in the field array
    int field;       // SWPC_* field id
    int start, count;
};
struct SrcContrib {
    int src;         // source number
    int comp;        // MT: 0..5 = xx yy zz yz xz xy ; BF: 0..2 = fx fy fz
    int div4;        // MT shear: divide by 4 (m_source.f90:815-839)
    int pad;
    long long rho_a, rho_b;  // BF: the two densities of the staggered buoyancy
};

// one enabled section of snap__write
struct SnapTask {
    int plane, q, ns1, ns2;
    int blk0;            // first block of the task in the snapshot launch
    float *frame;        // (ns1, ns2, ncomp, nframe)
    float *uacc;         // (ns1, ns2, 3) for u sections
    float *maxmap;       // (ns1, ns2, 3) for fs/ob v and u sections
};

struct swpc_ctx {
    swpc_desc desc;
    Grid g;
    int device;
    hipStream_t stream;
    hipStream_t stream2;            // PML kernels alongside the interior ones
    hipEvent_t ev_fork, ev_join;
    char err[512];

    float *fields[SWPC_NFIELD];
    // when set, the step reads the old values of the fields it updates
    // (and of the PML auxiliaries) here and writes the new ones to fields
    // (ade): the adjoint records steps next to their inputs this way
    float *fields_in[SWPC_NFIELD];
    const float *ade_in;
    float *rho, *lam, *mu, *taup, *taus;
    int *kfs_top, *kfs_bot, *kob_top, *kob_bot, *kob;
    float *R;          // 6*nm*rn
    ViscoCoef visco;
    bool has_medium, has_visco, has_pml, has_cerjan;

    // PML
    PmlBox box[5];
    long long npml;
    float *ade;
    float4 *gxc, *gxe, *gyc, *gye, *gzc, *gze;
    int *d_pml_tiles;  // int4 (i0 | width << 20 | xslab << 28, j0, k0, k1) per side-column tile
    int npml_tiles;
    // Cerjan
    float *cgx_c, *cgx_b, *cgy_c, *cgy_b, *cgz_c, *cgz_b;

    Coef4 coef;

    // sources
    int nsrc, bf_mode, src_nt;
    int ntarget;
    SrcTarget *d_targets;
    SrcContrib *d_contrib;
    float *d_comp;     // (ncomp, nsrc), normalised by M0 (m_source.f90:297-303)
    float *d_mo;       // nsrc, mo/M0
    float *d_stf;      // (nt, nsrc)
    float dt_dxyz;

    // receivers
    int nst, wav_what, ntdec_w, ntw;
    float wav_M0, wav_UC;
    int *d_ist, *d_jst, *d_kst;
    float *d_uacc, *d_eacc;
    float *d_wav[4];   // V, U, stress, strain: (ntw, ncmp, nst)

    // reciprocity Green's functions (m_green.f90)
    int green_on;
    int green_ng, green_ncmp, green_ntdec, green_ntw, green_nt;
    int green_src[3];  // virtual source cell (a station)
    float green_f[3];  // unit force direction (fx1, fy1, fz1)
    float green_dt_dxyz;
    int *d_green_ijk;  // ig | jg | kg, ng each
    float *d_green_acc;  // (12, ng)
    float *d_gf;       // (ntw, ncmp*ng)
    float *d_green_stf;  // nt

    // sections (snapshots)
    swpc_snap_desc snap;
    int snap_ntask, snap_nblk;
    SnapTask snap_task[SWPC_SNP_NSECTION];
    int snap_dims[SWPC_SNP_NSECTION][3];
    float *d_snap_frame[SWPC_SNP_NSECTION];
    float *d_snap_u[5];
    float *d_snap_max[SWPC_SNP_NSECTION];
    int *d_kfs;

    // graphs and device step counter
    int *d_it;
    hipGraphExec_t graph_exec;
    int graph_block;
    unsigned gen;      // bumped whenever launch arguments change

    // staging for layout transposes
    float *staging;
    long long staging_n;
    float *d_vmax;

    int kchunk, by;
    int nbz;           // k chunks of the interior-column kernels
    int *d_zchunks;    // first plane of each chunk, stress phase then velocity phase (2 x (nbz+1))
    int zband[2];      // per phase, the first chunk that holds planes of the bottom PML band
    int *d_fzc;        // chunks of the fused launch over the interior planes (nfz + 1 entries)
    int nfz;
    float *fields2[SWPC_NFIELD];  // second field set of the fused steps (variant bit 7)
    int4 *d_sched[2];  // merged interior + side-column schedules of the two phases (variant bit 5)
    int nsched[2];
    int variant;       // bit 0: non-temporal single-use streams in the interior stress kernel,
                       // bit 1: side-column PML kernels on a second stream,
                       // bit 2: 32-wide interior tiles, bit 3: velocity loads one plane ahead,
                       // bit 4: adjoint records copied, bit 5: interior and side columns in one launch,
                       // bit 7: the velocity update fused to the next stress update (k_fused, two field
                       // sets), bit 8: fused tiles of 8 rows instead of 4
    long long bytes;
    int own_arena;     // the context reserved the device's striped arena itself (swpc_create)
    struct AdjWorkspace *adj_ws;  // adjoint state, gradient sums and checkpoints (adjoint.hip)
};

// records msg as the context's last error and returns code (api.hip)
int ctx_fail(swpc_ctx *c, int code, const char *msg);
// launch arguments changed: drops the captured graphs (api.hip)
void ctx_reset_graph(swpc_ctx *c);

// ---- striped device arena (arena.hip) ----
// memory from the arena reserved on the device, or nullptr when there is
// none or it is full; an array of nseg segments of seg_bytes each puts
// segment s in region seg_region[s] (nseg = 1: the whole array in
// seg_region[0]; seg_region = nullptr: segment s in region s mod the
// region count); a context's private arena serves only `who`
bool arena_active(int device);
void *arena_alloc(int device, const void *who, size_t bytes, const int *seg_region, int nseg, size_t seg_bytes);
// swpc_stripe_reserve with an owner: a private arena that serves only it
int arena_reserve(int device, int64_t bytes_per_region, int nregions, const void *owner);
// bytes of whole free granules in the device's arena that `who` may use (0 without one)
int64_t arena_granule_bytes(int device, const void *who);
// returns the block to its arena; false when p is not arena memory
bool arena_free(void *p);

// ---- launchers (stencil.hip / aux.hip) ----
// mask: 1 interior, 2 PML, 3 both; the PML kernel goes to pml_stream when
// given (it updates cells disjoint from the interior kernel's)
hipError_t launch_stress(swpc_ctx *c, int mask, hipStream_t pml_stream);
hipError_t launch_vel(swpc_ctx *c, int mask, hipStream_t pml_stream);
hipError_t launch_stress_side(swpc_ctx *c, hipStream_t s);
hipError_t launch_vel_side(swpc_ctx *c, hipStream_t s);
hipError_t launch_stress_cols(swpc_ctx *c, int mask);
int build_schedule(swpc_ctx *c);
bool merged_launch_ready(const swpc_ctx *c);
hipError_t launch_stress_all(swpc_ctx *c);
hipError_t launch_vel_all(swpc_ctx *c);
hipError_t launch_vel_cols(swpc_ctx *c, int mask);
hipError_t launch_fused(swpc_ctx *c);
hipError_t launch_kernel_stress(swpc_ctx *c);
hipError_t launch_kernel_vel(swpc_ctx *c);
hipError_t launch_pml_stress(swpc_ctx *c);
hipError_t launch_pml_vel(swpc_ctx *c);
hipError_t launch_cerjan_stress(swpc_ctx *c, hipStream_t s = nullptr);
hipError_t launch_cerjan_vel(swpc_ctx *c, hipStream_t s = nullptr);
hipError_t launch_stressglut(swpc_ctx *c, int it, const int *d_it, hipStream_t s = nullptr);
hipError_t launch_bodyforce(swpc_ctx *c, int it, const int *d_it, hipStream_t s = nullptr);
hipError_t launch_wav_store(swpc_ctx *c, int it, const int *d_it, hipStream_t s = nullptr);
hipError_t launch_green_store(swpc_ctx *c, int it, const int *d_it, hipStream_t s = nullptr);
hipError_t launch_green_source(swpc_ctx *c, int it, const int *d_it, hipStream_t s = nullptr);
hipError_t launch_step_advance(swpc_ctx *c);
hipError_t launch_snap(swpc_ctx *c, int it, const int *d_it, hipStream_t s = nullptr);

hipError_t launch_set_counter(int *d, int v, hipStream_t s);
hipError_t launch_counter_advance(int *d, hipStream_t s);

// ---- batched launches (stencil.hip) ----
size_t step_args_size();
void fill_batch_args(swpc_ctx *c, void *out, int phase);
void member_tiles(swpc_ctx *c, int &nbx, int &nby, int &nbz);
// batched receivers, sources, Green's-function stores and sources (aux.hip,
// green.hip): kinds of per-member argument arrays
enum { BATCH_AUX_WAV = 0, BATCH_AUX_GLUT = 1, BATCH_AUX_BODYFORCE = 2, BATCH_AUX_GREEN_STORE = 3,
       BATCH_AUX_GREEN_SOURCE = 4, BATCH_AUX_N = 5 };
size_t batch_aux_size(int kind);
// fills one member's arguments of a kind; returns its thread count (0: nothing to launch)
int batch_aux_fill(swpc_ctx *c, int kind, void *out);
hipError_t launch_batch_aux(int kind, const void *d_args, int nmem, int nmax, int it, const int *d_it, hipStream_t s);
size_t batch_green_size(int kind);
int batch_green_fill(swpc_ctx *c, int kind, void *out);
hipError_t launch_batch_green(int kind, const void *d_args, int nmem, int nmax, int it, const int *d_it, hipStream_t s);
// batched sections (snap.hip): returns the member's block count (0: no sections)
size_t batch_snap_size();
int batch_snap_fill(swpc_ctx *c, void *out);
hipError_t launch_batch_snap(const void *d_args, int nmem, int maxblk, int it, const int *d_it, hipStream_t s);
// batched Cerjan sponges (stencil.hip): phase 0 stress, 1 velocity; returns the member's block count
size_t batch_cerjan_size();
int batch_cerjan_fill(swpc_ctx *c, int phase, void *out);
hipError_t launch_batch_cerjan(int phase, const void *d_args, const int *d_blk0, int nmem, int nblocks, hipStream_t s);
hipError_t launch_batch_phase_all(int phase, int nm, int variant, int by, int bxw, const void *d_args,
                                  const int4 *d_all, int nall, hipStream_t s);
hipError_t launch_batch_phase(int phase, int nm, int variant, int by, int bxw, const void *d_args, const int4 *d_cols,
                              int ncols, const int4 *d_side, int nside, hipStream_t s);
hipError_t launch_vmax(swpc_ctx *c);
hipError_t launch_ref_to_dev(swpc_ctx *c, const float *src_ref, float *dst, hipStream_t s);
hipError_t launch_dev_to_ref(swpc_ctx *c, const float *src, float *dst_ref, hipStream_t s);
