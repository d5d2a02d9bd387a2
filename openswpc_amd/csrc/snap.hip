// snap.hip -- wavefield sections of libswpc_gpu (snap__write,
// src/swpc_3d/m_snap.f90:919-2189, native-format semantics).
//
// Every enabled section is a task of ns1 x ns2 points; one launch per step
// covers all tasks (one thread per point).  A frame is taken when
// mod(it-1, ntdec_s) == 0 and stored in the section's frame store on the
// device (frame f = (it-1)/ntdec_s), so graph-replayed step blocks need no
// host round trip.  u sections integrate V dt every step and the fs/ob
// sections keep running maximum maps, as the reference does.
#include <cstring>

#include "swpc_internal.h"

struct SnapArgs {
    Grid g;
    const float *Vx, *Vy, *Vz, *Syz, *Sxz, *Sxy, *lam;
    const int *kfs, *kob;
    SnapTask task[SWPC_SNP_NSECTION];
    int ntask, nblk;
    int k0_xy, i0_yz, j0_xz, idec, jdec, kdec, ntdec_s, nframe;
    float r20x, r20y, r20z, M0, UC, dt;
    int it;
    const int *d_it;
};

#define EPSL 1.1920928955078125e-07f  // epsilon(1.0)

// correctly rounded single-precision square root, as the reference's sqrt:
// the double root of a float rounded once more to float is exact-rounded
// (53 >= 2*24 + 2 bits), which the single-precision hardware root is not
__device__ static inline float sqrt_rn(float x) { return (float)__dsqrt_rn((double)x); }

// block blk of one context's section launch at step it
__device__ static inline void snap_body(const SnapArgs &a, int blk, int it)
{
    int t = 0;
    while (t + 1 < a.ntask && blk >= a.task[t + 1].blk0) ++t;
    const SnapTask T = a.task[t];
    const long long npt = (long long)T.ns1 * T.ns2;
    const long long p = (long long)(blk - T.blk0) * blockDim.x + threadIdx.x;
    if (p >= npt) return;
    if (it < 1) return;
    const bool out = ((it - 1) % a.ntdec_s) == 0;
    const int f = (it - 1) / a.ntdec_s;
    const bool every = (T.q == SWPC_SNP_U) || (T.q == SWPC_SNP_V && T.plane >= SWPC_SNP_FS);
    if (!(out && f < a.nframe) && !every) return;

    const int sa = (int)(p % T.ns1) + 1, sb = (int)(p / T.ns1) + 1;
    const Grid &g = a.g;
    int i, j, k;
    switch (T.plane) {
    case SWPC_SNP_XY: i = sa * a.idec - a.idec / 2; j = sb * a.jdec - a.jdec / 2; k = a.k0_xy; break;
    case SWPC_SNP_YZ: i = a.i0_yz; j = sa * a.jdec - a.jdec / 2; k = sb * a.kdec - a.kdec / 2; break;
    case SWPC_SNP_XZ: i = sa * a.idec - a.idec / 2; j = a.j0_xz; k = sb * a.kdec - a.kdec / 2; break;
    case SWPC_SNP_FS:
        i = sa * a.idec - a.idec / 2; j = sb * a.jdec - a.jdec / 2; k = a.kfs[g2idx(g, i, j)] + 1; break;
    default:
        i = sa * a.idec - a.idec / 2; j = sb * a.jdec - a.jdec / 2; k = a.kob[g2idx(g, i, j)] + 1; break;
    }
    const long long q = gidx(g, i, j, k);
    const long long P = g.plane;
    const int Y = g.pitch;
    float *fr = (out && f < a.nframe) ? T.frame + (long long)f * npt * (T.q == SWPC_SNP_PS ? 4 : 3) : nullptr;

    if (T.q == SWPC_SNP_PS) {
        // m_snap.f90:1016-1036
        float div = (a.Vx[q] - a.Vx[q - 1]) * a.r20x + (a.Vy[q] - a.Vy[q - Y]) * a.r20y
                  + (a.Vz[q] - a.Vz[q - P]) * a.r20z;
        float rot_x = (a.Vz[q + Y] - a.Vz[q]) * a.r20y - (a.Vy[q + P] - a.Vy[q]) * a.r20z;
        float rot_y = (a.Vx[q + P] - a.Vx[q]) * a.r20z - (a.Vz[q + 1] - a.Vz[q]) * a.r20x;
        float rot_z = (a.Vy[q + 1] - a.Vy[q]) * a.r20x - (a.Vx[q + Y] - a.Vx[q]) * a.r20y;
        const float lam = a.lam[q];
        div = div * lam / fabsf(lam + EPSL);
        const float syz = a.Syz[q], sxz = a.Sxz[q], sxy = a.Sxy[q];
        if (T.plane == SWPC_SNP_YZ) {
            // the yz section masks with abs(S + eps) (m_snap.f90:1028-1030)
            rot_x = rot_x * fabsf(syz) / fabsf(syz + EPSL);
            rot_y = rot_y * fabsf(sxz) / fabsf(sxz + EPSL);
            rot_z = rot_z * fabsf(sxy) / fabsf(sxy + EPSL);
        } else {
            // xz, xy, fs and ob mask with abs(S) + eps (m_snap.f90:1121-1123)
            rot_x = rot_x * fabsf(syz) / (fabsf(syz) + EPSL);
            rot_y = rot_y * fabsf(sxz) / (fabsf(sxz) + EPSL);
            rot_z = rot_z * fabsf(sxy) / (fabsf(sxy) + EPSL);
        }
        fr[p] = div * a.UC * a.M0 * 1e-3f;
        fr[p + npt] = rot_x * a.UC * a.M0 * 1e-3f;
        fr[p + 2 * npt] = rot_y * a.UC * a.M0 * 1e-3f;
        fr[p + 3 * npt] = rot_z * a.UC * a.M0 * 1e-3f;
    } else if (T.q == SWPC_SNP_V) {
        // m_snap.f90:1678-1683, maps :1697-1707
        const float b1 = a.Vx[q] * a.UC * a.M0;
        const float b2 = a.Vy[q] * a.UC * a.M0;
        const float b3 = a.Vz[q] * a.UC * a.M0;
        if (T.maxmap) {
            float *mx = T.maxmap;
            const float h = sqrt_rn(b1 * b1 + b2 * b2);
            const float az = fabsf(b3);
            const float m1 = fmaxf(mx[p], az), m2 = fmaxf(mx[p + npt], h);
            mx[p] = m1;
            mx[p + npt] = m2;
            mx[p + 2 * npt] = sqrt_rn(m1 * m1 + m2 * m2);
        }
        if (fr) {
            fr[p] = b1;
            fr[p + npt] = b2;
            fr[p + 2 * npt] = b3;
        }
    } else {
        // m_snap.f90:1857-1859 (u accumulates every step)
        float *u = T.uacc;
        const float u1 = u[p] + a.Vx[q] * a.UC * a.M0 * a.dt;
        const float u2 = u[p + npt] + a.Vy[q] * a.UC * a.M0 * a.dt;
        const float u3 = u[p + 2 * npt] + a.Vz[q] * a.UC * a.M0 * a.dt;
        u[p] = u1;
        u[p + npt] = u2;
        u[p + 2 * npt] = u3;
        if (T.maxmap) {
            float *mx = T.maxmap;
            const float h = sqrt_rn(u1 * u1 + u2 * u2);
            const float az = fabsf(u3);
            const float m1 = fmaxf(mx[p], az), m2 = fmaxf(mx[p + npt], h);
            mx[p] = m1;
            mx[p + npt] = m2;
            mx[p + 2 * npt] = sqrt_rn(m1 * m1 + m2 * m2);
        }
        if (fr) {
            fr[p] = u1;
            fr[p + npt] = u2;
            fr[p + 2 * npt] = u3;
        }
    }
}

__global__ void __launch_bounds__(256) k_snap(SnapArgs a) { snap_body(a, blockIdx.x, a.d_it ? *a.d_it : a.it); }

// batched: context m = blockIdx.y, its arguments in HBM; blocks beyond the
// context's own count idle
__global__ void __launch_bounds__(256) k_snap_b(const SnapArgs *__restrict__ args, int it, const int *d_it)
{
    const SnapArgs &a = args[blockIdx.y];
    if ((int)blockIdx.x >= a.nblk) return;
    snap_body(a, blockIdx.x, d_it ? *d_it : it);
}

static void snap_fill(swpc_ctx *c, SnapArgs &a, int it, const int *d_it)
{
    a.g = c->g;
    a.Vx = c->fields[SWPC_VX]; a.Vy = c->fields[SWPC_VY]; a.Vz = c->fields[SWPC_VZ];
    a.Syz = c->fields[SWPC_SYZ]; a.Sxz = c->fields[SWPC_SXZ]; a.Sxy = c->fields[SWPC_SXY];
    a.lam = c->lam;
    a.kfs = c->d_kfs; a.kob = c->kob;
    for (int t = 0; t < c->snap_ntask; ++t) a.task[t] = c->snap_task[t];
    a.ntask = c->snap_ntask;
    a.nblk = c->snap_nblk;
    const swpc_snap_desc &d = c->snap;
    a.k0_xy = d.k0_xy; a.i0_yz = d.i0_yz; a.j0_xz = d.j0_xz;
    a.idec = d.idec; a.jdec = d.jdec; a.kdec = d.kdec; a.ntdec_s = d.ntdec_s; a.nframe = d.nframe;
    a.r20x = 1.0f / (float)c->desc.dx;
    a.r20y = 1.0f / (float)c->desc.dy;
    a.r20z = 1.0f / (float)c->desc.dz;
    a.M0 = d.M0; a.UC = d.UC; a.dt = c->desc.dt;
    a.it = it;
    a.d_it = d_it;
}

hipError_t launch_snap(swpc_ctx *c, int it, const int *d_it, hipStream_t s)
{
    if (c->snap_ntask <= 0) return hipSuccess;
    SnapArgs a;
    snap_fill(c, a, it, d_it);
    hipLaunchKernelGGL(k_snap, dim3(c->snap_nblk), dim3(256), 0, (s ? s : c->stream), a);
    return hipGetLastError();
}

// batched sections (swpc_batch): one launch for every member with sections,
// as many blocks as the member with the most section points
size_t batch_snap_size() { return sizeof(SnapArgs); }

int batch_snap_fill(swpc_ctx *c, void *out)
{
    if (c->snap_ntask <= 0) return 0;
    SnapArgs a;
    snap_fill(c, a, 0, nullptr);
    memcpy(out, &a, sizeof(a));
    return c->snap_nblk;
}

hipError_t launch_batch_snap(const void *d_args, int nmem, int maxblk, int it, const int *d_it, hipStream_t s)
{
    if (nmem <= 0 || maxblk <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_snap_b, dim3(maxblk, nmem), dim3(256), 0, s, (const SnapArgs *)d_args, it, d_it);
    return hipGetLastError();
}
