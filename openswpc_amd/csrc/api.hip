// api.hip -- C ABI of libswpc_gpu (declared in include/swpc_gpu.h).
#include <stdarg.h>
#include <stdlib.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <utility>
#include <vector>

#include "swpc_internal.h"

static int fail(swpc_ctx *c, int code, const char *fmt, ...)
{
    if (c) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(c->err, sizeof(c->err), fmt, ap);
        va_end(ap);
    }
    return code;
}

#define HIPCHK(c, call)                                                                          \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) return fail((c), SWPC_EHIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

static long long round_up(long long v, long long m) { return (v + m - 1) / m * m; }

static hipError_t dalloc(swpc_ctx *c, void **p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) return hipSuccess;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return e;
    c->bytes += (long long)bytes;
    return hipMemsetAsync(*p, 0, bytes, c->stream);
}

// a per-cell array: from the striped arena when the device has one (each
// segment in the HBM region the caller names), else hipMalloc
static hipError_t dalloc_striped(swpc_ctx *c, void **p, size_t bytes, const int *seg_region, int nseg, size_t seg_bytes)
{
    *p = arena_alloc(c->device, c, bytes, seg_region, nseg, seg_bytes);
    if (!*p) return dalloc(c, p, bytes);
    c->bytes += (long long)bytes;
    return hipMemsetAsync(*p, 0, bytes, c->stream);
}

static void dfree(void *p)
{
    if (p && !arena_free(p)) (void)hipFree(p);
}

// HBM regions of the per-cell arrays when the device has a striped arena:
// the stress update's 18 streams (V, S, lam, mu, taup, taus and the
// memory-variable packs) fall 6/6/6 on the three regions, the velocity
// update's 10 (V, S, rho) 4/3/3; PML auxiliary component c in region c mod 3
// (the nine of each phase 3/3/3)
static const int kFieldRegion[SWPC_NFIELD] = {0, 1, 2, 0, 1, 2, 0, 1, 2};  // Vx..Vz, Sxx..Szz, Syz, Sxz, Sxy
static const int kMedRegion[5] = {0, 0, 1, 2, 2};                           // rho, lam, mu, taup, taus
static const int kPackRegion[SWPC_MAXNM * 6 / 4 + 1] = {0, 1, 2, 0, 1, 2, 0, 1, 2, 0, 1, 2, 0};
static const int kAuxRegion[A_N] = {0, 1, 2, 0, 1, 2, 0, 1, 2, 0, 1, 2, 0, 1, 2, 0, 1, 2};
static const int64_t kGranuleBytes = (int64_t)256 << 20;                    // the arena's granule
static const int64_t kAutoStripeMin = (int64_t)4 << 30;                     // contexts this large stripe by themselves
static const int64_t kStripeSlack = 2 * kGranuleBytes;

// the bytes a context's arena arrays take in their busiest region: arrays of
// a granule or more in whole granules (a segmented array's granules in the
// regions of their segments), smaller ones in 64 KiB steps
static int64_t busiest_region_bytes(const swpc_desc *d, const Grid &g)
{
    int64_t per[3] = {0, 0, 0};
    auto add = [&](int64_t bytes, const int *reg, int nseg, int64_t seg) {
        if (bytes < kGranuleBytes) {
            per[reg[0] % 3] += (bytes + 65535) / 65536 * 65536;
            return;
        }
        const int64_t ng = (bytes + kGranuleBytes - 1) / kGranuleBytes;
        for (int64_t k = 0; k < ng; ++k) {
            const int64_t sg = seg > 0 ? std::min<int64_t>(k * kGranuleBytes / seg, nseg - 1) : 0;
            per[reg[sg] % 3] += kGranuleBytes;
        }
    };
    const int64_t fb = g.n * (int64_t)sizeof(float);
    for (int q = 0; q < SWPC_NFIELD; ++q) add(fb, &kFieldRegion[q], 1, 0);
    for (int q = 0; q < 5; ++q) add(fb, &kMedRegion[q], 1, 0);
    if (d->nm > 0) add((int64_t)6 * d->nm * g.rn * 4, kPackRegion, (6 * d->nm + 3) / 4, g.rn * 16);
    if (d->abc == SWPC_ABC_PML && d->na > 0) {
        const int64_t na = d->na, nx = d->nx, ny = d->ny, nz = d->nz;
        const int64_t npml = 2 * nx * na * nz + 2 * na * (ny - 2 * na) * nz + (int64_t)g.rpitch * (ny - 2 * na) * na;
        add((int64_t)A_N * npml * 4, kAuxRegion, A_N, npml * 4);
    }
    return std::max(per[0], std::max(per[1], per[2]));
}

static void reset_graph(swpc_ctx *c)
{
    ++c->gen;  // launch arguments may change: captured graphs (own or a batch's) are stale
    if (c->graph_exec) {
        (void)hipGraphExecDestroy(c->graph_exec);
        c->graph_exec = nullptr;
    }
    c->graph_block = 0;
}

// Tiles (64 x by columns, one k chunk) that hold PML cells: side columns
// over the full depth, interior columns over the bottom na planes.
static int build_pml_tiles(swpc_ctx *c)
{
    if (c->d_pml_tiles) c->bytes -= (long long)c->npml_tiles * 4 * (long long)sizeof(int);
    dfree(c->d_pml_tiles);
    c->d_pml_tiles = nullptr;
    c->npml_tiles = 0;
    if (c->desc.abc != SWPC_ABC_PML || c->npml == 0) return SWPC_OK;
    const int nx = c->g.nx, ny = c->g.ny, nz = c->g.nz, na = c->g.na, by = c->by, kc = c->kchunk;
    std::vector<int> t;
    // tile = (i0 | width << 20 | xslab << 28, j0, k0, k1)
    auto add = [&](int i0, int w, int xslab, int j0, int k0, int k1) {
        t.push_back(i0 | (w << 20) | (xslab << 28)); t.push_back(j0); t.push_back(k0); t.push_back(k1);
    };
    // x slabs (side columns of interior rows), width na, several rows per block
    const int slab_rows = (64 * by) / std::max(na, 1);
    const bool slab_tiles = na > 0 && na <= 64 && slab_rows > 0;
    for (int k0 = 1; k0 <= nz; k0 += kc) {
        const int k1 = std::min(nz, k0 + kc - 1);
        // y-slab rows as full-row tiles over the full depth; the bottom band
        // of interior columns is marched by the interior kernels
        for (int j0 = 1; j0 <= ny; j0 += by) {
            const int jl = std::min(j0 + by - 1, ny);
            if (j0 <= na || jl > ny - na)
                for (int i0 = 1; i0 <= nx; i0 += 64) add(i0, 64, 0, j0, k0, k1);
        }
        if (slab_tiles)
            for (int j0 = na + 1; j0 <= ny - na; j0 += slab_rows) {
                add(1, na, 1, j0, k0, k1);
                add(nx - na + 1, na, 1, j0, k0, k1);
            }
    }
    if (!slab_tiles && na > 0) return fail(c, SWPC_EINVAL, "PML thickness na must be 1..64");
    if (t.empty()) return SWPC_OK;
    if (dalloc(c, (void **)&c->d_pml_tiles, t.size() * sizeof(int)) != hipSuccess)
        return fail(c, SWPC_ENOMEM, "PML tile list allocation failed");
    HIPCHK(c, hipMemcpyAsync(c->d_pml_tiles, t.data(), t.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->npml_tiles = (int)(t.size() / 4);
    return SWPC_OK;
}

// k chunks of the interior-column kernels.  Blocks of one chunk form a
// contiguous run of logical tiles, so each XCD takes whole chunks; a plane
// of the bottom PML band costs a different time than an interior plane (more
// in the velocity phase, about the same in the stress phase), so chunk
// boundaries equalise estimated time per chunk rather than planes.
static int build_zchunks(swpc_ctx *c)
{
    const Grid &g = c->g;
    const int nz = g.nz, nm = g.nm;
    const int nbz = std::max(1, std::min(nz, (nz + c->kchunk - 1) / c->kchunk));
    const bool pml = c->desc.abc == SWPC_ABC_PML && c->npml > 0;
    // relative time per cell: algorithmic bytes over the streaming rate the
    // two loops reach (interior ~4.7 TB/s, PML band ~3.0 TB/s on MI355X)
    const double w_int[2] = {4.0 * (3 + 4 + 12 + 12 * nm), 4.0 * (6 + 1 + 6)};   // stress, velocity
    const double w_pml[2] = {4.0 * (3 + 2 + 12 + 18) * 1.55, 4.0 * (6 + 1 + 6 + 18) * 1.55};
    std::vector<int> zc(2 * (nbz + 1));
    for (int ph = 0; ph < 2; ++ph) {
        std::vector<double> W(nz + 1, 0.0);
        for (int k = 1; k <= nz; ++k) W[k] = W[k - 1] + ((pml && k > g.ke) ? w_pml[ph] : w_int[ph]);
        int *z = zc.data() + ph * (nbz + 1);
        z[0] = 1;
        int k = 0;
        for (int b = 1; b < nbz; ++b) {
            const double target = W[nz] * b / nbz;
            while (k < nz && W[k] < target) ++k;
            // chunk b starts after plane k; keep every chunk at least one plane deep
            z[b] = std::min(std::max(k + 1, z[b - 1] + 1), nz - (nbz - b) + 1);
        }
        z[nbz] = nz + 1;
        c->zband[ph] = 0;
        for (int b = 0; b < nbz; ++b)
            if (pml && z[b + 1] - 1 > g.ke) {
                c->zband[ph] = b;
                break;
            }
    }
    // fused launch: chunks of kchunk planes over the interior planes kb..ke
    // (each chunk recomputes the velocities of two planes on either side)
    const int nki = g.ke - g.kb + 1;
    const int nfz = nki > 0 ? (nki + c->kchunk - 1) / c->kchunk : 0;
    std::vector<int> fz(nfz + 1);
    for (int b = 0; b <= nfz; ++b) fz[b] = g.kb + (int)((long long)nki * b / std::max(nfz, 1));
    if (c->d_zchunks) c->bytes -= 2LL * (c->nbz + 1) * (long long)sizeof(int);
    dfree(c->d_zchunks);
    c->d_zchunks = nullptr;
    c->nbz = nbz;
    if (dalloc(c, (void **)&c->d_zchunks, zc.size() * sizeof(int)) != hipSuccess)
        return fail(c, SWPC_ENOMEM, "k-chunk table allocation failed");
    HIPCHK(c, hipMemcpyAsync(c->d_zchunks, zc.data(), zc.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    if (c->d_fzc) c->bytes -= (c->nfz + 1) * (long long)sizeof(int);
    dfree(c->d_fzc);
    c->d_fzc = nullptr;
    c->nfz = nfz;
    if (dalloc(c, (void **)&c->d_fzc, fz.size() * sizeof(int)) != hipSuccess)
        return fail(c, SWPC_ENOMEM, "k-chunk table allocation failed");
    HIPCHK(c, hipMemcpyAsync(c->d_fzc, fz.data(), fz.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SWPC_OK;
}

int ctx_fail(swpc_ctx *c, int code, const char *msg) { return fail(c, code, "%s", msg); }

void ctx_reset_graph(swpc_ctx *c) { reset_graph(c); }

extern "C" const char *swpc_last_error(const swpc_ctx *ctx) { return ctx ? ctx->err : "null context"; }

extern "C" int64_t swpc_device_bytes(const swpc_ctx *ctx) { return ctx ? ctx->bytes : 0; }

extern "C" void *swpc_stream(swpc_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

extern "C" void *swpc_staging_buffer(swpc_ctx *ctx, int64_t *nfloats)
{
    if (nfloats) *nfloats = ctx ? ctx->staging_n : 0;
    return ctx ? (void *)ctx->staging : nullptr;
}

// The default k chunk of a context.  16 planes was fastest on 1024x1024x512
// with NM = 3 among 8..128 (the stress and side-column launches favour short
// chunks); an elastic update streams far fewer arrays, so each chunk's
// priming loads weigh more and 32 planes is faster there (256x256x192,
// NM = 0).  Small grids halve it until the interior launch has two blocks per
// CU: 128^3 with a 20-cell PML took 0.233 ms/step in 352 blocks of 16 planes
// and 0.161 in 704 blocks of 8.
static int default_kchunk(const swpc_ctx *c)
{
    const Grid &g = c->g;
    int kc = g.nm > 0 ? 16 : 32;
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || ncu <= 0) {
        (void)hipGetLastError();
        ncu = 256;
    }
    const long long tiles =
        g.nxk > 0 && g.nyk > 0 ? (long long)((g.ie - 1) / 64 - (g.ib - 1) / 64 + 1) * ((g.nyk + 3) / 4) : 0;
    while (tiles > 0 && kc > 4 && tiles * ((g.nz + kc - 1) / kc) < 2LL * ncu) kc /= 2;
    return kc;
}

extern "C" int swpc_create(const swpc_desc *d, swpc_ctx **out)
{
    if (!d || !out) return SWPC_EINVAL;
    *out = nullptr;
    if (d->nx < 4 || d->ny < 4 || d->nz < 4 || d->na < 0 || d->nm < 0 || d->nm > SWPC_MAXNM)
        return SWPC_EINVAL;
    if (!(d->dx > 0) || !(d->dy > 0) || !(d->dz > 0) || !(d->dt > 0)) return SWPC_EINVAL;
    if (d->abc != SWPC_ABC_PML && d->abc != SWPC_ABC_CERJAN) return SWPC_EINVAL;
    if (d->abc == SWPC_ABC_PML && (d->nx <= 2 * d->na || d->ny <= 2 * d->na || d->nz <= d->na)) return SWPC_EINVAL;

    swpc_ctx *c = new swpc_ctx();
    memset(c, 0, sizeof(*c));
    c->desc = *d;
    c->device = d->device;
    c->by = 4;
    c->variant = 1 | 8;
    if (hipSetDevice(d->device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
        hipStreamCreate(&c->stream2) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess) {
        swpc_destroy(c);
        return SWPC_EHIP;
    }
    Grid &g = c->g;
    g.nx = d->nx; g.ny = d->ny; g.nz = d->nz; g.na = d->na; g.nm = d->nm;
    // i = 1 on a 128-byte boundary: every 64-cell tile row of the full-width
    // kernels starts a cache line; 32 halo cells below i = 1 (3 are used)
    g.ox = 32;
    g.pitch = (int)round_up(g.ox + d->nx + 3, 32);
    g.nyp = d->ny + 6;
    g.nzp = d->nz + 6;
    g.plane = (long long)g.pitch * g.nyp;
    g.n = g.plane * g.nzp;
    if (d->abc == SWPC_ABC_PML) {
        g.ib = d->na + 1; g.ie = d->nx - d->na;
        g.jb = d->na + 1; g.je = d->ny - d->na;
        g.kb = 1; g.ke = d->nz - d->na;
    } else {
        g.ib = 1; g.ie = d->nx; g.jb = 1; g.je = d->ny; g.kb = 1; g.ke = d->nz;
    }
    g.nxk = g.ie - g.ib + 1;
    g.nyk = g.je - g.jb + 1;
    c->kchunk = default_kchunk(c);
    g.nzk = g.ke - g.kb + 1;
    // memory variables keep full-width rows (i = 1 at offset 0) so that they
    // align with the field rows; only interior cells are ever touched
    g.rpitch = (int)round_up(d->nx, 32);
    g.rplane = (long long)g.rpitch * g.nyk;
    g.rn = g.rplane * g.nzk;

    int rc = SWPC_OK;
    const size_t fb = (size_t)g.n * sizeof(float);
    const size_t mb = (size_t)g.plane * sizeof(int);
#define ALLOC(ptr, bytes)                                                                   \
    do {                                                                                    \
        if (dalloc(c, (void **)&(ptr), (bytes)) != hipSuccess) { rc = SWPC_ENOMEM; goto bad; } \
    } while (0)
#define ALLOC_STRIPED(ptr, bytes, reg, nseg, segb)                                                  \
    do {                                                                                            \
        if (dalloc_striped(c, (void **)&(ptr), (bytes), (reg), (nseg), (segb)) != hipSuccess) {     \
            rc = SWPC_ENOMEM;                                                                       \
            goto bad;                                                                               \
        }                                                                                           \
    } while (0)
    // a large context on a device without an arena reserves one of its own
    // (SWPC_STRIPE=0 in the environment turns this off)
    {
        const char *e = getenv("SWPC_STRIPE");
        const int64_t busiest = busiest_region_bytes(d, g);
        if (!(e && e[0] == '0') && !arena_active(c->device) && busiest >= kAutoStripeMin / 3 &&
            arena_reserve(c->device, busiest + kStripeSlack, 3, c) == SWPC_OK)
            c->own_arena = 1;
    }
    for (int q = 0; q < SWPC_NFIELD; ++q) ALLOC_STRIPED(c->fields[q], fb, &kFieldRegion[q], 1, 0);
    ALLOC_STRIPED(c->rho, fb, &kMedRegion[0], 1, 0);
    ALLOC_STRIPED(c->lam, fb, &kMedRegion[1], 1, 0);
    ALLOC_STRIPED(c->mu, fb, &kMedRegion[2], 1, 0);
    ALLOC_STRIPED(c->taup, fb, &kMedRegion[3], 1, 0);
    ALLOC_STRIPED(c->taus, fb, &kMedRegion[4], 1, 0);
    ALLOC(c->kfs_top, mb);
    ALLOC(c->kfs_bot, mb);
    ALLOC(c->kob_top, mb);
    ALLOC(c->kob_bot, mb);
    ALLOC(c->kob, mb);
    if (d->nm > 0) {
        // memory variables: float4 packs of rn cells each, then a float2 pack
        // (odd 6*nm/4); pack s in region s mod 3
        const int npack = (6 * d->nm + 3) / 4;
        ALLOC_STRIPED(c->R, (size_t)6 * d->nm * g.rn * sizeof(float), kPackRegion, npack, (size_t)g.rn * 16);
    }
    if (d->abc == SWPC_ABC_PML && d->na > 0) {
        const int na = d->na, nx = d->nx, ny = d->ny, nz = d->nz;
        PmlBox b[5] = {
            {1, 1, 1, nx, na, nz, 0},
            {1, ny - na + 1, 1, nx, na, nz, 0},
            {1, na + 1, 1, na, ny - 2 * na, nz, 0},
            {nx - na + 1, na + 1, 1, na, ny - 2 * na, nz, 0},
            // bottom band under the interior columns: rows padded to the
            // full aligned width (i from 1) so a wave's auxiliaries start
            // on a 128-byte line like its field rows
            {1, na + 1, nz - na + 1, g.rpitch, ny - 2 * na, na, 0},
        };
        long long off = 0;
        for (int q = 0; q < 5; ++q) {
            b[q].off = off;
            off += (long long)b[q].nxb * b[q].nyb * b[q].nzb;
            c->box[q] = b[q];
        }
        c->npml = off;
        ALLOC_STRIPED(c->ade, (size_t)A_N * c->npml * sizeof(float), kAuxRegion, A_N, (size_t)c->npml * sizeof(float));
        ALLOC(c->gxc, (size_t)nx * sizeof(float4));
        ALLOC(c->gxe, (size_t)nx * sizeof(float4));
        ALLOC(c->gyc, (size_t)ny * sizeof(float4));
        ALLOC(c->gye, (size_t)ny * sizeof(float4));
        ALLOC(c->gzc, (size_t)nz * sizeof(float4));
        ALLOC(c->gze, (size_t)nz * sizeof(float4));
    }
    if (d->abc == SWPC_ABC_CERJAN) {
        ALLOC(c->cgx_c, (size_t)(d->nx + 6) * sizeof(float));
        ALLOC(c->cgx_b, (size_t)(d->nx + 6) * sizeof(float));
        ALLOC(c->cgy_c, (size_t)(d->ny + 6) * sizeof(float));
        ALLOC(c->cgy_b, (size_t)(d->ny + 6) * sizeof(float));
        ALLOC(c->cgz_c, (size_t)(d->nz + 6) * sizeof(float));
        ALLOC(c->cgz_b, (size_t)(d->nz + 6) * sizeof(float));
    }
    c->staging_n = (long long)(d->nx + 6) * (d->ny + 6) * (d->nz + 6);
    ALLOC(c->staging, (size_t)c->staging_n * sizeof(float));
    ALLOC(c->d_vmax, 4 * sizeof(float));
    ALLOC(c->d_it, 4 * sizeof(int));
#undef ALLOC
#undef ALLOC_STRIPED
    {
        // 4th-order coefficients of both band states (m_kernel.f90:43-54,106-111)
        const float dx = (float)d->dx, dy = (float)d->dy, dz = (float)d->dz;
        const float h[3] = {dx, dy, dz};
        float *re40[3] = {c->coef.re40x, c->coef.re40y, c->coef.re40z};
        float *re41[3] = {c->coef.re41x, c->coef.re41y, c->coef.re41z};
        for (int a = 0; a < 3; ++a) {
            const float rc40 = 17.0f / 16.0f / h[a], rc41 = 1.0f / 48.0f / h[a];
            const float rd40 = -1.0f / 16.0f / h[a], rd41 = -1.0f / 48.0f / h[a];
            re40[a][0] = rc40 + (-1.0f) * rd40;
            re41[a][0] = rc41 + (-1.0f) * rd41;
            re40[a][1] = rc40 + (1.0f) * rd40;
            re41[a][1] = rc41 + (1.0f) * rd41;
        }
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess) { rc = SWPC_EHIP; goto bad; }
    if ((rc = build_pml_tiles(c)) != SWPC_OK) goto bad;
    if ((rc = build_zchunks(c)) != SWPC_OK) goto bad;
    if ((rc = build_schedule(c)) != SWPC_OK) goto bad;
    *out = c;
    return SWPC_OK;
bad:
    swpc_destroy(c);
    return rc;
}

static void free_snapshots(swpc_ctx *c)
{
    for (int s = 0; s < SWPC_SNP_NSECTION; ++s) {
        dfree(c->d_snap_frame[s]);
        dfree(c->d_snap_max[s]);
        c->d_snap_frame[s] = c->d_snap_max[s] = nullptr;
    }
    for (int p = 0; p < 5; ++p) {
        dfree(c->d_snap_u[p]);
        c->d_snap_u[p] = nullptr;
    }
    dfree(c->d_kfs);
    c->d_kfs = nullptr;
    c->snap_ntask = 0;
    c->snap_nblk = 0;
}

extern "C" int swpc_destroy(swpc_ctx *c)
{
    if (!c) return SWPC_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    reset_graph(c);
    swpc_adj_release(c);
    for (int q = 0; q < SWPC_NFIELD; ++q) dfree(c->fields[q]);
    for (int q = 0; q < SWPC_NFIELD; ++q) dfree(c->fields2[q]);
    dfree(c->d_fzc);
    dfree(c->rho); dfree(c->lam); dfree(c->mu); dfree(c->taup); dfree(c->taus);
    dfree(c->kfs_top); dfree(c->kfs_bot); dfree(c->kob_top); dfree(c->kob_bot); dfree(c->kob);
    dfree(c->R); dfree(c->ade); dfree(c->d_pml_tiles); dfree(c->d_zchunks);
    dfree(c->d_sched[0]); dfree(c->d_sched[1]);
    dfree(c->gxc); dfree(c->gxe); dfree(c->gyc); dfree(c->gye); dfree(c->gzc); dfree(c->gze);
    dfree(c->cgx_c); dfree(c->cgx_b); dfree(c->cgy_c); dfree(c->cgy_b); dfree(c->cgz_c); dfree(c->cgz_b);
    dfree(c->d_targets); dfree(c->d_contrib); dfree(c->d_comp); dfree(c->d_mo); dfree(c->d_stf);
    dfree(c->d_ist); dfree(c->d_jst); dfree(c->d_kst); dfree(c->d_uacc); dfree(c->d_eacc);
    for (int q = 0; q < 4; ++q) dfree(c->d_wav[q]);
    dfree(c->d_green_ijk); dfree(c->d_green_acc); dfree(c->d_gf); dfree(c->d_green_stf);
    dfree(c->staging); dfree(c->d_vmax); dfree(c->d_it);
    free_snapshots(c);
    if (c->stream2) (void)hipStreamSynchronize(c->stream2);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    // the arena this context reserved for itself goes with it (it stays, to
    // be released at process exit or by swpc_stripe_release, while another
    // context still holds memory from it)
    if (c->own_arena) (void)swpc_stripe_release(c->device);
    delete c;
    return SWPC_OK;
}

// upload one reference-layout (k,i,j) halo-3 array into a device array
static int upload_ref(swpc_ctx *c, const float *host, float *dev)
{
    HIPCHK(c, hipMemcpyAsync(c->staging, host, (size_t)c->staging_n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_ref_to_dev(c, c->staging, dev, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SWPC_OK;
}

static int upload_map(swpc_ctx *c, const int32_t *host, int *dev)
{
    const Grid &g = c->g;
    std::vector<int> buf((size_t)g.plane, 0);
    const int nxr = g.nx + 6;
    for (int j = -2; j <= g.ny + 3; ++j)
        for (int i = -2; i <= g.nx + 3; ++i) buf[(size_t)g2idx(g, i, j)] = host[(size_t)(j + 2) * nxr + (i + 2)];
    HIPCHK(c, hipMemcpyAsync(dev, buf.data(), buf.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SWPC_OK;
}

extern "C" int swpc_set_medium(swpc_ctx *c, const float *rho, const float *lam, const float *mu, const float *taup,
                               const float *taus, const int32_t *kfs_top, const int32_t *kfs_bot,
                               const int32_t *kob_top, const int32_t *kob_bot, const int32_t *kob)
{
    if (!c) return SWPC_EINVAL;
    if (!rho || !lam || !mu || !taup || !taus || !kfs_top || !kfs_bot || !kob_top || !kob_bot || !kob)
        return fail(c, SWPC_EINVAL, "swpc_set_medium: null array");
    HIPCHK(c, hipSetDevice(c->device));
    int rc;
    if ((rc = upload_ref(c, rho, c->rho)) || (rc = upload_ref(c, lam, c->lam)) || (rc = upload_ref(c, mu, c->mu)) ||
        (rc = upload_ref(c, taup, c->taup)) || (rc = upload_ref(c, taus, c->taus)))
        return rc;
    if ((rc = upload_map(c, kfs_top, c->kfs_top)) || (rc = upload_map(c, kfs_bot, c->kfs_bot)) ||
        (rc = upload_map(c, kob_top, c->kob_top)) || (rc = upload_map(c, kob_bot, c->kob_bot)) ||
        (rc = upload_map(c, kob, c->kob)))
        return rc;
    c->has_medium = true;
    reset_graph(c);
    return SWPC_OK;
}

extern "C" int swpc_set_visco(swpc_ctx *c, const float *c1, const float *c2, const float *d1, float d2)
{
    if (!c) return SWPC_EINVAL;
    const int nm = c->desc.nm;
    if (nm > 0 && (!c1 || !c2 || !d1)) return fail(c, SWPC_EINVAL, "swpc_set_visco: null array");
    for (int m = 0; m < nm; ++m) {
        c->visco.c1[m] = c1[m];
        c->visco.c2[m] = c2[m];
        c->visco.d1[m] = d1[m];
    }
    c->visco.d2 = d2;
    c->has_visco = true;
    reset_graph(c);
    return SWPC_OK;
}

extern "C" int swpc_set_pml(swpc_ctx *c, const float *gxc, const float *gxe, const float *gyc, const float *gye,
                            const float *gzc, const float *gze)
{
    if (!c) return SWPC_EINVAL;
    if (c->desc.abc != SWPC_ABC_PML) return fail(c, SWPC_ESTATE, "swpc_set_pml: context uses the Cerjan sponge");
    if (c->npml == 0) { c->has_pml = true; return SWPC_OK; }
    if (!gxc || !gxe || !gyc || !gye || !gzc || !gze) return fail(c, SWPC_EINVAL, "swpc_set_pml: null array");
    HIPCHK(c, hipSetDevice(c->device));
    const Grid &g = c->g;
    // g(4,n) is four consecutive floats per cell: the float4 layout
    HIPCHK(c, hipMemcpyAsync(c->gxc, gxc, (size_t)g.nx * 16, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->gxe, gxe, (size_t)g.nx * 16, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->gyc, gyc, (size_t)g.ny * 16, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->gye, gye, (size_t)g.ny * 16, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->gzc, gzc, (size_t)g.nz * 16, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->gze, gze, (size_t)g.nz * 16, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->has_pml = true;
    reset_graph(c);
    return SWPC_OK;
}

extern "C" int swpc_set_cerjan(swpc_ctx *c, const float *gx_c, const float *gx_b, const float *gy_c,
                               const float *gy_b, const float *gz_c, const float *gz_b)
{
    if (!c) return SWPC_EINVAL;
    if (c->desc.abc != SWPC_ABC_CERJAN) return fail(c, SWPC_ESTATE, "swpc_set_cerjan: context uses the PML");
    if (!gx_c || !gx_b || !gy_c || !gy_b || !gz_c || !gz_b) return fail(c, SWPC_EINVAL, "swpc_set_cerjan: null array");
    HIPCHK(c, hipSetDevice(c->device));
    const Grid &g = c->g;
    HIPCHK(c, hipMemcpyAsync(c->cgx_c, gx_c, (size_t)(g.nx + 6) * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->cgx_b, gx_b, (size_t)(g.nx + 6) * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->cgy_c, gy_c, (size_t)(g.ny + 6) * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->cgy_b, gy_b, (size_t)(g.ny + 6) * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->cgz_c, gz_c, (size_t)(g.nz + 6) * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->cgz_b, gz_b, (size_t)(g.nz + 6) * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->has_cerjan = true;
    reset_graph(c);
    return SWPC_OK;
}

extern "C" int swpc_set_sources(swpc_ctx *c, int nsrc, int bf_mode, const int32_t *isrc, const int32_t *jsrc,
                                const int32_t *ksrc, const float *mo, const float *comp, float M0, int nt,
                                const float *stf)
{
    if (!c) return SWPC_EINVAL;
    if (nsrc < 0 || nt < 0) return fail(c, SWPC_EINVAL, "swpc_set_sources: negative count");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dfree(c->d_targets); dfree(c->d_contrib); dfree(c->d_comp); dfree(c->d_mo); dfree(c->d_stf);
    c->d_targets = nullptr; c->d_contrib = nullptr; c->d_comp = nullptr; c->d_mo = nullptr; c->d_stf = nullptr;
    c->nsrc = 0;
    c->ntarget = 0;
    reset_graph(c);
    if (nsrc == 0) return SWPC_OK;
    if (!isrc || !jsrc || !ksrc || !comp || !stf || (!bf_mode && !mo) || !(M0 != 0.0f))
        return fail(c, SWPC_EINVAL, "swpc_set_sources: null array or zero M0");
    const Grid &g = c->g;
    for (int n = 0; n < nsrc; ++n) {
        if (isrc[n] < 2 || isrc[n] > g.nx || jsrc[n] < 2 || jsrc[n] > g.ny || ksrc[n] < 2 || ksrc[n] > g.nz)
            return fail(c, SWPC_EINVAL, "swpc_set_sources: source %d at (%d,%d,%d) outside [2,nx]x[2,ny]x[2,nz]", n,
                        isrc[n], jsrc[n], ksrc[n]);
    }
    const int ncomp = bf_mode ? 3 : 6;
    // targets grouped by (field, cell) in first-touch order, contributions in source order
    std::map<std::pair<int, long long>, int> where;
    std::vector<std::vector<SrcContrib>> lists;
    std::vector<std::pair<int, long long>> keys;
    auto add = [&](int field, int i, int j, int k, SrcContrib ct) {
        const std::pair<int, long long> key(field, gidx(g, i, j, k));
        auto it = where.find(key);
        int id;
        if (it == where.end()) {
            id = (int)lists.size();
            where[key] = id;
            lists.emplace_back();
            keys.push_back(key);
        } else {
            id = it->second;
        }
        lists[id].push_back(ct);
    };
    for (int n = 0; n < nsrc; ++n) {
        const int ii = isrc[n], jj = jsrc[n], kk = ksrc[n];
        SrcContrib ct;
        memset(&ct, 0, sizeof(ct));
        ct.src = n;
        if (!bf_mode) {  // m_source.f90:808-839
            ct.div4 = 0;
            ct.comp = 0; add(SWPC_SXX, ii, jj, kk, ct);
            ct.comp = 1; add(SWPC_SYY, ii, jj, kk, ct);
            ct.comp = 2; add(SWPC_SZZ, ii, jj, kk, ct);
            ct.div4 = 1;
            ct.comp = 5;
            add(SWPC_SXY, ii, jj, kk, ct); add(SWPC_SXY, ii, jj - 1, kk, ct);
            add(SWPC_SXY, ii - 1, jj, kk, ct); add(SWPC_SXY, ii - 1, jj - 1, kk, ct);
            ct.comp = 4;
            add(SWPC_SXZ, ii, jj, kk, ct); add(SWPC_SXZ, ii, jj, kk - 1, ct);
            add(SWPC_SXZ, ii - 1, jj, kk, ct); add(SWPC_SXZ, ii - 1, jj, kk - 1, ct);
            ct.comp = 3;
            add(SWPC_SYZ, ii, jj, kk, ct); add(SWPC_SYZ, ii, jj, kk - 1, ct);
            add(SWPC_SYZ, ii, jj - 1, kk, ct); add(SWPC_SYZ, ii, jj - 1, kk - 1, ct);
        } else {  // m_source.f90:880-890
            const long long p0 = gidx(g, ii, jj, kk);
            ct.comp = 0;
            ct.rho_a = p0; ct.rho_b = gidx(g, ii + 1, jj, kk); add(SWPC_VX, ii, jj, kk, ct);
            ct.rho_a = p0; ct.rho_b = gidx(g, ii - 1, jj, kk); add(SWPC_VX, ii - 1, jj, kk, ct);
            ct.comp = 1;
            ct.rho_a = p0; ct.rho_b = gidx(g, ii, jj + 1, kk); add(SWPC_VY, ii, jj, kk, ct);
            ct.rho_a = p0; ct.rho_b = gidx(g, ii, jj - 1, kk); add(SWPC_VY, ii, jj - 1, kk, ct);
            ct.comp = 2;
            ct.rho_a = p0; ct.rho_b = gidx(g, ii, jj, kk + 1); add(SWPC_VZ, ii, jj, kk, ct);
            ct.rho_a = p0; ct.rho_b = gidx(g, ii, jj, kk - 1); add(SWPC_VZ, ii, jj, kk - 1, ct);
        }
    }
    std::vector<SrcTarget> tg(lists.size());
    std::vector<SrcContrib> ct;
    for (size_t q = 0; q < lists.size(); ++q) {
        tg[q].field = keys[q].first;
        tg[q].idx = keys[q].second;
        tg[q].start = (int)ct.size();
        tg[q].count = (int)lists[q].size();
        ct.insert(ct.end(), lists[q].begin(), lists[q].end());
    }
    // normalisation by M0 in single precision (m_source.f90:297-303) and
    // dt_dxyz = dt / (dx*dy*dz) (m_source.f90:306)
    std::vector<float> cm((size_t)ncomp * nsrc), mon((size_t)nsrc, 1.0f);
    for (int q = 0; q < ncomp; ++q)
        for (int n = 0; n < nsrc; ++n) {
            const float v = comp[(size_t)q * nsrc + n];
            cm[(size_t)q * nsrc + n] = bf_mode ? v / M0 : v;
        }
    if (!bf_mode)
        for (int n = 0; n < nsrc; ++n) mon[n] = mo[n] / M0;
    const float dx = (float)c->desc.dx, dy = (float)c->desc.dy, dz = (float)c->desc.dz;
    c->dt_dxyz = c->desc.dt / (dx * dy * dz);
    if (dalloc(c, (void **)&c->d_targets, tg.size() * sizeof(SrcTarget)) != hipSuccess ||
        dalloc(c, (void **)&c->d_contrib, ct.size() * sizeof(SrcContrib)) != hipSuccess ||
        dalloc(c, (void **)&c->d_comp, cm.size() * sizeof(float)) != hipSuccess ||
        dalloc(c, (void **)&c->d_mo, mon.size() * sizeof(float)) != hipSuccess ||
        dalloc(c, (void **)&c->d_stf, (size_t)nt * nsrc * sizeof(float) + 4) != hipSuccess)
        return fail(c, SWPC_ENOMEM, "swpc_set_sources: device allocation failed");
    HIPCHK(c, hipMemcpyAsync(c->d_targets, tg.data(), tg.size() * sizeof(SrcTarget), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_contrib, ct.data(), ct.size() * sizeof(SrcContrib), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_comp, cm.data(), cm.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_mo, mon.data(), mon.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    if (nt > 0)
        HIPCHK(c, hipMemcpyAsync(c->d_stf, stf, (size_t)nt * nsrc * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->nsrc = nsrc;
    c->bf_mode = bf_mode ? 1 : 0;
    c->src_nt = nt;
    c->ntarget = (int)tg.size();
    return SWPC_OK;
}

extern "C" int swpc_set_receivers(swpc_ctx *c, int nst, const int32_t *ist, const int32_t *jst, const int32_t *kst,
                                  int what, int ntdec_w, int ntw, float M0, float UC)
{
    if (!c) return SWPC_EINVAL;
    if (nst < 0 || (nst > 0 && (!ist || !jst || !kst || ntdec_w < 1 || ntw < 1)))
        return fail(c, SWPC_EINVAL, "swpc_set_receivers: bad arguments");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dfree(c->d_ist); dfree(c->d_jst); dfree(c->d_kst); dfree(c->d_uacc); dfree(c->d_eacc);
    c->d_ist = c->d_jst = c->d_kst = nullptr;
    c->d_uacc = c->d_eacc = nullptr;
    for (int q = 0; q < 4; ++q) { dfree(c->d_wav[q]); c->d_wav[q] = nullptr; }
    c->nst = 0;
    reset_graph(c);
    if (nst == 0) return SWPC_OK;
    const Grid &g = c->g;
    for (int n = 0; n < nst; ++n)
        if (ist[n] < 1 || ist[n] > g.nx || jst[n] < 1 || jst[n] > g.ny || kst[n] < 1 || kst[n] > g.nz)
            return fail(c, SWPC_EINVAL, "swpc_set_receivers: station %d outside the grid", n);
    if (dalloc(c, (void **)&c->d_ist, nst * sizeof(int)) != hipSuccess ||
        dalloc(c, (void **)&c->d_jst, nst * sizeof(int)) != hipSuccess ||
        dalloc(c, (void **)&c->d_kst, nst * sizeof(int)) != hipSuccess ||
        dalloc(c, (void **)&c->d_uacc, 3 * nst * sizeof(float)) != hipSuccess ||
        dalloc(c, (void **)&c->d_eacc, 6 * nst * sizeof(float)) != hipSuccess)
        return fail(c, SWPC_ENOMEM, "swpc_set_receivers: device allocation failed");
    const int ncmp[4] = {3, 3, 6, 6};
    for (int q = 0; q < 4; ++q)
        if (what & (1 << q))
            if (dalloc(c, (void **)&c->d_wav[q], (size_t)ntw * ncmp[q] * nst * sizeof(float)) != hipSuccess)
                return fail(c, SWPC_ENOMEM, "swpc_set_receivers: device allocation failed");
    HIPCHK(c, hipMemcpyAsync(c->d_ist, ist, nst * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_jst, jst, nst * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_kst, kst, nst * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->nst = nst;
    c->wav_what = what;
    c->ntdec_w = ntdec_w;
    c->ntw = ntw;
    c->wav_M0 = M0;
    c->wav_UC = UC;
    return SWPC_OK;
}

static void free_green(swpc_ctx *c)
{
    dfree(c->d_green_ijk); dfree(c->d_green_acc); dfree(c->d_gf); dfree(c->d_green_stf);
    c->d_green_ijk = nullptr;
    c->d_green_acc = c->d_gf = c->d_green_stf = nullptr;
    c->green_on = 0;
    c->green_ng = c->green_nt = 0;
}

extern "C" int swpc_set_green(swpc_ctx *c, int ng, const int32_t *ig, const int32_t *jg, const int32_t *kg,
                              int isrc, int jsrc, int ksrc, int cmp, int bforce, int ntdec_w, int ntw, int nt,
                              const float *stf)
{
    if (!c) return SWPC_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    free_green(c);
    reset_graph(c);
    if (cmp < 0) return SWPC_OK;
    if (cmp > 2 || ng < 0 || nt < 0 || ntdec_w < 1 || ntw < 1 || (ng > 0 && (!ig || !jg || !kg)) ||
        (nt > 0 && !stf))
        return fail(c, SWPC_EINVAL, "swpc_set_green: bad arguments");
    const Grid &g = c->g;
    if (isrc < 1 || isrc > g.nx || jsrc < 1 || jsrc > g.ny || ksrc < 1 || ksrc > g.nz)
        return fail(c, SWPC_EINVAL, "swpc_set_green: virtual source cell outside the grid");
    for (int n = 0; n < ng; ++n)
        if (ig[n] < 1 || ig[n] > g.nx || jg[n] < 1 || jg[n] > g.ny || kg[n] < 1 || kg[n] > g.nz)
            return fail(c, SWPC_EINVAL, "swpc_set_green: GF point %d outside the grid", n);
    const int ncmp = bforce ? 9 : 6;
    // accumulators and output start at zero (m_green.f90:424-441, :324)
    if (dalloc(c, (void **)&c->d_green_ijk, (size_t)3 * ng * sizeof(int)) != hipSuccess ||
        dalloc(c, (void **)&c->d_green_acc, (size_t)12 * ng * sizeof(float)) != hipSuccess ||
        dalloc(c, (void **)&c->d_gf, (size_t)ntw * ncmp * ng * sizeof(float)) != hipSuccess ||
        dalloc(c, (void **)&c->d_green_stf, (size_t)nt * sizeof(float)) != hipSuccess) {
        free_green(c);
        return fail(c, SWPC_ENOMEM, "swpc_set_green: device allocation failed");
    }
    if (ng > 0) {
        HIPCHK(c, hipMemcpyAsync(c->d_green_ijk, ig, ng * sizeof(int), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->d_green_ijk + ng, jg, ng * sizeof(int), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->d_green_ijk + 2 * (size_t)ng, kg, ng * sizeof(int), hipMemcpyHostToDevice,
                                 c->stream));
    }
    if (nt > 0)
        HIPCHK(c, hipMemcpyAsync(c->d_green_stf, stf, (size_t)nt * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->green_ng = ng;
    c->green_ncmp = ncmp;
    c->green_ntdec = ntdec_w;
    c->green_ntw = ntw;
    c->green_nt = nt;
    c->green_src[0] = isrc; c->green_src[1] = jsrc; c->green_src[2] = ksrc;
    // unit force along green_cmp (m_green.f90:136-143)
    for (int q = 0; q < 3; ++q) c->green_f[q] = q == cmp ? 1.0f : 0.0f;
    // dt_dxyz = real(dt / (dx*dy*dz)) (m_green.f90:145)
    const float dx = (float)c->desc.dx, dy = (float)c->desc.dy, dz = (float)c->desc.dz;
    c->green_dt_dxyz = c->desc.dt / (dx * dy * dz);
    c->green_on = 1;
    return SWPC_OK;
}

extern "C" int swpc_green_store(swpc_ctx *c, int it)
{
    if (!c) return SWPC_EINVAL;
    HIPCHK(c, launch_green_store(c, it, nullptr));
    return SWPC_OK;
}

extern "C" int swpc_green_source(swpc_ctx *c, int it)
{
    if (!c) return SWPC_EINVAL;
    HIPCHK(c, launch_green_source(c, it, nullptr));
    return SWPC_OK;
}

extern "C" int swpc_get_green(swpc_ctx *c, float *out)
{
    if (!c || !out) return SWPC_EINVAL;
    if (!c->green_on) return fail(c, SWPC_ESTATE, "swpc_get_green: Green's-function mode not set");
    const size_t n = (size_t)c->green_ntw * c->green_ncmp * c->green_ng;
    if (n == 0) return SWPC_OK;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(out, c->d_gf, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SWPC_OK;
}

static int ready(swpc_ctx *c)
{
    if (!c->has_medium) return fail(c, SWPC_ESTATE, "medium not set");
    if (c->desc.nm > 0 && !c->has_visco) return fail(c, SWPC_ESTATE, "visco coefficients not set");
    if (c->desc.abc == SWPC_ABC_PML && !c->has_pml) return fail(c, SWPC_ESTATE, "PML profiles not set");
    if (c->desc.abc == SWPC_ABC_CERJAN && !c->has_cerjan) return fail(c, SWPC_ESTATE, "Cerjan profiles not set");
    return SWPC_OK;
}

extern "C" int swpc_set_snapshots(swpc_ctx *c, const swpc_snap_desc *d, const int32_t *kfs)
{
    if (!c) return SWPC_EINVAL;
    if (!d || !kfs) return fail(c, SWPC_EINVAL, "swpc_set_snapshots: null argument");
    const Grid &g = c->g;
    if (d->idec < 1 || d->jdec < 1 || d->kdec < 1 || d->ntdec_s < 1 || d->nframe < 1)
        return fail(c, SWPC_EINVAL, "swpc_set_snapshots: decimations and nframe must be >= 1");
    if (d->k0_xy < 1 || d->k0_xy > g.nz || d->i0_yz < 1 || d->i0_yz > g.nx || d->j0_xz < 1 || d->j0_xz > g.ny)
        return fail(c, SWPC_EINVAL, "swpc_set_snapshots: section plane outside the grid");
    for (int j = 1; j <= g.ny; ++j)
        for (int i = 1; i <= g.nx; ++i) {
            const int k = kfs[(size_t)(j + 2) * (g.nx + 6) + (i + 2)];
            if (k < 0 || k > g.nz) return fail(c, SWPC_EINVAL, "swpc_set_snapshots: kfs outside 0..nz");
        }
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    reset_graph(c);
    free_snapshots(c);
    c->snap = *d;
    const int nxs = (g.nx + d->idec / 2) / d->idec, nys = (g.ny + d->jdec / 2) / d->jdec;
    const int nzs = (g.nz + d->kdec / 2) / d->kdec;
    int nblk = 0, ntask = 0;
    for (int s = 0; s < SWPC_SNP_NSECTION; ++s) {
        const int plane = s / 3, q = s % 3;
        const int ns1 = plane == SWPC_SNP_YZ ? nys : nxs;
        const int ns2 = plane == SWPC_SNP_XY || plane >= SWPC_SNP_FS ? nys : nzs;
        const int ncomp = q == SWPC_SNP_PS ? 4 : 3;
        c->snap_dims[s][0] = ns1;
        c->snap_dims[s][1] = ns2;
        c->snap_dims[s][2] = ncomp;
        if (!d->sw[s]) continue;
        const size_t npt = (size_t)ns1 * ns2;
        SnapTask &T = c->snap_task[ntask];
        T.plane = plane;
        T.q = q;
        T.ns1 = ns1;
        T.ns2 = ns2;
        T.blk0 = nblk;
        T.uacc = T.maxmap = nullptr;
        if (dalloc(c, (void **)&c->d_snap_frame[s], npt * ncomp * (size_t)d->nframe * sizeof(float)) != hipSuccess)
            return fail(c, SWPC_ENOMEM, "snapshot frame store allocation failed (%zu MB)",
                        npt * ncomp * (size_t)d->nframe * 4 >> 20);
        T.frame = c->d_snap_frame[s];
        if (q == SWPC_SNP_U) {
            if (dalloc(c, (void **)&c->d_snap_u[plane], npt * 3 * sizeof(float)) != hipSuccess)
                return fail(c, SWPC_ENOMEM, "snapshot accumulator allocation failed");
            T.uacc = c->d_snap_u[plane];
        }
        if (q != SWPC_SNP_PS && plane >= SWPC_SNP_FS) {
            if (dalloc(c, (void **)&c->d_snap_max[s], npt * 3 * sizeof(float)) != hipSuccess)
                return fail(c, SWPC_ENOMEM, "snapshot maximum map allocation failed");
            T.maxmap = c->d_snap_max[s];
        }
        nblk += (int)((npt + 255) / 256);
        ++ntask;
    }
    if (dalloc(c, (void **)&c->d_kfs, (size_t)g.plane * sizeof(int)) != hipSuccess)
        return fail(c, SWPC_ENOMEM, "kfs map allocation failed");
    int rc = upload_map(c, kfs, c->d_kfs);
    if (rc) return rc;
    c->snap_ntask = ntask;
    c->snap_nblk = nblk;
    return SWPC_OK;
}

// frames that steps it0 .. it0+nsteps-1 would take must fit the frame store
static int check_snap_frames(swpc_ctx *c, int it0, int nsteps)
{
    if (c->snap_ntask <= 0 || nsteps <= 0) return SWPC_OK;
    const int last = it0 + nsteps - 1;
    if (last < 1) return SWPC_OK;
    const int f = (last - 1) / c->snap.ntdec_s;
    if (f >= c->snap.nframe)
        return fail(c, SWPC_EINVAL, "snapshot frame store holds %d frames; step %d would take frame %d",
                    c->snap.nframe, last, f);
    return SWPC_OK;
}

extern "C" int swpc_snap_write(swpc_ctx *c, int it)
{
    if (!c) return SWPC_EINVAL;
    int rc = check_snap_frames(c, it, 1);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, launch_snap(c, it, nullptr));
    return SWPC_OK;
}

extern "C" int swpc_snap_dims(const swpc_ctx *c, int section, int dims[3])
{
    if (!c || !dims || section < 0 || section >= SWPC_SNP_NSECTION) return SWPC_EINVAL;
    for (int q = 0; q < 3; ++q) dims[q] = c->snap_dims[section][q];
    return c->d_snap_frame[section] ? SWPC_OK : SWPC_ESTATE;
}

extern "C" int swpc_get_snapshot(swpc_ctx *c, int section, int frame, float *out)
{
    if (!c || !out || section < 0 || section >= SWPC_SNP_NSECTION) return SWPC_EINVAL;
    if (!c->d_snap_frame[section]) return fail(c, SWPC_ESTATE, "swpc_get_snapshot: section %d not enabled", section);
    if (frame < 0 || frame >= c->snap.nframe) return fail(c, SWPC_EINVAL, "swpc_get_snapshot: frame out of range");
    const size_t n = (size_t)c->snap_dims[section][0] * c->snap_dims[section][1] * c->snap_dims[section][2];
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(out, c->d_snap_frame[section] + n * frame, n * sizeof(float), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SWPC_OK;
}

extern "C" int swpc_get_snap_max(swpc_ctx *c, int section, float *out)
{
    if (!c || !out || section < 0 || section >= SWPC_SNP_NSECTION) return SWPC_EINVAL;
    if (!c->d_snap_max[section])
        return fail(c, SWPC_ESTATE, "swpc_get_snap_max: no maximum map for section %d (fs/ob v or u only)", section);
    const size_t n = (size_t)c->snap_dims[section][0] * c->snap_dims[section][1] * 3;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(out, c->d_snap_max[section], n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SWPC_OK;
}

extern "C" int swpc_kernel_update_stress(swpc_ctx *c)
{
    if (!c) return SWPC_EINVAL;
    int rc = ready(c);
    if (rc) return rc;
    HIPCHK(c, launch_kernel_stress(c));
    return SWPC_OK;
}

extern "C" int swpc_absorb_update_stress(swpc_ctx *c)
{
    if (!c) return SWPC_EINVAL;
    int rc = ready(c);
    if (rc) return rc;
    if (c->desc.abc == SWPC_ABC_PML) HIPCHK(c, launch_pml_stress(c));
    else HIPCHK(c, launch_cerjan_stress(c));
    return SWPC_OK;
}

extern "C" int swpc_kernel_update_vel(swpc_ctx *c)
{
    if (!c) return SWPC_EINVAL;
    int rc = ready(c);
    if (rc) return rc;
    HIPCHK(c, launch_kernel_vel(c));
    return SWPC_OK;
}

extern "C" int swpc_absorb_update_vel(swpc_ctx *c)
{
    if (!c) return SWPC_EINVAL;
    int rc = ready(c);
    if (rc) return rc;
    if (c->desc.abc == SWPC_ABC_PML) HIPCHK(c, launch_pml_vel(c));
    else HIPCHK(c, launch_cerjan_vel(c));
    return SWPC_OK;
}

extern "C" int swpc_update_stress(swpc_ctx *c)
{
    if (!c) return SWPC_EINVAL;
    int rc = ready(c);
    if (rc) return rc;
    if (c->desc.abc == SWPC_ABC_PML) {
        HIPCHK(c, launch_stress(c, 3, nullptr));
    } else {
        HIPCHK(c, launch_stress(c, 1, nullptr));
        HIPCHK(c, launch_cerjan_stress(c));
    }
    return SWPC_OK;
}

extern "C" int swpc_update_vel(swpc_ctx *c)
{
    if (!c) return SWPC_EINVAL;
    int rc = ready(c);
    if (rc) return rc;
    if (c->desc.abc == SWPC_ABC_PML) {
        HIPCHK(c, launch_vel(c, 3, nullptr));
    } else {
        HIPCHK(c, launch_vel(c, 1, nullptr));
        HIPCHK(c, launch_cerjan_vel(c));
    }
    return SWPC_OK;
}

extern "C" int swpc_source_stressglut(swpc_ctx *c, int it)
{
    if (!c) return SWPC_EINVAL;
    HIPCHK(c, launch_stressglut(c, it, nullptr));
    return SWPC_OK;
}

extern "C" int swpc_source_bodyforce(swpc_ctx *c, int it)
{
    if (!c) return SWPC_EINVAL;
    HIPCHK(c, launch_bodyforce(c, it, nullptr));
    return SWPC_OK;
}

extern "C" int swpc_wav_store(swpc_ctx *c, int it)
{
    if (!c) return SWPC_EINVAL;
    HIPCHK(c, launch_wav_store(c, it, nullptr));
    return SWPC_OK;
}

// one step in the order of src/swpc_3d/main.f90:122-138.  d_it != nullptr:
// kernels read the step number from device memory and the step ends by
// advancing it (graph-capturable).  marks != nullptr: an event is recorded
// before the step and after each of the SWPC_NTIMER launch slots (timing;
// no second stream then).
static hipError_t enqueue_step(swpc_ctx *c, int it, const int *d_it, hipEvent_t *marks = nullptr)
{
    // kernel__update_stress and absorb__update_stress touch disjoint cells
    // and only read V: with the PML the interior columns (interior planes and
    // bottom band) are one launch and the side columns another; the Cerjan
    // taper is a separate sweep over the updated stress, as in the reference.
    hipError_t e;
    const bool pml = c->desc.abc == SWPC_ABC_PML;
    const bool fork = !marks && pml && (c->variant & 2) && c->npml_tiles > 0;
    hipStream_t side = fork ? c->stream2 : c->stream;
    int slot = 0;
    auto mark = [&]() -> hipError_t { return marks ? hipEventRecord(marks[slot++], c->stream) : hipSuccess; };
    auto fork_to_side = [&]() -> hipError_t {
        if (!fork) return hipSuccess;
        hipError_t r = hipEventRecord(c->ev_fork, c->stream);
        return r != hipSuccess ? r : hipStreamWaitEvent(c->stream2, c->ev_fork, 0);
    };
    auto join_side = [&]() -> hipError_t {
        if (!fork) return hipSuccess;
        hipError_t r = hipEventRecord(c->ev_join, c->stream2);
        return r != hipSuccess ? r : hipStreamWaitEvent(c->stream, c->ev_join, 0);
    };
    if ((e = mark()) != hipSuccess) return e;
    if ((e = launch_green_store(c, it, d_it)) != hipSuccess) return e;
    if ((e = launch_wav_store(c, it, d_it)) != hipSuccess || (e = mark()) != hipSuccess) return e;
    if ((e = launch_snap(c, it, d_it)) != hipSuccess || (e = mark()) != hipSuccess) return e;
    const bool merged = pml && merged_launch_ready(c);
    if ((e = fork_to_side()) != hipSuccess) return e;
    if (pml && !merged && (e = launch_stress_side(c, side)) != hipSuccess) return e;
    if ((e = mark()) != hipSuccess) return e;
    if (merged) {
        if ((e = launch_stress_all(c)) != hipSuccess) return e;
    } else if ((e = launch_stress_cols(c, pml ? 3 : 1)) != hipSuccess) {
        return e;
    }
    if ((e = join_side()) != hipSuccess) return e;
    if ((e = mark()) != hipSuccess) return e;
    if (!pml && (e = launch_cerjan_stress(c)) != hipSuccess) return e;
    if ((e = mark()) != hipSuccess) return e;
    if ((e = launch_stressglut(c, it, d_it)) != hipSuccess || (e = mark()) != hipSuccess) return e;
    if ((e = fork_to_side()) != hipSuccess) return e;
    if (pml && !merged && (e = launch_vel_side(c, side)) != hipSuccess) return e;
    if ((e = mark()) != hipSuccess) return e;
    if (merged) {
        if ((e = launch_vel_all(c)) != hipSuccess) return e;
    } else if ((e = launch_vel_cols(c, pml ? 3 : 1)) != hipSuccess) {
        return e;
    }
    if ((e = join_side()) != hipSuccess) return e;
    if ((e = mark()) != hipSuccess) return e;
    if (!pml && (e = launch_cerjan_vel(c)) != hipSuccess) return e;
    if ((e = mark()) != hipSuccess) return e;
    if ((e = launch_bodyforce(c, it, d_it)) != hipSuccess) return e;
    if ((e = launch_green_source(c, it, d_it)) != hipSuccess || (e = mark()) != hipSuccess) return e;
    if (d_it) e = launch_step_advance(c);
    return e;
}

// variant bit 7: the velocity update of a step and the stress update of the
// next in one interior launch (k_fused), between the side-column and band
// launches of the two phases.  It needs the second field set and the PML,
// and no source that acts between the two updates (body forces, the
// reciprocity force) and no routine that reads the fields between them.
static bool fused_ready(const swpc_ctx *c)
{
    return (c->variant & 128) && c->fields2[0] && c->desc.abc == SWPC_ABC_PML && c->npml > 0 && c->nfz > 0 &&
           c->g.nxk > 0 && c->g.nyk > 0 && !c->fields_in[0] && !(c->nsrc > 0 && c->bf_mode) && !c->green_on;
}

// steps it..it+n-1 (n >= 3) with fused launches.  The first step's stress
// update and the last step's velocity update run in place on the current
// fields; each fused launch between them reads one field set and writes the
// other, an even number of times, so the fields end where they started.
// The outputs of a step (receivers, sections) read the velocities and
// stresses of the step before it, which the two sets hold then.
static hipError_t enqueue_fused_steps(swpc_ctx *c, int it, int n, const int *d_it)
{
    float *A[SWPC_NFIELD], *B[SWPC_NFIELD];
    for (int q = 0; q < SWPC_NFIELD; ++q) {
        A[q] = c->fields[q];
        B[q] = c->fields2[q];
    }
    auto view = [&](float *const *v, float *const *s, float *const *in) {
        for (int q = 0; q < 3; ++q) c->fields[q] = v[q];
        for (int q = 3; q < SWPC_NFIELD; ++q) c->fields[q] = s[q];
        for (int q = 0; q < SWPC_NFIELD; ++q) c->fields_in[q] = in ? in[q] : nullptr;
    };
    auto step_no = [&](int s) { return d_it ? 0 : it + s; };
    hipError_t e;
    // step it up to its stress glut, as in enqueue_step
    if ((e = launch_green_store(c, step_no(0), d_it)) != hipSuccess ||
        (e = launch_wav_store(c, step_no(0), d_it)) != hipSuccess || (e = launch_snap(c, step_no(0), d_it)) != hipSuccess ||
        (e = launch_stress_side(c, nullptr)) != hipSuccess || (e = launch_stress_cols(c, 3)) != hipSuccess ||
        (e = launch_stressglut(c, step_no(0), d_it)) != hipSuccess)
        return e;
    const int nf = (n - 1) & ~1;
    for (int s = 0; s < nf && e == hipSuccess; ++s) {
        float *const *cur = (s & 1) ? B : A;
        float *const *nxt = (s & 1) ? A : B;
        // step it+s: velocities of the side columns and the bottom band (2: band only)
        view(nxt, cur, cur);
        if ((e = launch_vel_side(c, nullptr)) != hipSuccess || (e = launch_vel_cols(c, 2)) != hipSuccess) break;
        // interior: velocities of step it+s and stresses of step it+s+1
        view(nxt, nxt, cur);
        if ((e = launch_fused(c)) != hipSuccess) break;
        // step it+s+1: stresses of the side columns and the bottom band
        if ((e = launch_stress_side(c, nullptr)) != hipSuccess || (e = launch_stress_cols(c, 2)) != hipSuccess) break;
        if (d_it && (e = launch_step_advance(c)) != hipSuccess) break;
        view(nxt, cur, nullptr);
        if ((e = launch_wav_store(c, step_no(s + 1), d_it)) != hipSuccess ||
            (e = launch_snap(c, step_no(s + 1), d_it)) != hipSuccess)
            break;
        view(nxt, nxt, nullptr);
        e = launch_stressglut(c, step_no(s + 1), d_it);
    }
    view(A, A, nullptr);
    if (e != hipSuccess) return e;
    // step it+nf from its velocity update on, in place
    if ((e = launch_vel_side(c, nullptr)) != hipSuccess || (e = launch_vel_cols(c, 3)) != hipSuccess ||
        (e = launch_bodyforce(c, step_no(nf), d_it)) != hipSuccess ||
        (e = launch_green_source(c, step_no(nf), d_it)) != hipSuccess || (d_it && (e = launch_step_advance(c)) != hipSuccess))
        return e;
    for (int s = nf + 1; s < n && e == hipSuccess; ++s) e = enqueue_step(c, step_no(s), d_it);
    return e;
}

static hipError_t enqueue_steps(swpc_ctx *c, int it, int n, const int *d_it)
{
    if (n >= 3 && fused_ready(c)) return enqueue_fused_steps(c, it, n, d_it);
    hipError_t e = hipSuccess;
    for (int s = 0; s < n && e == hipSuccess; ++s) e = enqueue_step(c, d_it ? 0 : it + s, d_it);
    return e;
}

__global__ void k_set_int(int *p, int v)
{
    if (threadIdx.x == 0) *p = v;
}

extern "C" int swpc_step(swpc_ctx *c, int it0, int nsteps, int use_graph, int block)
{
    if (!c) return SWPC_EINVAL;
    if (nsteps < 0) return fail(c, SWPC_EINVAL, "swpc_step: negative step count");
    int rc = ready(c);
    if (rc) return rc;
    if ((rc = check_snap_frames(c, it0, nsteps))) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    int done = 0;
    if (use_graph && block > 0 && nsteps >= block) {
        if (!c->graph_exec || c->graph_block != block) {
            reset_graph(c);
            hipGraph_t graph;
            HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = enqueue_steps(c, 0, block, c->d_it);
            hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
            if (e != hipSuccess) return fail(c, SWPC_EHIP, "graph capture: %s", hipGetErrorString(e));
            if (e2 != hipSuccess) return fail(c, SWPC_EHIP, "graph capture end: %s", hipGetErrorString(e2));
            e = hipGraphInstantiate(&c->graph_exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) {
                c->graph_exec = nullptr;
                return fail(c, SWPC_EHIP, "graph instantiate: %s", hipGetErrorString(e));
            }
            c->graph_block = block;
        }
        hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, c->stream, c->d_it, it0);
        HIPCHK(c, hipGetLastError());
        for (; done + block <= nsteps; done += block) HIPCHK(c, hipGraphLaunch(c->graph_exec, c->stream));
    }
    if (done < nsteps) HIPCHK(c, enqueue_steps(c, it0 + done, nsteps - done, nullptr));
    return SWPC_OK;
}

extern "C" int swpc_step_timed(swpc_ctx *c, int it0, int nsteps, float *ms)
{
    if (!c || !ms) return SWPC_EINVAL;
    if (nsteps < 0) return fail(c, SWPC_EINVAL, "swpc_step_timed: negative step count");
    int rc = ready(c);
    if (rc) return rc;
    if ((rc = check_snap_frames(c, it0, nsteps))) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    for (int t = 0; t < SWPC_NTIMER; ++t) ms[t] = 0.0f;
    hipEvent_t ev[SWPC_NTIMER + 1] = {};
    hipError_t e = hipSuccess;
    for (int t = 0; t <= SWPC_NTIMER && e == hipSuccess; ++t) e = hipEventCreate(&ev[t]);
    for (int s = 0; s < nsteps && e == hipSuccess; ++s) {
        e = enqueue_step(c, it0 + s, nullptr, ev);
        if (e == hipSuccess) e = hipEventSynchronize(ev[SWPC_NTIMER]);
        for (int t = 0; t < SWPC_NTIMER && e == hipSuccess; ++t) {
            float x = 0.0f;
            e = hipEventElapsedTime(&x, ev[t], ev[t + 1]);
            ms[t] += x;
        }
    }
    for (int t = 0; t <= SWPC_NTIMER; ++t)
        if (ev[t]) (void)hipEventDestroy(ev[t]);
    if (e != hipSuccess) return fail(c, SWPC_EHIP, "swpc_step_timed: %s", hipGetErrorString(e));
    return SWPC_OK;
}

extern "C" int swpc_kernel_vmax(swpc_ctx *c, float out[3])
{
    if (!c || !out) return SWPC_EINVAL;
    if (!c->has_medium) return fail(c, SWPC_ESTATE, "medium not set");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, launch_vmax(c));
    HIPCHK(c, hipMemcpyAsync(out, c->d_vmax, 3 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SWPC_OK;
}

extern "C" int swpc_get_field(swpc_ctx *c, int field, float *out)
{
    if (!c || !out || field < 0 || field >= SWPC_NFIELD) return SWPC_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, launch_dev_to_ref(c, c->fields[field], c->staging, c->stream));
    HIPCHK(c, hipMemcpyAsync(out, c->staging, (size_t)c->staging_n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SWPC_OK;
}

extern "C" int swpc_get_medium(swpc_ctx *c, int which, void *out)
{
    if (!c || !out || which < 0 || which >= SWPC_NMEDIUM) return SWPC_EINVAL;
    if (!c->has_medium) return fail(c, SWPC_ESTATE, "swpc_get_medium: medium not set");
    HIPCHK(c, hipSetDevice(c->device));
    if (which <= SWPC_MED_TAUS) {
        float *src[5] = {c->rho, c->lam, c->mu, c->taup, c->taus};
        HIPCHK(c, launch_dev_to_ref(c, src[which], c->staging, c->stream));
        HIPCHK(c, hipMemcpyAsync(out, c->staging, (size_t)c->staging_n * sizeof(float), hipMemcpyDeviceToHost,
                                 c->stream));
    } else {
        const int *src[5] = {c->kfs_top, c->kfs_bot, c->kob_top, c->kob_bot, c->kob};
        const Grid &g = c->g;
        const size_t w = (size_t)(g.nx + 6) * sizeof(int), sp = (size_t)g.pitch * sizeof(int);
        HIPCHK(c, hipMemcpy2DAsync(out, w, src[which - SWPC_MED_KFS_TOP] + g2idx(g, -2, -2), sp, w, g.ny + 6,
                                   hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SWPC_OK;
}

extern "C" int swpc_set_field(swpc_ctx *c, int field, const float *in)
{
    if (!c || !in || field < 0 || field >= SWPC_NFIELD) return SWPC_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    return upload_ref(c, in, c->fields[field]);
}

extern "C" int swpc_get_seismograms(swpc_ctx *c, int kind, float *out)
{
    if (!c || !out) return SWPC_EINVAL;
    int q = -1;
    for (int b = 0; b < 4; ++b)
        if (kind == (1 << b)) q = b;
    if (q < 0) return fail(c, SWPC_EINVAL, "swpc_get_seismograms: kind must be one SWPC_WAV_* flag");
    if (c->nst == 0 || !c->d_wav[q]) return fail(c, SWPC_ESTATE, "swpc_get_seismograms: output not enabled");
    const int ncmp[4] = {3, 3, 6, 6};
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(out, c->d_wav[q], (size_t)c->ntw * ncmp[q] * c->nst * sizeof(float),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SWPC_OK;
}

extern "C" int swpc_synchronize(swpc_ctx *c)
{
    if (!c) return SWPC_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SWPC_OK;
}

extern "C" int swpc_set_tuning(swpc_ctx *c, int kchunk, int by, int variant)
{
    if (!c) return SWPC_EINVAL;
    if (kchunk < -1 || by < 0 || by == 3 || by > 4)
        return fail(c, SWPC_EINVAL, "swpc_set_tuning: kchunk >= -1, by in {1, 2, 4} (launch bound 256)");
    HIPCHK(c, hipSetDevice(c->device));
    // the tables rebuilt below may be read by work on any stream of the
    // device (the context's own, a batch's that steps this member)
    HIPCHK(c, hipDeviceSynchronize());
    if (kchunk > 0) c->kchunk = kchunk;
    else if (kchunk == -1) c->kchunk = default_kchunk(c);
    if (by > 0) c->by = by;
    if (variant >= 0) c->variant = variant;
    reset_graph(c);
    // the second field set of the fused steps, held while variant bit 7 is set
    const size_t fb = (size_t)c->g.n * sizeof(float);
    const bool want2 = (c->variant & 128) && c->desc.abc == SWPC_ABC_PML;
    for (int q = 0; q < SWPC_NFIELD; ++q) {
        if (want2 && !c->fields2[q] && dalloc_striped(c, (void **)&c->fields2[q], fb, &kFieldRegion[q], 1, 0) != hipSuccess) {
            c->fields2[q] = nullptr;
            for (int r = 0; r < SWPC_NFIELD; ++r)
                if (c->fields2[r]) {
                    dfree(c->fields2[r]);
                    c->fields2[r] = nullptr;
                    c->bytes -= (long long)fb;
                }
            return fail(c, SWPC_ENOMEM, "swpc_set_tuning: no device memory for the second field set (variant bit 7)");
        }
        if (!want2 && c->fields2[q]) {
            dfree(c->fields2[q]);
            c->fields2[q] = nullptr;
            c->bytes -= (long long)fb;
        }
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int rc = build_pml_tiles(c);
    if (rc == SWPC_OK) rc = build_zchunks(c);
    if (rc == SWPC_OK && (rc = build_schedule(c)) != SWPC_OK) rc = fail(c, rc, "swpc_set_tuning: launch schedule");
    return rc;
}
