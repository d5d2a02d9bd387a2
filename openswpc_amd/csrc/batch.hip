// batch.hip -- many independent simulations stepped together (swpc_batch).
//
// A batch references member contexts that were created and set up as usual
// (swpc_create, swpc_set_*).  Each phase of a step is one launch for all
// members: the launch walks a tile table whose entries carry the member
// index, and each member's launch arguments sit in HBM.  Uneven grids pack
// into one grid of blocks with no per-member launch tail.  Receivers,
// sources, sections, Green's-function stores and Cerjan sponges are
// member-tagged launches too: one per kind and step for every member that
// has them.
// Steps replay from a HIP graph of `block` steps driven by one device step
// counter shared by all members.
//
// Tile packing, chosen by the autotuner among:
//   0: member-major (each member's tiles contiguous, so a member's
//      neighbouring tiles share an XCD's L2), members ordered by size,
//      largest first, so the smallest members fill the end of the launch;
//   1: chunk-major: the tiles of every member's first k chunk, then of every
//      member's second chunk, and so on (members largest first within a
//      chunk), so each stretch of the launch streams many members at once;
//   2: member-major with the side-column tiles in the interior launch: one
//      launch per phase, each member's x-slab tiles next to the interior tile
//      rows they border, so the cache lines the two share are fetched once.
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <numeric>
#include <vector>

#include "swpc_internal.h"

struct swpc_batch {
    std::vector<swpc_ctx *> mem;
    std::vector<unsigned> gen;     // member config generations the tables were built for
    int device;
    hipStream_t stream;
    int nm, by;
    int packing;                   // tile order of the tables (see above)
    void *d_args[2];               // StepArgs of every member: stress, velocity
    int4 *d_cols, *d_side;
    int ncols, nside;
    int4 *d_all;                   // packing 2: interior and side tiles in one table
    int nall;
    void *d_aux[BATCH_AUX_N];      // receiver, source and Green's-function arguments of the members that have them
    int naux[BATCH_AUX_N], maxaux[BATCH_AUX_N];
    void *d_snap;                  // section arguments of the members with sections
    int nsnap, maxsnap;
    void *d_cerj[2];               // sponge arguments (stress, velocity phase) of the Cerjan members
    int *d_cerj_blk0[2];           // their first blocks in the flat sponge launch (ncerj + 1 entries)
    int ncerj, ncerj_blocks[2];
    int *d_it;
    hipGraphExec_t graph;
    int graph_block;
    char err[512];
};

static int bfail(swpc_batch *b, int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(b->err, sizeof(b->err), fmt, ap);
    va_end(ap);
    return code;
}

#define BCHK(b, call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) return bfail(b, SWPC_EHIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

static void bfree(void *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

static void breset_graph(swpc_batch *b)
{
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    b->graph_block = 0;
}

static void free_tables(swpc_batch *b)
{
    bfree(b->d_args[0]);
    bfree(b->d_args[1]);
    void *p = b->d_cols;
    bfree(p);
    p = b->d_side;
    bfree(p);
    b->d_cols = b->d_side = nullptr;
    b->ncols = b->nside = 0;
    p = b->d_all;
    bfree(p);
    b->d_all = nullptr;
    b->nall = 0;
    for (int k = 0; k < BATCH_AUX_N; ++k) {
        bfree(b->d_aux[k]);
        b->d_aux[k] = nullptr;
        b->naux[k] = b->maxaux[k] = 0;
    }
    bfree(b->d_snap);
    b->nsnap = b->maxsnap = 0;
    for (int ph = 0; ph < 2; ++ph) {
        bfree(b->d_cerj[ph]);
        p = b->d_cerj_blk0[ph];
        bfree(p);
        b->d_cerj_blk0[ph] = nullptr;
        b->ncerj_blocks[ph] = 0;
    }
    b->ncerj = 0;
}

// interior tile width of the members (variant bit 2: 32 cells, else 64)
static int batch_bxw(const swpc_batch *b) { return (b->mem[0]->variant & 4) ? 32 : 64; }

// (re)build the argument arrays and tile tables from the members' current state
static int build_tables(swpc_batch *b)
{
    // the batch stream may still run replays that read the old tables
    BCHK(b, hipStreamSynchronize(b->stream));
    breset_graph(b);
    free_tables(b);
    const int n = (int)b->mem.size();
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        const Grid &gx = b->mem[x]->g, &gy = b->mem[y]->g;
        return (long long)gx.nx * gx.ny * gx.nz > (long long)gy.nx * gy.ny * gy.nz;
    });
    const size_t asz = step_args_size();
    std::vector<char> args[2] = {std::vector<char>(asz * n), std::vector<char>(asz * n)};
    // per member, its interior tiles and side tiles grouped by k chunk
    std::vector<std::vector<std::vector<int4>>> mcols(n), mside(n);
    int maxchunks = 0;
    for (int m : order) {
        swpc_ctx *c = b->mem[m];
        if (c->by != b->by) return bfail(b, SWPC_ESTATE, "member %d has tile rows %d, batch %d", m, c->by, b->by);
        fill_batch_args(c, args[0].data() + asz * m, 0);
        fill_batch_args(c, args[1].data() + asz * m, 1);
        int nbx, nby, nbz;
        member_tiles(c, nbx, nby, nbz);
        const int nch = (c->g.nz + c->kchunk - 1) / c->kchunk;
        mcols[m].assign(std::max(nbz, nch), {});
        mside[m].assign(std::max(nbz, nch), {});
        maxchunks = std::max(maxchunks, (int)mcols[m].size());
        if (c->g.nxk > 0 && c->g.nyk > 0)
            for (int bz = 0; bz < nbz; ++bz)
                for (int by = 0; by < nby; ++by)
                    for (int bx = 0; bx < nbx; ++bx) mcols[m][bz].push_back(make_int4(m, bx, by, bz));
        if (c->desc.abc == SWPC_ABC_PML && c->npml_tiles > 0) {
            std::vector<int4> t(c->npml_tiles);
            BCHK(b, hipMemcpy(t.data(), c->d_pml_tiles, t.size() * sizeof(int4), hipMemcpyDeviceToHost));
            for (const int4 &x : t) {
                if (x.z > 0xFFFF || x.w > 0x7FFF) return bfail(b, SWPC_EINVAL, "member %d: nz too large", m);
                const int ch = std::min((x.z - 1) / c->kchunk, (int)mside[m].size() - 1);
                mside[m][ch].push_back(make_int4(x.x, x.y, x.z | (x.w << 16), m));
            }
        }
    }
    std::vector<int4> cols, side, all;
    if (b->packing == 2) {
        const int byi = b->by * (64 / batch_bxw(b));
        for (int m : order) {
            const Grid &g = b->mem[m]->g;
            for (size_t ch = 0; ch < mcols[m].size(); ++ch) {
                std::vector<int4> low, high, xs;
                for (const int4 &t : mside[m][ch]) {
                    const int4 e = make_int4(-(m + 1), t.x, t.y, t.z);
                    if ((t.x >> 28) & 1) xs.push_back(e);
                    else if (t.y <= g.na) low.push_back(e);
                    else high.push_back(e);
                }
                all.insert(all.end(), low.begin(), low.end());
                size_t q = 0;
                int row = -1;
                for (const int4 &t : mcols[m][ch]) {
                    if (t.z != row) {
                        row = t.z;
                        const int j1 = g.jb + (row + 1) * byi - 1;
                        while (q < xs.size() && xs[q].z <= j1) all.push_back(xs[q++]);
                    }
                    all.push_back(t);
                }
                for (; q < xs.size(); ++q) all.push_back(xs[q]);
                all.insert(all.end(), high.begin(), high.end());
            }
        }
    } else if (b->packing == 1) {
        for (int ch = 0; ch < maxchunks; ++ch)
            for (int m : order) {
                if (ch < (int)mcols[m].size()) cols.insert(cols.end(), mcols[m][ch].begin(), mcols[m][ch].end());
                if (ch < (int)mside[m].size()) side.insert(side.end(), mside[m][ch].begin(), mside[m][ch].end());
            }
    } else {
        for (int m : order)
            for (size_t ch = 0; ch < mcols[m].size(); ++ch) {
                cols.insert(cols.end(), mcols[m][ch].begin(), mcols[m][ch].end());
                side.insert(side.end(), mside[m][ch].begin(), mside[m][ch].end());
            }
    }
    for (int ph = 0; ph < 2; ++ph) {
        BCHK(b, hipMalloc(&b->d_args[ph], asz * n));
        BCHK(b, hipMemcpy(b->d_args[ph], args[ph].data(), asz * n, hipMemcpyHostToDevice));
    }
    if (!cols.empty()) {
        BCHK(b, hipMalloc((void **)&b->d_cols, cols.size() * sizeof(int4)));
        BCHK(b, hipMemcpy(b->d_cols, cols.data(), cols.size() * sizeof(int4), hipMemcpyHostToDevice));
    }
    if (!side.empty()) {
        BCHK(b, hipMalloc((void **)&b->d_side, side.size() * sizeof(int4)));
        BCHK(b, hipMemcpy(b->d_side, side.data(), side.size() * sizeof(int4), hipMemcpyHostToDevice));
    }
    if (!all.empty()) {
        BCHK(b, hipMalloc((void **)&b->d_all, all.size() * sizeof(int4)));
        BCHK(b, hipMemcpy(b->d_all, all.data(), all.size() * sizeof(int4), hipMemcpyHostToDevice));
    }
    b->ncols = (int)cols.size();
    b->nside = (int)side.size();
    b->nall = (int)all.size();
    // receivers and sources: one launch per kind for every member that has them
    for (int k = 0; k < BATCH_AUX_N; ++k) {
        const size_t sz = batch_aux_size(k);
        std::vector<char> buf(sz * n);
        int cnt = 0, mx = 0;
        for (int m = 0; m < n; ++m) {
            const int t = batch_aux_fill(b->mem[m], k, buf.data() + sz * cnt);
            if (t > 0) {
                ++cnt;
                mx = std::max(mx, t);
            }
        }
        if (cnt > 0) {
            BCHK(b, hipMalloc(&b->d_aux[k], sz * cnt));
            BCHK(b, hipMemcpy(b->d_aux[k], buf.data(), sz * cnt, hipMemcpyHostToDevice));
        }
        b->naux[k] = cnt;
        b->maxaux[k] = mx;
    }
    // sections: one launch for every member that has them
    {
        const size_t sz = batch_snap_size();
        std::vector<char> buf(sz * n);
        int cnt = 0, mx = 0;
        for (int m = 0; m < n; ++m) {
            const int t = batch_snap_fill(b->mem[m], buf.data() + sz * cnt);
            if (t > 0) {
                ++cnt;
                mx = std::max(mx, t);
            }
        }
        if (cnt > 0) {
            BCHK(b, hipMalloc(&b->d_snap, sz * cnt));
            BCHK(b, hipMemcpy(b->d_snap, buf.data(), sz * cnt, hipMemcpyHostToDevice));
        }
        b->nsnap = cnt;
        b->maxsnap = mx;
    }
    // Cerjan sponges: one flat launch per phase over the sponge members' blocks
    for (int ph = 0; ph < 2; ++ph) {
        const size_t sz = batch_cerjan_size();
        std::vector<char> buf(sz * n);
        std::vector<int> blk0(1, 0);
        long long total = 0;
        int cnt = 0;
        for (int m = 0; m < n; ++m) {
            const int t = batch_cerjan_fill(b->mem[m], ph, buf.data() + sz * cnt);
            if (t > 0) {
                ++cnt;
                total += t;
                if (total > 0x7FFFFFFFLL) return bfail(b, SWPC_EINVAL, "sponge members too large for one launch");
                blk0.push_back((int)total);
            }
        }
        if (cnt > 0) {
            BCHK(b, hipMalloc(&b->d_cerj[ph], sz * cnt));
            BCHK(b, hipMemcpy(b->d_cerj[ph], buf.data(), sz * cnt, hipMemcpyHostToDevice));
            BCHK(b, hipMalloc((void **)&b->d_cerj_blk0[ph], blk0.size() * sizeof(int)));
            BCHK(b, hipMemcpy(b->d_cerj_blk0[ph], blk0.data(), blk0.size() * sizeof(int), hipMemcpyHostToDevice));
        }
        b->ncerj = cnt;
        b->ncerj_blocks[ph] = (int)total;
    }
    for (int m = 0; m < n; ++m) b->gen[m] = b->mem[m]->gen;
    return SWPC_OK;
}

static bool tables_stale(const swpc_batch *b)
{
    for (size_t m = 0; m < b->mem.size(); ++m)
        if (b->mem[m]->gen != b->gen[m]) return true;
    return false;
}


// one step of every member, in the order of src/swpc_3d/main.f90:122-138
static hipError_t enqueue_batch_step(swpc_batch *b, int it, const int *d_it)
{
    hipError_t e;
    hipStream_t s = b->stream;
    if ((e = launch_batch_aux(BATCH_AUX_GREEN_STORE, b->d_aux[BATCH_AUX_GREEN_STORE], b->naux[BATCH_AUX_GREEN_STORE],
                              b->maxaux[BATCH_AUX_GREEN_STORE], it, d_it, s)) != hipSuccess)
        return e;
    if ((e = launch_batch_snap(b->d_snap, b->nsnap, b->maxsnap, it, d_it, s)) != hipSuccess) return e;
    if ((e = launch_batch_aux(BATCH_AUX_WAV, b->d_aux[BATCH_AUX_WAV], b->naux[BATCH_AUX_WAV],
                              b->maxaux[BATCH_AUX_WAV], it, d_it, s)) != hipSuccess)
        return e;
    if (b->packing == 2) e = launch_batch_phase_all(0, b->nm, b->mem[0]->variant, b->by, batch_bxw(b), b->d_args[0],
                                                    b->d_all, b->nall, s);
    else e = launch_batch_phase(0, b->nm, b->mem[0]->variant, b->by, batch_bxw(b), b->d_args[0], b->d_cols, b->ncols,
                                b->d_side, b->nside, s);
    if (e != hipSuccess) return e;
    if ((e = launch_batch_cerjan(0, b->d_cerj[0], b->d_cerj_blk0[0], b->ncerj, b->ncerj_blocks[0], s)) != hipSuccess)
        return e;
    // each member's sponge before its stress glut, as in its own step
    if ((e = launch_batch_aux(BATCH_AUX_GLUT, b->d_aux[BATCH_AUX_GLUT], b->naux[BATCH_AUX_GLUT],
                              b->maxaux[BATCH_AUX_GLUT], it, d_it, s)) != hipSuccess)
        return e;
    if (b->packing == 2) e = launch_batch_phase_all(1, b->nm, b->mem[0]->variant, b->by, batch_bxw(b), b->d_args[1],
                                                    b->d_all, b->nall, s);
    else e = launch_batch_phase(1, b->nm, b->mem[0]->variant, b->by, batch_bxw(b), b->d_args[1], b->d_cols, b->ncols,
                                b->d_side, b->nside, s);
    if (e != hipSuccess) return e;
    if ((e = launch_batch_cerjan(1, b->d_cerj[1], b->d_cerj_blk0[1], b->ncerj, b->ncerj_blocks[1], s)) != hipSuccess)
        return e;
    // sponge, body force, then the Green's-function source, per member as in its own step
    if ((e = launch_batch_aux(BATCH_AUX_BODYFORCE, b->d_aux[BATCH_AUX_BODYFORCE], b->naux[BATCH_AUX_BODYFORCE],
                              b->maxaux[BATCH_AUX_BODYFORCE], it, d_it, s)) != hipSuccess)
        return e;
    if ((e = launch_batch_aux(BATCH_AUX_GREEN_SOURCE, b->d_aux[BATCH_AUX_GREEN_SOURCE],
                              b->naux[BATCH_AUX_GREEN_SOURCE], b->maxaux[BATCH_AUX_GREEN_SOURCE], it, d_it, s)) != hipSuccess)
        return e;
    if (d_it) e = launch_counter_advance(b->d_it, s);
    return e;
}

extern "C" const char *swpc_batch_last_error(const swpc_batch *b) { return b ? b->err : "null batch"; }

extern "C" void *swpc_batch_stream(swpc_batch *b) { return b ? (void *)b->stream : nullptr; }

extern "C" int swpc_batch_destroy(swpc_batch *b)
{
    if (!b) return SWPC_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    breset_graph(b);
    free_tables(b);
    void *p = b->d_it;
    bfree(p);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return SWPC_OK;
}

extern "C" int swpc_batch_create(swpc_ctx *const *members, int n, swpc_batch **out)
{
    if (!members || n < 1 || !out) return SWPC_EINVAL;
    *out = nullptr;
    for (int m = 0; m < n; ++m) {
        swpc_ctx *c = members[m];
        if (!c) return SWPC_EINVAL;
        if (c->desc.device != members[0]->desc.device || c->desc.nm != members[0]->desc.nm) return SWPC_EINVAL;
        if (!c->has_medium || (c->desc.nm > 0 && !c->has_visco) ||
            (c->desc.abc == SWPC_ABC_PML && !c->has_pml) || (c->desc.abc == SWPC_ABC_CERJAN && !c->has_cerjan))
            return SWPC_ESTATE;
        for (int q = 0; q < m; ++q)
            if (members[q] == c) return SWPC_EINVAL;
    }
    swpc_batch *b = new swpc_batch();
    b->mem.assign(members, members + n);
    b->gen.assign(n, 0);
    b->device = members[0]->desc.device;
    b->nm = members[0]->desc.nm;
    b->by = members[0]->by;
    int rc = SWPC_OK;
    if (hipSetDevice(b->device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void **)&b->d_it, sizeof(int)) != hipSuccess) {
        swpc_batch_destroy(b);
        return SWPC_EHIP;
    }
    // members share the batch's tile rows and variant
    for (swpc_ctx *c : b->mem)
        if (c->by != b->by || c->variant != members[0]->variant) {
            if ((rc = swpc_set_tuning(c, 0, b->by, members[0]->variant)) != SWPC_OK) {
                swpc_batch_destroy(b);
                return rc;
            }
        }
    if ((rc = build_tables(b)) != SWPC_OK) {
        swpc_batch_destroy(b);
        return rc;
    }
    *out = b;
    return SWPC_OK;
}

static int batch_sync_members(swpc_batch *b)
{
    // members may have work queued on their own streams (uploads, own steps)
    for (swpc_ctx *c : b->mem) BCHK(b, hipStreamSynchronize(c->stream));
    return SWPC_OK;
}

extern "C" int swpc_batch_step(swpc_batch *b, int it0, int nsteps, int use_graph, int block)
{
    if (!b) return SWPC_EINVAL;
    if (nsteps < 0) return bfail(b, SWPC_EINVAL, "swpc_batch_step: negative step count");
    BCHK(b, hipSetDevice(b->device));
    int rc;
    if ((rc = batch_sync_members(b))) return rc;
    if (tables_stale(b) && (rc = build_tables(b))) return rc;
    for (swpc_ctx *c : b->mem)
        if (c->snap_ntask > 0 && nsteps > 0 && it0 + nsteps - 1 >= 1 &&
            (it0 + nsteps - 2) / c->snap.ntdec_s >= c->snap.nframe)
            return bfail(b, SWPC_EINVAL, "a member's snapshot frame store is too small for these steps");
    int done = 0;
    if (use_graph && block > 0 && nsteps >= block) {
        if (!b->graph || b->graph_block != block) {
            breset_graph(b);
            hipGraph_t graph;
            BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = hipSuccess;
            for (int s = 0; s < block && e == hipSuccess; ++s) e = enqueue_batch_step(b, 0, b->d_it);
            hipError_t e2 = hipStreamEndCapture(b->stream, &graph);
            if (e != hipSuccess) return bfail(b, SWPC_EHIP, "batch graph capture: %s", hipGetErrorString(e));
            if (e2 != hipSuccess) return bfail(b, SWPC_EHIP, "batch graph capture end: %s", hipGetErrorString(e2));
            e = hipGraphInstantiate(&b->graph, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) {
                b->graph = nullptr;
                return bfail(b, SWPC_EHIP, "batch graph instantiate: %s", hipGetErrorString(e));
            }
            b->graph_block = block;
        }
        BCHK(b, launch_set_counter(b->d_it, it0, b->stream));
        for (; done + block <= nsteps; done += block) BCHK(b, hipGraphLaunch(b->graph, b->stream));
    }
    for (; done < nsteps; ++done) BCHK(b, enqueue_batch_step(b, it0 + done, nullptr));
    return SWPC_OK;
}

extern "C" int swpc_batch_set_packing(swpc_batch *b, int packing)
{
    if (!b) return SWPC_EINVAL;
    if (packing < 0 || packing > 2) return bfail(b, SWPC_EINVAL, "swpc_batch_set_packing: packing 0, 1 or 2");
    BCHK(b, hipSetDevice(b->device));
    if (packing == b->packing) return SWPC_OK;
    b->packing = packing;
    return build_tables(b);
}

extern "C" int swpc_batch_get_packing(const swpc_batch *b) { return b ? b->packing : SWPC_EINVAL; }

extern "C" int swpc_batch_synchronize(swpc_batch *b)
{
    if (!b) return SWPC_EINVAL;
    BCHK(b, hipSetDevice(b->device));
    BCHK(b, hipStreamSynchronize(b->stream));
    return SWPC_OK;
}

// Autotuner.  Candidate launch shapes (k chunk depth, tile rows) are timed
// on dry steps: step index 0 injects no source and records nothing, so on
// the all-zero state before the first step the trials change no field.
extern "C" int swpc_batch_tune(swpc_batch *b, int trial_steps, int *kchunk_out, int *by_out, int *width_out,
                               float *ms_out)
{
    if (!b || trial_steps < 1) return SWPC_EINVAL;
    BCHK(b, hipSetDevice(b->device));
    int rc;
    if ((rc = batch_sync_members(b))) return rc;
    const int kc_cand[] = {8, 12, 16, 24, 32, 48, 64};
    const int by_cand[] = {2, 4};
    const int w_cand[] = {64, 32};
    const int pk_cand[] = {0, 1, 2};
    const int base = b->mem[0]->variant & ~4;
    hipEvent_t e0, e1;
    BCHK(b, hipEventCreate(&e0));
    BCHK(b, hipEventCreate(&e1));
    float best = 1e30f;
    int best_kc = b->mem[0]->kchunk, best_by = b->by, best_w = batch_bxw(b), best_pk = b->packing;
    for (int w : w_cand)
        for (int byv : by_cand)
            for (int kc : kc_cand) {
                const int variant = base | (w == 32 ? 4 : 0);
                for (swpc_ctx *c : b->mem)
                    if ((rc = swpc_set_tuning(c, kc, byv, variant)) != SWPC_OK) goto out;
                b->by = byv;
                for (int pk : pk_cand) {
                    b->packing = pk;
                    if ((rc = build_tables(b)) != SWPC_OK) goto out;
                    {
                        hipError_t e = enqueue_batch_step(b, 0, nullptr);  // warm
                        if (e == hipSuccess) e = hipEventRecord(e0, b->stream);
                        for (int s = 0; s < trial_steps && e == hipSuccess; ++s) e = enqueue_batch_step(b, 0, nullptr);
                        if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
                        if (e == hipSuccess) e = hipEventSynchronize(e1);
                        float ms = 0.0f;
                        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
                        if (e != hipSuccess) {
                            rc = bfail(b, SWPC_EHIP, "tuning trial: %s", hipGetErrorString(e));
                            goto out;
                        }
                        ms /= trial_steps;
                        if (ms < best) {
                            best = ms;
                            best_kc = kc;
                            best_by = byv;
                            best_w = w;
                            best_pk = pk;
                        }
                    }
                }
            }
    for (swpc_ctx *c : b->mem)
        if ((rc = swpc_set_tuning(c, best_kc, best_by, base | (best_w == 32 ? 4 : 0))) != SWPC_OK) goto out;
    b->by = best_by;
    b->packing = best_pk;
    rc = build_tables(b);
    if (kchunk_out) *kchunk_out = best_kc;
    if (by_out) *by_out = best_by;
    if (width_out) *width_out = best_w;
    if (ms_out) *ms_out = best;
out:
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}
