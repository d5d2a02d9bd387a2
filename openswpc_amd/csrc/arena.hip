// arena.hip -- device memory striped over the HBM regions of the device.
//
// The 288 GiB of an MI355X are eight 12-high HBM3E stacks; the physical
// address space falls into three 96 GiB regions (one per stack-ID die
// group).  A kernel that streams many arrays at the same index at once
// (the z-marching stencils stream 10 to 18) runs 12-14 % faster when the
// arrays lie in different regions than when they share one: measured on
// the production grid, the stress update takes 19.2 ms with its arrays
// spread over the three regions and 21.7-22.0 ms with all of them in one.
// A fresh process on an idle device gets its first ~90 GiB from the first
// region, so plain hipMalloc puts a whole simulation there.
//
// swpc_stripe_reserve takes every free granule of the device (hipMemCreate,
// 256 MiB each; on an idle device they come in physical order), keeps
// `bytes_per_region` from the middle of each third of that sequence and
// releases the rest.  Contexts created afterwards on the device take their
// per-cell arrays from the arena: an array of 256 MiB or more gets its own
// virtual range mapped from granules of the region(s) its caller names
// (a multi-stream array such as the packed memory variables names one
// region per stream segment); smaller arrays are sub-allocated from
// per-region slabs.  When the arena is missing or full, allocation falls
// back to hipMalloc.  The arena is process-wide, one per device; contexts
// hold no reference to it, so swpc_stripe_release refuses while any of its
// memory is in use.  A large context created on a device without an arena
// reserves a private one sized for itself (swpc_create), which only that
// context draws from and which goes when it is destroyed.
#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "swpc_internal.h"

namespace {

constexpr size_t kGranule = (size_t)256 << 20;
constexpr size_t kSmallAlign = (size_t)64 << 10;
constexpr int kMaxRegions = 8;
constexpr long long kSpareGranules = 4;  // left free for the runtime while every other granule is held

struct Slab {
    char *va = nullptr;
    hipMemGenericAllocationHandle_t h{};
    std::map<size_t, size_t> free_;  // offset -> bytes
};

struct Region {
    std::vector<hipMemGenericAllocationHandle_t> granules;  // free, in allocation order
    std::vector<Slab *> slabs;
};

struct Live {
    bool big = false;
    size_t bytes = 0;  // big: mapped bytes; small: block bytes
    std::vector<std::pair<int, hipMemGenericAllocationHandle_t>> granules;  // big: (region, handle) per granule
    Slab *slab = nullptr;
    size_t off = 0;
};

struct Arena {
    int device = 0;
    const void *owner = nullptr;  // a context's private arena (swpc_create), nullptr: the caller's
    int nreg = 0;
    size_t gran = 0;
    hipMemAccessDesc access{};
    Region reg[kMaxRegions];
    std::map<char *, Live> live;
};

std::mutex g_mu;
std::map<int, Arena *> g_arena;

// makes a device current for a scope and restores the caller's afterwards
struct DeviceScope {
    int prev = -1;
    hipError_t err = hipSuccess;
    explicit DeviceScope(int device)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        err = hipSetDevice(device);
    }
    ~DeviceScope()
    {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

hipMemAllocationProp device_prop(int device)
{
    hipMemAllocationProp prop = {};
    prop.type = hipMemAllocationTypePinned;
    prop.location.type = hipMemLocationTypeDevice;
    prop.location.id = device;
    return prop;
}

// map n granules at va (one virtual range) and open it for the device
hipError_t map_range(Arena &a, char *va, const hipMemGenericAllocationHandle_t *h, size_t n)
{
    for (size_t k = 0; k < n; ++k) {
        hipError_t e = hipMemMap(va + k * a.gran, a.gran, 0, h[k], 0);
        if (e != hipSuccess) {
            for (size_t j = 0; j < k; ++j) (void)hipMemUnmap(va + j * a.gran, a.gran);
            return e;
        }
    }
    hipError_t e = hipMemSetAccess(va, n * a.gran, &a.access, 1);
    if (e != hipSuccess) (void)hipMemUnmap(va, n * a.gran);
    return e;
}

void *alloc_big(Arena &a, size_t bytes, const int *seg_region, int nseg, size_t seg_bytes)
{
    const size_t ng = (bytes + a.gran - 1) / a.gran;
    size_t need[kMaxRegions] = {};
    std::vector<int> rk(ng);
    size_t avail[kMaxRegions] = {};
    for (int r = 0; r < a.nreg; ++r) avail[r] = a.reg[r].granules.size();
    for (size_t k = 0; k < ng; ++k) {
        // the granule's segment names its region (no table: segments rotate
        // over the regions); a region that has run out yields to the region
        // with the most granules left
        const size_t seg = seg_bytes > 0 ? k * a.gran / seg_bytes : 0;
        int r = seg_region ? seg_region[std::min<size_t>(seg, (size_t)nseg - 1)] : (int)(seg % (size_t)a.nreg);
        r = ((r % a.nreg) + a.nreg) % a.nreg;
        if (avail[r] == 0)
            for (int q = 0; q < a.nreg; ++q)
                if (avail[q] > avail[r]) r = q;
        if (avail[r] == 0) return nullptr;
        --avail[r];
        rk[k] = r;
        ++need[r];
    }
    void *va = nullptr;
    if (hipMemAddressReserve(&va, ng * a.gran, 0, nullptr, 0) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    std::vector<hipMemGenericAllocationHandle_t> hs(ng);
    size_t pos[kMaxRegions] = {};
    for (size_t k = 0; k < ng; ++k) {
        std::vector<hipMemGenericAllocationHandle_t> &g = a.reg[rk[k]].granules;
        hs[k] = g[g.size() - need[rk[k]] + pos[rk[k]]++];
    }
    if (map_range(a, (char *)va, hs.data(), ng) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipMemAddressFree(va, ng * a.gran);
        return nullptr;
    }
    for (int r = 0; r < a.nreg; ++r) a.reg[r].granules.resize(a.reg[r].granules.size() - need[r]);
    Live L;
    L.big = true;
    L.bytes = ng * a.gran;
    for (size_t k = 0; k < ng; ++k) L.granules.push_back({rk[k], hs[k]});
    a.live[(char *)va] = std::move(L);
    return va;
}

void *alloc_small(Arena &a, size_t bytes, int region)
{
    const int r = ((region % a.nreg) + a.nreg) % a.nreg;
    const size_t need = (bytes + kSmallAlign - 1) / kSmallAlign * kSmallAlign;
    Region &R = a.reg[r];
    Slab *slab = nullptr;
    std::map<size_t, size_t>::iterator f;
    for (Slab *s : R.slabs) {
        for (f = s->free_.begin(); f != s->free_.end(); ++f)
            if (f->second >= need) break;
        if (f != s->free_.end()) { slab = s; break; }
    }
    if (!slab) {
        if (R.granules.empty()) return nullptr;
        void *va = nullptr;
        if (hipMemAddressReserve(&va, a.gran, 0, nullptr, 0) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        const hipMemGenericAllocationHandle_t h = R.granules.back();
        if (map_range(a, (char *)va, &h, 1) != hipSuccess) {
            (void)hipGetLastError();
            (void)hipMemAddressFree(va, a.gran);
            return nullptr;
        }
        R.granules.pop_back();
        slab = new Slab;
        slab->va = (char *)va;
        slab->h = h;
        slab->free_[0] = a.gran;
        R.slabs.push_back(slab);
        f = slab->free_.begin();
    }
    const size_t off = f->first, sz = f->second;
    slab->free_.erase(f);
    if (sz > need) slab->free_[off + need] = sz - need;
    Live L;
    L.big = false;
    L.bytes = need;
    L.slab = slab;
    L.off = off;
    char *p = slab->va + off;
    a.live[p] = std::move(L);
    return p;
}

void release_all(Arena &a)
{
    for (int r = 0; r < a.nreg; ++r) {
        for (Slab *s : a.reg[r].slabs) {
            (void)hipMemUnmap(s->va, a.gran);
            (void)hipMemAddressFree(s->va, a.gran);
            (void)hipMemRelease(s->h);
            delete s;
        }
        a.reg[r].slabs.clear();
        for (auto h : a.reg[r].granules) (void)hipMemRelease(h);
        a.reg[r].granules.clear();
    }
}

}  // namespace

bool arena_active(int device)
{
    std::lock_guard<std::mutex> lk(g_mu);
    return g_arena.count(device) != 0;
}

void *arena_alloc(int device, const void *who, size_t bytes, const int *seg_region, int nseg, size_t seg_bytes)
{
    if (bytes == 0 || (seg_region && nseg < 1)) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_arena.find(device);
    if (it == g_arena.end()) return nullptr;
    Arena &a = *it->second;
    if (a.owner && a.owner != who) return nullptr;
    if (bytes >= a.gran) return alloc_big(a, bytes, seg_region, nseg, seg_bytes);
    return alloc_small(a, bytes, seg_region ? seg_region[0] : 0);
}

int64_t arena_granule_bytes(int device, const void *who)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_arena.find(device);
    if (it == g_arena.end() || (it->second->owner && it->second->owner != who)) return 0;
    int64_t b = 0;
    for (int r = 0; r < it->second->nreg; ++r) b += (int64_t)(it->second->reg[r].granules.size() * it->second->gran);
    return b;
}

bool arena_free(void *p)
{
    if (!p) return false;
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto &kv : g_arena) {
        Arena &a = *kv.second;
        auto it = a.live.find((char *)p);
        if (it == a.live.end()) continue;
        Live &L = it->second;
        // kernels still in flight may read the block: wait for the device
        DeviceScope ds(a.device);
        (void)hipDeviceSynchronize();
        if (L.big) {
            (void)hipMemUnmap(p, L.bytes);
            (void)hipMemAddressFree(p, L.bytes);
            for (auto &g : L.granules) a.reg[g.first].granules.push_back(g.second);
        } else {
            Slab *s = L.slab;
            size_t off = L.off, sz = L.bytes;
            auto nx = s->free_.lower_bound(off);
            if (nx != s->free_.end() && nx->first == off + sz) {
                sz += nx->second;
                s->free_.erase(nx);
            }
            auto pv = s->free_.lower_bound(off);
            if (pv != s->free_.begin()) {
                --pv;
                if (pv->first + pv->second == off) {
                    off = pv->first;
                    sz += pv->second;
                    s->free_.erase(pv);
                }
            }
            s->free_[off] = sz;
        }
        a.live.erase(it);
        return true;
    }
    return false;
}

int arena_reserve(int device, int64_t bytes_per_region, int nregions, const void *owner)
{
    if (nregions < 1 || nregions > kMaxRegions || bytes_per_region <= 0) return SWPC_EINVAL;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_arena.count(device)) return SWPC_ESTATE;
    DeviceScope ds(device);
    if (ds.err != hipSuccess) return SWPC_EHIP;
    int vmm = 0;
    if (hipDeviceGetAttribute(&vmm, hipDeviceAttributeVirtualMemoryManagementSupported, device) != hipSuccess || !vmm)
        return SWPC_EHIP;
    const hipMemAllocationProp prop = device_prop(device);
    size_t gmin = 0;
    if (hipMemGetAllocationGranularity(&gmin, &prop, hipMemAllocationGranularityMinimum) != hipSuccess || gmin == 0)
        return SWPC_EHIP;
    const size_t G = std::max(gmin, kGranule / gmin * gmin);
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return SWPC_EHIP;
    const size_t per = ((size_t)bytes_per_region + G - 1) / G;
    const long long n = (long long)(fr / G) - kSpareGranules;
    if (n < (long long)per * nregions) return SWPC_ENOMEM;
    // hold every free granule, in allocation order
    std::vector<hipMemGenericAllocationHandle_t> all;
    all.reserve((size_t)n);
    for (long long i = 0; i < n; ++i) {
        hipMemGenericAllocationHandle_t h;
        if (hipMemCreate(&h, G, &prop, 0) != hipSuccess) {
            (void)hipGetLastError();
            break;
        }
        all.push_back(h);
    }
    const size_t third = all.size() / nregions;
    const bool ok = third >= per;
    Arena *a = ok ? new Arena : nullptr;
    std::vector<char> keep(all.size(), 0);
    if (ok) {
        a->device = device;
        a->nreg = nregions;
        a->gran = G;
        a->access.location = prop.location;
        a->access.flags = hipMemAccessFlagsProtReadWrite;
        // the middle of each third: clear of the region boundaries even when
        // the first granule does not sit at the bottom of the address space
        for (int r = 0; r < nregions; ++r) {
            const size_t s = (size_t)r * third + (third - per) / 2;
            for (size_t k = 0; k < per; ++k) {
                keep[s + k] = 1;
                a->reg[r].granules.push_back(all[s + k]);
            }
            std::reverse(a->reg[r].granules.begin(), a->reg[r].granules.end());  // taken from the back
        }
    }
    for (size_t i = 0; i < all.size(); ++i)
        if (!keep[i]) (void)hipMemRelease(all[i]);
    if (!ok) return SWPC_ENOMEM;
    a->owner = owner;
    g_arena[device] = a;
    return SWPC_OK;
}

extern "C" int swpc_stripe_reserve(int device, int64_t bytes_per_region, int nregions)
{
    return arena_reserve(device, bytes_per_region, nregions, nullptr);
}

extern "C" int swpc_stripe_release(int device)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_arena.find(device);
    if (it == g_arena.end()) return SWPC_OK;
    Arena *a = it->second;
    if (!a->live.empty()) return SWPC_ESTATE;
    DeviceScope ds(device);
    (void)hipDeviceSynchronize();
    release_all(*a);
    delete a;
    g_arena.erase(it);
    return SWPC_OK;
}

extern "C" int64_t swpc_stripe_free_bytes(int device, int region)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_arena.find(device);
    if (it == g_arena.end()) return 0;
    const Arena &a = *it->second;
    if (region < 0 || region >= a.nreg) return SWPC_EINVAL;
    int64_t b = (int64_t)(a.reg[region].granules.size() * a.gran);
    for (const Slab *s : a.reg[region].slabs)
        for (const auto &f : s->free_) b += (int64_t)f.second;
    return b;
}
