// Batched step over caller-owned grids without allocation (se_grids_step_noalloc
// in include/sandengine_hip.h).  No counterpart in the reference: it serves
// callers that cannot allocate device memory (code under HIP graph capture,
// serving loops with their own arena, harnesses that forbid hipMalloc).
//
// The per-call device data of the batched LDS kernel -- launch records (one
// se_bworld per grid), tile prefix sums and modifications -- lives in
// `scratch`, a fixed array of this translation unit's code object.  In host
// code `scratch` names a host-side stand-in, so its device address is looked up
// once per device with hipGetSymbolAddress when a context is created on that
// device (se_rt::noalloc_scratch, called by se_ctx_create) and only that
// address is used afterwards.  Each call fills the context's pinned staging
// buffer (allocated with the context), copies it into the array on the call's
// stream and launches the rule set's se_batch_lds_* kernels with pointers into
// the array: no hipMalloc, hipFree or other allocation happens per call.
//
// Nothing in the array has to start at zero: every call writes the records,
// prefix sums and modifications it launches with before its kernels read them.
// Calls that overlap on one device share the array and race (see the header).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "se_internal.hpp"

using se_rt::DeviceGuard;
using se_rt::fail;
using se_rt::launch;
using se_rt::launch_ex;

namespace {

// == se_bworld in se_kernels.hip
struct BWorld {
    uint32_t* buf[2];
    int W, H, pitch, tiles_x, ntiles;
    int frame0, steps, cur0;
    const void* mods;
    int nmods, pad;
};
static_assert(sizeof(BWorld) == 64, "se_bworld layout");

// Layout of the array (and of the pinned staging buffer): records, the two
// prefix arrays (single-tile grids, multi-tile grids), modifications.
constexpr size_t kRecBytes = sizeof(BWorld) * SE_NOALLOC_MAX_GRIDS;
constexpr size_t kPreBytes = (((size_t)SE_NOALLOC_MAX_GRIDS + 1) * sizeof(int) + 255) & ~(size_t)255;
constexpr size_t kOffPreSmall = kRecBytes;
constexpr size_t kOffPreLarge = kOffPreSmall + kPreBytes;
constexpr size_t kOffMods = kOffPreLarge + kPreBytes;
constexpr size_t kUsed = kOffMods + sizeof(se_modification) * SE_NOALLOC_MAX_MODS;
static_assert(kUsed <= SE_NOALLOC_SCRATCH_BYTES, "scratch layout exceeds SE_NOALLOC_SCRATCH_BYTES");

}  // namespace

static __device__ __align__(256) unsigned char scratch[SE_NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[64] = {};  // device address of `scratch` per device
}  // namespace

// Device address of this device's `scratch` (the device must be current):
// looked up on the first call per device and kept.
int se_rt::noalloc_scratch(int device, void** out) {
    if (device < 0 || device >= 64) return fail(SE_ERR_ARG, "device index out of range for the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(SE_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return SE_OK;
}

extern "C" int se_grids_step_noalloc(se_ctx* ctx, const se_ruleset* rs, se_grid_step* grids, uint32_t n,
                                     const se_modification* mods, uint32_t nmods, void* stream) {
    if (!ctx || !rs || (!grids && n)) return fail(SE_ERR_ARG, "null argument");
    if (rs->ctx != ctx) return fail(SE_ERR_ARG, "rule set belongs to another context");
    if (n > SE_NOALLOC_MAX_GRIDS)
        return fail(SE_ERR_ARG, "more than SE_NOALLOC_MAX_GRIDS (" + std::to_string(SE_NOALLOC_MAX_GRIDS) +
                                    ") grids in one call");
    if (nmods > SE_NOALLOC_MAX_MODS)
        return fail(SE_ERR_ARG, "more than SE_NOALLOC_MAX_MODS (" + std::to_string(SE_NOALLOC_MAX_MODS) +
                                    ") modifications in one call");
    if (nmods && !mods) return fail(SE_ERR_ARG, "null modification array");
    if (!ctx->scratch || !ctx->staging || !ctx->staging_done)
        return fail(SE_ERR_STATE, "the scratch array or its staging was not set up when the context was created");
    for (uint32_t i = 0; i < n; i++) {
        const se_grid_step& g = grids[i];
        if (!g.grid[0] || !g.grid[1] || g.w == 0 || g.h == 0 || g.w > (1u << 20) || g.h > (1u << 20) ||
            g.pitch < g.w || (g.cur != 0 && g.cur != 1) || g.frame < 0)
            return fail(SE_ERR_ARG, "grid " + std::to_string(i) + ": bad buffers, size, pitch, cur or frame");
        if ((int64_t)g.frame + g.nsteps > INT32_MAX)
            return fail(SE_ERR_ARG, "grid " + std::to_string(i) + ": frame counter would overflow");
        if (g.mod_count > SE_MAX_MODIFICATIONS || (uint64_t)g.mod_first + g.mod_count > nmods)
            return fail(SE_ERR_ARG, "grid " + std::to_string(i) + ": modification range outside the array");
    }
    DeviceGuard dg(ctx->device);
    se_module* mod = rs->mod.get();
    const int vi = se_rt::default_lds_index(mod);
    const se::LdsVariant& v = mod->lds[vi].first;
    const uint32_t k = (uint32_t)v.hb;
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    char* dev = (char*)ctx->scratch;

    // steps and modification counts left per grid after the clearing frame
    uint32_t steps[SE_NOALLOC_MAX_GRIDS], mcount[SE_NOALLOC_MAX_GRIDS];
    for (uint32_t i = 0; i < n; i++) {
        steps[i] = grids[i].nsteps;
        mcount[i] = grids[i].mod_count;
    }
    // Grids whose next frame is the clearing frame take it first with the fill
    // kernel (falling_sand.glsl:128-131); their modifications are dropped.
    for (uint32_t i = 0; i < n; i++) {
        se_grid_step& g = grids[i];
        if (steps[i] == 0 || g.frame + 1 != 1) continue;
        size_t cells = (size_t)g.pitch * g.h;
        unsigned blocks = (unsigned)std::min<size_t>((cells + 255) / 256, 65536);
        int rc = launch(mod->fill, blocks, st, g.grid[g.cur ^ 1], cells, 0u);
        if (rc) return rc;
        g.cur ^= 1;
        g.frame = 1;
        steps[i]--;
        mcount[i] = 0;
    }

    std::lock_guard<std::mutex> lk(ctx->staging_mu);
    HIPCHECK(hipEventSynchronize(ctx->staging_done));  // the previous call's copy has read the staging buffer
    char* host = (char*)ctx->staging;
    BWorld* recs = (BWorld*)host;
    int* pre_small = (int*)(host + kOffPreSmall);
    int* pre_large = (int*)(host + kOffPreLarge);
    se_modification* hm = (se_modification*)(host + kOffMods);
    if (nmods) std::memcpy(hm, mods, sizeof(se_modification) * nmods);

    // single-tile grids first (all their steps in one work item), then the
    // multi-tile ones, longest first, so that each round's active set is a prefix
    uint32_t order[SE_NOALLOC_MAX_GRIDS];
    uint32_t ns = 0, nl = 0;
    auto single = [&](uint32_t i) { return grids[i].w <= (uint32_t)v.tw && grids[i].h <= (uint32_t)v.th; };
    for (uint32_t i = 0; i < n; i++)
        if (steps[i] && single(i)) order[ns++] = i;
    for (uint32_t i = 0; i < n; i++)
        if (steps[i] && !single(i)) order[ns + nl++] = i;
    std::sort(order + ns, order + ns + nl, [&](uint32_t a, uint32_t b) {  // in place: no temporary buffer
        return steps[a] != steps[b] ? steps[a] > steps[b] : a < b;
    });
    pre_small[0] = 0;
    pre_large[0] = 0;
    for (uint32_t j = 0; j < ns + nl; j++) {
        const uint32_t i = order[j];
        const se_grid_step& g = grids[i];
        BWorld b{};
        b.buf[0] = g.grid[0];
        b.buf[1] = g.grid[1];
        b.W = (int)g.w;
        b.H = (int)g.h;
        b.pitch = (int)g.pitch;
        b.tiles_x = (int)((g.w + v.tw - 1) / v.tw);
        b.ntiles = b.tiles_x * (int)((g.h + v.th - 1) / v.th);
        b.frame0 = g.frame + 1;
        b.steps = (int)steps[i];
        b.cur0 = g.cur;
        int nm = 0;  // the kernels' modification loop ends at a size-0 entry (falling_sand.glsl:139-141)
        while (nm < (int)mcount[i] && hm[g.mod_first + nm].mod_size != 0) nm++;
        b.nmods = nm;
        b.mods = nm ? (const void*)(dev + kOffMods + sizeof(se_modification) * g.mod_first) : nullptr;
        recs[j] = b;
        if (j < ns) pre_small[j + 1] = pre_small[j] + b.ntiles;
        else pre_large[j - ns + 1] = pre_large[j - ns] + b.ntiles;
    }
    if (ns + nl == 0) return SE_OK;
    HIPCHECK(hipMemcpyAsync(dev, host, kUsed, hipMemcpyHostToDevice, st));
    HIPCHECK(hipEventRecord(ctx->staging_done, st));

    const uint32_t* tbl = mod->table;
    hipFunction_t f = mod->batch[vi];
    if (ns) {
        int rc = launch_ex(f, (unsigned)pre_small[ns], (unsigned)v.nt, 0u, st, (const void*)dev,
                           (const int*)(dev + kOffPreSmall), (int)ns, 0, (int)(1 << 30), tbl);
        if (rc) return rc;
    }
    if (nl) {
        const BWorld* lrec = (const BWorld*)dev + ns;
        uint32_t rounds = (steps[order[ns]] + k - 1) / k;
        for (uint32_t r = 0; r < rounds; r++) {
            int act = 0;
            while (act < (int)nl && steps[order[ns + act]] > r * k) act++;
            int rc = launch_ex(f, (unsigned)pre_large[act], (unsigned)v.nt, 0u, st, (const void*)lrec,
                               (const int*)(dev + kOffPreLarge), act, (int)r, (int)k, tbl);
            if (rc) return rc;
        }
    }
    for (uint32_t j = 0; j < ns + nl; j++) {
        const uint32_t i = order[j];
        grids[i].frame += (int32_t)steps[i];
        grids[i].cur ^= j < ns ? 1 : (int32_t)(((steps[i] + k - 1) / k) & 1);
    }
    return SE_OK;
}
