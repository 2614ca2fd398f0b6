// Batched multi-world execution (se_batch_* in include/sandengine_hip.h).
//
// Many independent worlds advance together.  Worlds are grouped by rule set
// (one JIT module each).  Within a group:
//   * a world that fits one LDS tile keeps its WALL border in LDS, so one work
//     item advances it through all of its steps in a single launch;
//   * larger worlds advance in rounds of k sub-steps (k <= the tile halo): round
//     r covers every world with more than r*k steps; worlds are ordered by step
//     count (longest first) so the active set of a round is a prefix, and one
//     persistent launch per round strides over all of their tiles.
// The single-tile launch and the rounds of every group go to their own streams
// and overlap on the device; the whole submission can be captured in one
// hipGraph.  The reference has no batched mode: each Simulation is one world
// (sandengine-core/src/simulation.rs:97-126); the per-world step is the same.
#include <algorithm>
#include <cstdio>

#include "se_internal.hpp"

using se_rt::DeviceGuard;
using se_rt::fail;
using se_rt::launch_ex;

namespace {

// == se_bworld in se_kernels.hip
struct BWorld {
    uint32_t* buf[2];
    int W, H, pitch, tiles_x, ntiles;
    int frame0;
    int steps;
    int cur0;
    const void* mods;
    int nmods;
    int pad;
};
static_assert(sizeof(BWorld) == 64, "se_bworld layout");

struct DevArray {
    void* p = nullptr;
    ~DevArray() {
        if (p) (void)hipFree(p);
    }
};

// Events, a captured graph and its executable, released on every exit path.
struct RunResources {
    std::vector<hipEvent_t> events;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipStream_t> streams;  // drained first: launches may still read the plans' arrays
    RunResources() = default;
    RunResources(const RunResources&) = delete;
    RunResources& operator=(const RunResources&) = delete;
    ~RunResources() {
        for (auto s : streams)
            if (s) (void)hipStreamSynchronize(s);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto e : events) (void)hipEventDestroy(e);
    }
    int event(hipEvent_t* out, unsigned flags) {
        HIPCHECK(hipEventCreateWithFlags(out, flags));
        events.push_back(*out);
        return SE_OK;
    }
};

}  // namespace

struct se_batch {
    se_ctx* ctx = nullptr;
    std::vector<std::pair<se_world*, uint32_t>> items;
    struct Tuned {
        int vi;           // LDS variant index
        uint32_t k;       // sub-steps per round
        bool persistent;  // launch shape (see enqueue_plan)
    };
    std::map<se_module*, Tuned> tuned;
    hipStream_t streams[8] = {};
    ~se_batch() {
        for (auto s : streams)
            if (s) (void)hipStreamDestroy(s);
    }
};

namespace {

struct Plan {
    se_module* mod = nullptr;
    int vi = 0;
    uint32_t k = 8;
    std::vector<std::pair<se_world*, uint32_t>> small, large;  // (world, steps), steps descending
    std::vector<int> n_active;                                  // per round, large list
    DevArray d_small, d_large, d_pre_small, d_pre_large;
    std::vector<int> pre_small, pre_large;                      // host copies of the tile prefix sums
    bool persistent = false;                                     // launch shape (see enqueue_plan)
};

bool fits_one_tile(const se_world* w, const se::LdsVariant& v) {
    return w->W <= (uint32_t)v.tw && w->H <= (uint32_t)v.th;
}

int upload_list(const std::vector<std::pair<se_world*, uint32_t>>& list, const se::LdsVariant& v, DevArray& dw,
                DevArray& dp, std::vector<int>& pre) {
    std::vector<BWorld> recs;
    pre.assign(1, 0);
    for (auto& it : list) {
        se_world* w = it.first;
        BWorld b;
        b.buf[0] = w->buf[0];
        b.buf[1] = w->buf[1];
        b.W = (int)w->W;
        b.H = (int)w->H;
        b.pitch = (int)w->pitch;
        b.tiles_x = (int)((w->W + v.tw - 1) / v.tw);
        b.ntiles = b.tiles_x * (int)((w->H + v.th - 1) / v.th);
        b.frame0 = w->frame + 1;
        b.steps = (int)it.second;
        b.cur0 = w->cur;
        b.mods = nullptr;
        b.nmods = 0;
        b.pad = 0;
        recs.push_back(b);
        pre.push_back(pre.back() + b.ntiles);
    }
    if (recs.empty()) return SE_OK;
    HIPCHECK(hipMalloc(&dw.p, recs.size() * sizeof(BWorld)));
    HIPCHECK(hipMalloc(&dp.p, pre.size() * sizeof(int)));
    HIPCHECK(hipMemcpy(dw.p, recs.data(), recs.size() * sizeof(BWorld), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(dp.p, pre.data(), pre.size() * sizeof(int), hipMemcpyHostToDevice));
    return SE_OK;
}

int make_plan(se_ctx* ctx, se_module* mod, int vi, uint32_t k, const std::vector<std::pair<se_world*, uint32_t>>& ws,
              Plan& P) {
    const se::LdsVariant& v = mod->lds[vi].first;
    if (k == 0 || k > (uint32_t)v.hb) k = (uint32_t)v.hb;
    P.mod = mod;
    P.vi = vi;
    P.k = k;
    for (auto& it : ws) {
        if (it.second == 0) continue;
        (fits_one_tile(it.first, v) ? P.small : P.large).push_back(it);
    }
    auto by_steps = [](const std::pair<se_world*, uint32_t>& a, const std::pair<se_world*, uint32_t>& b) {
        return a.second > b.second;
    };
    std::stable_sort(P.small.begin(), P.small.end(), by_steps);
    std::stable_sort(P.large.begin(), P.large.end(), by_steps);
    if (!P.large.empty()) {
        uint32_t rounds = (P.large[0].second + k - 1) / k;
        for (uint32_t r = 0; r < rounds; r++) {
            int n = 0;
            while (n < (int)P.large.size() && P.large[n].second > r * k) n++;
            P.n_active.push_back(n);
        }
    }
    int rc = upload_list(P.small, v, P.d_small, P.d_pre_small, P.pre_small);
    if (!rc) rc = upload_list(P.large, v, P.d_large, P.d_pre_large, P.pre_large);
    return rc;
}

// Enqueues a plan's launches: single-tile worlds on `s_small`, rounds on `s_large`.
// Two launch shapes of the same kernel:
//   * one workgroup per work item (the kernel's item loop runs once);
//     workgroups that start as others retire keep the loads and compute of
//     different items overlapped, and the longest-first order makes the
//     dispatch order a longest-first schedule;
//   * persistent: as many workgroups as are resident at once, each striding
//     over the items.
// The autotuner measures both.
int enqueue_plan(Plan& P, hipStream_t s_small, hipStream_t s_large) {
    const se::LdsVariant& v = P.mod->lds[P.vi].first;
    hipFunction_t f = P.mod->batch[P.vi];
    const unsigned shmem = 0;  // the LDS kernels' LDS is static
    const uint32_t* tbl = P.mod->table;
    auto grid = [&](int items) {
        return (unsigned)(P.persistent ? std::min(items, P.mod->lds_grid[P.vi]) : items);
    };
    if (!P.small.empty()) {
        int n = (int)P.small.size();
        int rc = launch_ex(f, grid(P.pre_small.back()), (unsigned)v.nt, shmem, s_small, (const void*)P.d_small.p,
                           (const int*)P.d_pre_small.p, n, 0, (int)(1 << 30), tbl);
        if (rc) return rc;
    }
    for (size_t r = 0; r < P.n_active.size(); r++) {
        int rc = launch_ex(f, grid(P.pre_large[P.n_active[r]]), (unsigned)v.nt, shmem, s_large,
                           (const void*)P.d_large.p, (const int*)P.d_pre_large.p, P.n_active[r], (int)r, (int)P.k,
                           tbl);
        if (rc) return rc;
    }
    return SE_OK;
}

void finish_plan(Plan& P) {
    for (auto& it : P.small) se_rt::sleep_reset(it.first);
    for (auto& it : P.large) se_rt::sleep_reset(it.first);
    for (auto& it : P.small) {
        it.first->frame += (int32_t)it.second;
        it.first->cur ^= 1;
    }
    for (auto& it : P.large) {
        uint32_t rounds = (it.second + P.k - 1) / P.k;
        it.first->frame += (int32_t)it.second;
        it.first->cur ^= (int)(rounds & 1);
    }
}

int ensure_streams(se_batch* b, int n) {
    for (int i = 0; i < n && i < 8; i++)
        if (!b->streams[i]) HIPCHECK(hipStreamCreateWithFlags(&b->streams[i], hipStreamNonBlocking));
    return SE_OK;
}

// Runs the batch: plans are built for every rule-set group, then all launches
// are submitted (optionally through a hipGraph) between two events.
// launch: 0 one workgroup per item, 1 persistent, -1 the tuned choice (else one per item)
int run_batched(se_batch* b, const std::vector<std::pair<se_world*, uint32_t>>& work, int32_t variant, uint32_t k,
                int launch, int graph, float* gpu_ms) {
    se_ctx* ctx = b->ctx;
    std::map<se_module*, std::vector<std::pair<se_world*, uint32_t>>> groups;
    for (auto& it : work) groups[it.first->mod.get()].push_back(it);
    std::vector<std::unique_ptr<Plan>> plans;
    for (auto& g : groups) {
        se_module* mod = g.first;
        int vi = variant;
        uint32_t kk = k;
        int la = launch;
        auto t = b->tuned.find(mod);
        if (vi < 0) {
            if (t != b->tuned.end()) {
                vi = t->second.vi;
                if (kk == 0) kk = t->second.k;
            } else {
                vi = se_rt::default_lds_index(mod);
            }
        }
        if (la < 0) la = t != b->tuned.end() ? (int)t->second.persistent : 0;
        if (vi < 0 || vi >= (int)mod->lds.size()) return fail(SE_ERR_ARG, "LDS variant out of range");
        plans.emplace_back(new Plan());
        plans.back()->persistent = la == 1;
        int rc = make_plan(ctx, mod, vi, kk, g.second, *plans.back());
        if (rc) return rc;
    }
    int rc = ensure_streams(b, 2 * (int)plans.size());
    if (rc) return rc;
    RunResources res;  // declared after `plans`: destroyed (and the stream drained) first
    res.streams.assign(b->streams, b->streams + 8);
    res.streams.push_back(ctx->stream);
    hipEvent_t e0, e1, fork;
    std::vector<hipEvent_t> joins(2 * plans.size());
    rc = res.event(&e0, hipEventDefault);
    if (!rc) rc = res.event(&e1, hipEventDefault);
    if (!rc) rc = res.event(&fork, hipEventDisableTiming);
    for (auto& j : joins)
        if (!rc) rc = res.event(&j, hipEventDisableTiming);
    if (rc) return rc;
    auto submit = [&]() -> int {
        HIPCHECK(hipEventRecord(fork, ctx->stream));
        for (size_t i = 0; i < plans.size(); i++) {
            hipStream_t sa = b->streams[(2 * i) % 8], sb = b->streams[(2 * i + 1) % 8];
            HIPCHECK(hipStreamWaitEvent(sa, fork, 0));
            HIPCHECK(hipStreamWaitEvent(sb, fork, 0));
            int r2 = enqueue_plan(*plans[i], sa, sb);
            if (r2) return r2;
            HIPCHECK(hipEventRecord(joins[2 * i], sa));
            HIPCHECK(hipEventRecord(joins[2 * i + 1], sb));
        }
        for (auto& j : joins) HIPCHECK(hipStreamWaitEvent(ctx->stream, j, 0));
        return SE_OK;
    };
    if (graph) {
        HIPCHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        int r2 = submit();
        hipError_t ce = hipStreamEndCapture(ctx->stream, &res.graph);
        if (r2) return r2;
        if (ce != hipSuccess) return fail(SE_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ce));
        HIPCHECK(hipGraphInstantiate(&res.exec, res.graph, nullptr, nullptr, 0));
    }
    HIPCHECK(hipEventRecord(e0, ctx->stream));
    if (graph) {
        HIPCHECK(hipGraphLaunch(res.exec, ctx->stream));
    } else {
        rc = submit();
        if (rc) return rc;
    }
    HIPCHECK(hipEventRecord(e1, ctx->stream));
    HIPCHECK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
    for (auto& p : plans) finish_plan(*p);
    if (gpu_ms) *gpu_ms = ms;
    return SE_OK;
}

// Worlds whose next step is the clearing frame, or that carry queued
// modifications, take that one step with the plain kernel first.
int pre_steps(std::vector<std::pair<se_world*, uint32_t>>& work) {
    for (auto& it : work) {
        se_world* w = it.first;
        if (it.second == 0) continue;
        if (w->frame + 1 == 1 || !w->pending.empty()) {
            int rc = se_rt::plain_step(w);
            if (rc) return rc;
            it.second--;
        }
    }
    return SE_OK;
}

}  // namespace

extern "C" {

int se_batch_create(se_ctx* ctx, se_batch** out) {
    if (!ctx || !out) return fail(SE_ERR_ARG, "null argument");
    auto* b = new se_batch();
    b->ctx = ctx;
    *out = b;
    return SE_OK;
}

void se_batch_destroy(se_batch* b) {
    if (!b) return;
    DeviceGuard dg(b->ctx->device);
    delete b;
}

int se_batch_add(se_batch* b, se_world* w, uint32_t nsteps) {
    if (!b || !w) return fail(SE_ERR_ARG, "null argument");
    if (w->ctx != b->ctx) return fail(SE_ERR_ARG, "world belongs to another context");
    if ((int64_t)w->frame + nsteps > INT32_MAX) return fail(SE_ERR_ARG, "frame counter would overflow");
    if (w->color) return fail(SE_ERR_STATE, "a world with render outputs steps alone (one step per launch)");
    for (auto& it : b->items)
        if (it.first == w) return fail(SE_ERR_ARG, "world already in the batch");
    b->items.push_back({w, nsteps});
    return SE_OK;
}

int se_batch_clear(se_batch* b) {
    if (!b) return fail(SE_ERR_ARG, "null batch");
    b->items.clear();
    return SE_OK;
}

int se_batch_run(se_batch* b, int mode, int32_t lds_variant, uint32_t k, int graph, float* gpu_ms) {
    if (!b) return fail(SE_ERR_ARG, "null batch");
    DeviceGuard dg(b->ctx->device);
    std::vector<std::pair<se_world*, uint32_t>> work = b->items;
    // every run advances the worlds again: the frame check of se_batch_add holds
    // for the first run only
    for (auto& it : work)
        if ((int64_t)it.first->frame + it.second > INT32_MAX)
            return fail(SE_ERR_ARG, "frame counter would overflow for a world of the batch");
    if (mode == 0) {  // loop the plain kernel world by world, one launch per step
        RunResources res;
        hipEvent_t e0, e1;
        int rc = res.event(&e0, hipEventDefault);
        if (!rc) rc = res.event(&e1, hipEventDefault);
        if (rc) return rc;
        HIPCHECK(hipEventRecord(e0, b->ctx->stream));
        for (auto& it : work)
            for (uint32_t s = 0; s < it.second; s++) {
                rc = se_rt::plain_step(it.first);
                if (rc) return rc;
            }
        HIPCHECK(hipEventRecord(e1, b->ctx->stream));
        HIPCHECK(hipEventSynchronize(e1));
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
        if (gpu_ms) *gpu_ms = ms;
        return SE_OK;
    }
    if (mode < 1 || mode > 3)
        return fail(SE_ERR_ARG, "mode must be 0 (looped plain), 1 (batched, tuned launch), 2 (batched, one "
                                "workgroup per item) or 3 (batched, persistent)");
    int rc = pre_steps(work);
    if (rc) return rc;
    HIPCHECK(hipStreamSynchronize(b->ctx->stream));
    return run_batched(b, work, lds_variant, k, mode == 1 ? -1 : mode - 2, graph, gpu_ms);
}

int se_batch_tuned(const se_batch* b, const se_ruleset* rs, int32_t* variant, uint32_t* k, int32_t* persistent) {
    if (!b || !rs) return fail(SE_ERR_ARG, "null argument");
    auto it = b->tuned.find(rs->mod.get());
    if (it == b->tuned.end()) return fail(SE_ERR_STATE, "this rule set's group has not been tuned");
    if (variant) *variant = it->second.vi;
    if (k) *k = it->second.k;
    if (persistent) *persistent = it->second.persistent ? 1 : 0;
    return SE_OK;
}

int se_batch_autotune(se_batch* b, uint32_t rounds, int32_t* best_variant, uint32_t* best_k,
                      int32_t* best_persistent, float* best_ns_per_kcell) {
    if (!b || rounds == 0) return fail(SE_ERR_ARG, "bad arguments");
    DeviceGuard dg(b->ctx->device);
    int rc = SE_OK;
    // trial runs must leave the worlds untouched: worlds whose next step needs
    // the plain kernel (clearing frame, queued modifications) sit out
    std::map<se_module*, std::vector<std::pair<se_world*, uint32_t>>> groups;
    for (auto& it : b->items)
        if (it.second > 0 && it.first->frame + 1 != 1 && it.first->pending.empty())
            groups[it.first->mod.get()].push_back(it);
    for (auto& g : groups) {
        se_module* mod = g.first;
        // snapshot the group's grids and frames
        std::vector<std::pair<void*, size_t>> snaps;
        std::vector<std::pair<int32_t, int>> state;
        for (auto& it : g.second) {
            se_world* w = it.first;
            size_t bytes = (size_t)w->pitch * w->H * 4;
            void* p = nullptr;
            if (hipMalloc(&p, bytes) != hipSuccess) {
                for (auto& s : snaps) (void)hipFree(s.first);
                return fail(SE_ERR_OOM, "autotune snapshot allocation failed");
            }
            HIPCHECK(hipMemcpyAsync(p, w->buf[w->cur], bytes, hipMemcpyDeviceToDevice, b->ctx->stream));
            snaps.push_back({p, bytes});
            state.push_back({w->frame, w->cur});
        }
        HIPCHECK(hipStreamSynchronize(b->ctx->stream));
        double best = 1e300;
        int bv = 0;
        uint32_t bk = 0;
        bool bp = false;
        for (int vi = 0; vi < (int)mod->lds.size(); vi++) {
            const int hb = mod->lds[vi].first.hb;
            for (uint32_t kk : {(uint32_t)hb, (uint32_t)std::max(2, hb / 2)})
            for (int la = 0; la < 2; la++) {
                std::vector<std::pair<se_world*, uint32_t>> trial;
                double cells = 0;
                for (auto& it : g.second) {
                    uint32_t st = std::min<uint32_t>(it.second, rounds * kk);
                    trial.push_back({it.first, st});
                    cells += (double)it.first->W * it.first->H * st;
                }
                float ms = 0;
                rc = run_batched(b, trial, vi, kk, la, 0, &ms);
                // restore
                for (size_t i = 0; i < g.second.size(); i++) {
                    se_world* w = g.second[i].first;
                    w->frame = state[i].first;
                    w->cur = state[i].second;
                    HIPCHECK(hipMemcpyAsync(w->buf[w->cur], snaps[i].first, snaps[i].second,
                                            hipMemcpyDeviceToDevice, b->ctx->stream));
                }
                HIPCHECK(hipStreamSynchronize(b->ctx->stream));
                if (rc) {
                    for (auto& s : snaps) (void)hipFree(s.first);
                    return rc;
                }
                double score = cells > 0 ? ms * 1e6 / (cells / 1e3) : 0;  // ns per 1000 cell-updates
                if (score < best) {
                    best = score;
                    bv = vi;
                    bk = kk;
                    bp = la == 1;
                }
            }
        }
        for (auto& s : snaps) (void)hipFree(s.first);
        b->tuned[mod] = {bv, bk, bp};
        if (best_variant) *best_variant = bv;
        if (best_k) *best_k = bk;
        if (best_persistent) *best_persistent = bp ? 1 : 0;
        if (best_ns_per_kcell) *best_ns_per_kcell = (float)best;
    }
    return SE_OK;
}

}  // extern "C"
