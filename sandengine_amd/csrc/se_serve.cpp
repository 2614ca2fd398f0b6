// Serving loop with continuous batching (se_serve_* in include/sandengine_hip.h).
//
// Requests (a world, a rule set, a step count and a list of modifications at
// given steps) arrive by a seeded Poisson process.  The loop:
//   * admits arrived requests in FIFO order while their state fits the HBM
//     budget: both grids come from one pre-allocated arena (no hipMalloc per
//     request) and the seeded fill is launched asynchronously;
//   * every round advances each admitted world by a quantum (at most the tile
//     halo for worlds spanning several LDS tiles, `quantum_small` for worlds
//     that fit one tile), cut so that a world's modifications always fall on the
//     first step of a round, which is where the kernels apply them (the
//     reference applies queued modifications to the next step only,
//     sandengine-core/src/simulation.rs:203-208);
//   * keeps up to `max_inflight` rounds queued on the device, so rounds overlap
//     with the host building the next one, and requests arriving meanwhile join
//     the next round (continuous batching, not a queue of whole runs);
//   * retires a request when the round holding its last step completes, stamps
//     its latency (arrival -> completion, host clock) and frees its arena space.
// kernel 1 uses the batched LDS kernel (one launch per rule set per round, one
// workgroup per work item); kernel 0 loops the plain kernel world by world
// inside each round.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <thread>

#include "se_internal.hpp"

using se_rt::DeviceGuard;
using se_rt::fail;
using se_rt::launch;
using se_rt::launch_ex;

namespace {

struct Rng {  // SplitMix64 stream
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    uint32_t range(uint32_t lo, uint32_t hi) { return lo + (uint32_t)(next() % (uint64_t)(hi - lo + 1)); }
};

struct Arena {
    char* base = nullptr;
    size_t size = 0, used = 0, peak = 0;
    std::map<size_t, size_t> free_;  // offset -> length
    int init(size_t bytes) {
        HIPCHECK(hipMalloc((void**)&base, bytes));
        size = bytes;
        free_[0] = bytes;
        return SE_OK;
    }
    ~Arena() {
        if (base) (void)hipFree(base);
    }
    size_t alloc(size_t n) {  // first fit; returns SIZE_MAX when full
        n = (n + 255) & ~(size_t)255;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            size_t off = it->first, len = it->second;
            free_.erase(it);
            if (len > n) free_[off + n] = len - n;
            used += n;
            peak = std::max(peak, used);
            return off;
        }
        return SIZE_MAX;
    }
    void release(size_t off, size_t n) {
        n = (n + 255) & ~(size_t)255;
        used -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct FillSpec {  // == se_fillspec in se_kernels.hip
    int n, stone_rows;
    uint32_t stone_id, pad;
    uint32_t ids[16], thr[16];
};

struct BWorld {  // == se_bworld in se_kernels.hip
    uint32_t* buf[2];
    int W, H, pitch, tiles_x, ntiles;
    int frame0, steps, cur0;
    const void* mods;
    int nmods, pad;
};
static_assert(sizeof(BWorld) == 64, "se_bworld layout");

struct Mod {
    uint32_t step;
    se_modification m;
};

struct Request {
    double arrival = 0;
    uint32_t side = 0, steps = 0, rs = 0;
    uint64_t fill_seed = 0;
    std::vector<Mod> mods;  // sorted by step
    // runtime
    int state = 0;  // 0 waiting, 1 admitted, 2 done
    se_world* w = nullptr;
    size_t off = 0, bytes = 0;
    uint32_t scheduled = 0;  // steps enqueued so far
    size_t next_mod = 0;
    double done_t = 0;
    int verified = 0;
    uint64_t hash = 0;
};

struct Slot {  // one in-flight round
    hipEvent_t ev = nullptr;
    bool busy = false;
    std::vector<size_t> finishing;
    std::vector<BWorld> h_recs;
    std::vector<se_modification> h_mods;
    std::vector<int> h_pre;
    BWorld* p_recs = nullptr;  // pinned staging
    se_modification* p_mods = nullptr;
    int* p_pre = nullptr;
    void* d_recs = nullptr;
    void* d_mods = nullptr;
    void* d_pre = nullptr;
    size_t cap_recs = 0, cap_mods = 0, cap_pre = 0;
};

int grow(void** dev, void** pinned, size_t* cap, size_t need, size_t elem) {
    if (need <= *cap) return SE_OK;
    size_t n = std::max(need, *cap * 2 + 64);
    if (*dev) (void)hipFree(*dev);
    if (*pinned) (void)hipHostFree(*pinned);
    *dev = nullptr;
    *pinned = nullptr;
    HIPCHECK(hipMalloc(dev, n * elem));
    HIPCHECK(hipHostMalloc(pinned, n * elem, hipHostMallocDefault));
    *cap = n;
    return SE_OK;
}

uint64_t fnv64(const uint32_t* p, size_t n) {
    uint64_t h = 1469598103934665603ull;
    const unsigned char* b = (const unsigned char*)p;
    for (size_t i = 0; i < n * 4; i++) h = (h ^ b[i]) * 1099511628211ull;
    return h;
}

}  // namespace

struct se_serve {
    se_ctx* ctx = nullptr;
    std::vector<const se_ruleset*> rs;
    std::vector<FillSpec> fills;
    se_serve_config cfg{};
    std::vector<Request> reqs;
};

namespace {

double now_s(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void generate(se_serve* S) {
    Rng rng{S->cfg.seed * 0x2545F4914F6CDD1Dull + 12345};
    const auto& c = S->cfg;
    double t = 0;
    uint64_t i = 0;
    for (;;) {
        t += -std::log(1.0 - rng.uniform()) / c.rate;
        if (t >= c.duration_s) break;
        Request r;
        r.arrival = t;
        double side = std::exp(std::log((double)c.side_min) +
                               rng.uniform() * (std::log((double)c.side_max) - std::log((double)c.side_min)));
        r.side = std::max<uint32_t>(2, (uint32_t)std::lround(side / 2.0) * 2);
        r.rs = (uint32_t)(rng.next() % S->rs.size());
        double st = std::exp(std::log((double)c.steps_min) +
                             rng.uniform() * (std::log((double)c.steps_max) - std::log((double)c.steps_min)));
        r.steps = std::max<uint32_t>(1, (uint32_t)std::lround(st));
        r.fill_seed = c.seed * 1000003ull + i;
        uint32_t nm = c.mods_max ? rng.range(0, c.mods_max) : 0;
        const int nmat = S->rs[r.rs]->nmat;
        for (uint32_t j = 0; j < nm; j++) {
            Mod m{};
            m.step = (uint32_t)(rng.next() % r.steps);
            m.m.position[0] = (int32_t)(rng.next() % r.side);
            m.m.position[1] = (int32_t)(rng.next() % r.side);
            m.m.mod_shape = 0;  // circle (MODSHAPE_CIRCLE)
            if (rng.next() & 1) {  // brush: a random user material
                m.m.mod_size = (int32_t)rng.range(4, 64);
                m.m.mod_matID = nmat > 3 ? (int32_t)rng.range(3, (uint32_t)nmat - 1) : 0;
            } else {  // explosion: empty the disc
                m.m.mod_size = (int32_t)rng.range(8, 128);
                m.m.mod_matID = 0;
            }
            r.mods.push_back(m);
        }
        std::stable_sort(r.mods.begin(), r.mods.end(), [](const Mod& a, const Mod& b) { return a.step < b.step; });
        S->reqs.push_back(std::move(r));
        i++;
    }
}

int admit(se_serve* S, Request& r, Arena& A) {
    const se_ruleset* rs = S->rs[r.rs];
    uint32_t pitch = (r.side + 63u) & ~63u;
    size_t grid = (size_t)pitch * r.side * 4;
    size_t need = 2 * grid + sizeof(se_modification) * SE_MAX_MODIFICATIONS;
    size_t off = A.alloc(need);
    if (off == SIZE_MAX) return 1;  // does not fit now
    auto* w = new se_world();
    w->ctx = S->ctx;
    w->rs = rs;
    w->mod = rs->mod;
    w->W = w->H = r.side;
    w->pitch = pitch;
    w->buf[0] = (uint32_t*)(A.base + off);
    w->buf[1] = (uint32_t*)(A.base + off + grid);
    w->d_mods = A.base + off + 2 * grid;
    w->owns = false;
    w->frame = 1;  // the clearing frame has run; the seeded grid starts at frame 2
    w->cur = 0;
    w->stream = S->ctx->stream;
    r.w = w;
    r.off = off;
    r.bytes = need;
    r.state = 1;
    const FillSpec& fs = S->fills[r.rs];
    size_t cells = (size_t)r.side * r.side;
    unsigned blocks = (unsigned)std::min<size_t>((cells + 255) / 256, 1u << 16);
    return launch(w->mod->fill_spec, blocks, w->stream, w->buf[0], (int)r.side, (int)r.side, (int)pitch,
                  (unsigned long long)r.fill_seed, fs);
}

// Steps the world takes in this round, starting at step r.scheduled.
uint32_t quantum(const se_serve* S, const Request& r, const se::LdsVariant& v) {
    uint32_t q = r.steps - r.scheduled;
    bool single = r.side <= (uint32_t)v.tw && r.side <= (uint32_t)v.th;
    uint32_t cap = S->cfg.kernel == 1 ? (single ? S->cfg.quantum_small : std::min<uint32_t>(S->cfg.quantum, v.hb))
                                      : S->cfg.quantum;
    q = std::min(q, std::max<uint32_t>(1, cap));
    // the next modification step after this round's first step ends the round
    for (size_t j = r.next_mod; j < r.mods.size(); j++)
        if (r.mods[j].step > r.scheduled) {
            q = std::min(q, r.mods[j].step - r.scheduled);
            break;
        }
    return q;
}

int lds_index(const se_module* m) { return se_rt::default_lds_index(m); }

// Builds and enqueues one round over every admitted world with steps left.
int enqueue_round(se_serve* S, Slot& sl, std::vector<size_t>& active, bool* any) {
    *any = false;
    sl.finishing.clear();
    sl.h_recs.clear();
    sl.h_mods.clear();
    hipStream_t st = S->ctx->stream;
    // group by module (rule set)
    std::map<se_module*, std::vector<size_t>> groups;
    for (size_t i : active) {
        Request& r = S->reqs[i];
        if (r.scheduled < r.steps) groups[r.w->mod.get()].push_back(i);
    }
    if (groups.empty()) return SE_OK;
    *any = true;
    if (S->cfg.kernel == 0) {  // plain kernel, world by world
        for (auto& g : groups)
            for (size_t i : g.second) {
                Request& r = S->reqs[i];
                const se::LdsVariant& v = r.w->mod->lds[lds_index(r.w->mod.get())].first;
                uint32_t q = quantum(S, r, v);
                while (r.next_mod < r.mods.size() && r.mods[r.next_mod].step == r.scheduled)
                    r.w->pending.push_back(r.mods[r.next_mod++].m);
                for (uint32_t s = 0; s < q; s++) {
                    int rc = se_rt::plain_step(r.w);
                    if (rc) return rc;
                }
                r.scheduled += q;
                if (r.scheduled == r.steps) sl.finishing.push_back(i);
            }
        HIPCHECK(hipEventRecord(sl.ev, st));
        sl.busy = true;
        return SE_OK;
    }
    // batched LDS kernel: one launch per rule set; records and mods staged per slot
    std::vector<std::pair<se_module*, std::pair<size_t, size_t>>> ranges;  // module -> [rec begin, end)
    sl.h_pre.clear();
    for (auto& g : groups) {
        se_module* mod = g.first;
        const se::LdsVariant& v = mod->lds[lds_index(mod)].first;
        size_t b0 = sl.h_recs.size();
        for (size_t i : g.second) {
            Request& r = S->reqs[i];
            uint32_t q = quantum(S, r, v);
            BWorld b{};
            b.buf[0] = r.w->buf[0];
            b.buf[1] = r.w->buf[1];
            b.W = b.H = (int)r.side;
            b.pitch = (int)r.w->pitch;
            b.tiles_x = (int)((r.side + v.tw - 1) / v.tw);
            b.ntiles = b.tiles_x * (int)((r.side + v.th - 1) / v.th);
            b.frame0 = r.w->frame + 1;
            b.steps = (int)q;
            b.cur0 = r.w->cur;
            size_t m0 = sl.h_mods.size();
            while (r.next_mod < r.mods.size() && r.mods[r.next_mod].step == r.scheduled)
                sl.h_mods.push_back(r.mods[r.next_mod++].m);
            b.nmods = (int)std::min<size_t>(sl.h_mods.size() - m0, SE_MAX_MODIFICATIONS);
            b.mods = (const void*)(uintptr_t)m0;  // index now, device address below
            sl.h_recs.push_back(b);
            // host-side state after this round
            r.w->frame += (int32_t)q;
            r.w->cur ^= 1;
            r.scheduled += q;
            if (r.scheduled == r.steps) sl.finishing.push_back(i);
        }
        ranges.push_back({mod, {b0, sl.h_recs.size()}});
    }
    int rc = grow(&sl.d_recs, (void**)&sl.p_recs, &sl.cap_recs, sl.h_recs.size(), sizeof(BWorld));
    if (!rc) rc = grow(&sl.d_mods, (void**)&sl.p_mods, &sl.cap_mods, std::max<size_t>(1, sl.h_mods.size()),
                       sizeof(se_modification));
    if (!rc) rc = grow(&sl.d_pre, (void**)&sl.p_pre, &sl.cap_pre, sl.h_recs.size() + ranges.size(), sizeof(int));
    if (rc) return rc;
    for (auto& b : sl.h_recs) {
        size_t m0 = (size_t)(uintptr_t)b.mods;
        b.mods = b.nmods ? (const void*)((se_modification*)sl.d_mods + m0) : nullptr;
    }
    // prefix arrays (one per group, packed)
    std::vector<size_t> pre_off;
    for (auto& rg : ranges) {
        pre_off.push_back(sl.h_pre.size());
        sl.h_pre.push_back(0);
        for (size_t j = rg.second.first; j < rg.second.second; j++)
            sl.h_pre.push_back(sl.h_pre.back() + sl.h_recs[j].ntiles);
    }
    std::memcpy(sl.p_recs, sl.h_recs.data(), sl.h_recs.size() * sizeof(BWorld));
    if (!sl.h_mods.empty()) std::memcpy(sl.p_mods, sl.h_mods.data(), sl.h_mods.size() * sizeof(se_modification));
    std::memcpy(sl.p_pre, sl.h_pre.data(), sl.h_pre.size() * sizeof(int));
    HIPCHECK(hipMemcpyAsync(sl.d_recs, sl.p_recs, sl.h_recs.size() * sizeof(BWorld), hipMemcpyHostToDevice, st));
    if (!sl.h_mods.empty())
        HIPCHECK(hipMemcpyAsync(sl.d_mods, sl.p_mods, sl.h_mods.size() * sizeof(se_modification),
                                hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(sl.d_pre, sl.p_pre, sl.h_pre.size() * sizeof(int), hipMemcpyHostToDevice, st));
    for (size_t gi = 0; gi < ranges.size(); gi++) {
        se_module* mod = ranges[gi].first;
        int vi = lds_index(mod);
        const se::LdsVariant& v = mod->lds[vi].first;
        size_t b0 = ranges[gi].second.first, n = ranges[gi].second.second - b0;
        // one workgroup per work item (measured faster than persistent workgroups:
        // 2.51 against 2.15 Tcell/s processed under overload)
        const unsigned items = (unsigned)sl.h_pre[pre_off[gi] + n];
        rc = launch_ex(mod->batch[vi], items ? items : 1u, (unsigned)v.nt, 0u, st,
                       (const void*)((BWorld*)sl.d_recs + b0), (const int*)((int*)sl.d_pre + pre_off[gi]), (int)n, 0,
                       (int)(1 << 30), (const uint32_t*)mod->table);
        if (rc) return rc;
    }
    HIPCHECK(hipEventRecord(sl.ev, st));
    sl.busy = true;
    return SE_OK;
}

int retire(se_serve* S, Slot& sl, Arena& A, std::chrono::steady_clock::time_point t0, uint64_t* done,
           double* cells_done) {
    double t = now_s(t0);
    for (size_t i : sl.finishing) {
        Request& r = S->reqs[i];
        r.done_t = t;
        r.state = 2;
        if (S->cfg.verify_every && i % S->cfg.verify_every == 0) {
            std::vector<uint32_t> g((size_t)r.side * r.side);
            HIPCHECK(hipMemcpy2D(g.data(), (size_t)r.side * 4, r.w->buf[r.w->cur], (size_t)r.w->pitch * 4,
                                 (size_t)r.side * 4, r.side, hipMemcpyDeviceToHost));
            r.hash = fnv64(g.data(), g.size());
            r.verified = 1;
        }
        *cells_done += (double)r.side * r.side * r.steps;
        delete r.w;  // buffers belong to the arena
        r.w = nullptr;
        A.release(r.off, r.bytes);
        (*done)++;
    }
    sl.finishing.clear();
    sl.busy = false;
    return SE_OK;
}

}  // namespace

extern "C" {

int se_serve_create(se_ctx* ctx, const se_ruleset* const* rulesets, int n_rulesets, const se_serve_config* cfg,
                    se_serve** out) {
    if (!ctx || !rulesets || n_rulesets < 1 || n_rulesets > 16 || !cfg || !out) return fail(SE_ERR_ARG, "bad arguments");
    if (!(cfg->rate > 0) || !(cfg->duration_s > 0) || cfg->side_min < 2 || cfg->side_max < cfg->side_min ||
        cfg->side_max > 16384 || cfg->steps_min < 1 || cfg->steps_max < cfg->steps_min || cfg->mods_max > 256 ||
        cfg->hbm_budget < (1u << 20) || (cfg->kernel != 0 && cfg->kernel != 1))
        return fail(SE_ERR_ARG, "bad serving configuration");
    auto* S = new se_serve();
    S->ctx = ctx;
    S->cfg = *cfg;
    if (S->cfg.quantum == 0) S->cfg.quantum = 8;
    if (S->cfg.quantum_small == 0) S->cfg.quantum_small = 64;
    if (S->cfg.max_inflight < 1) S->cfg.max_inflight = 2;
    if (S->cfg.max_inflight > 8) S->cfg.max_inflight = 8;
    for (int i = 0; i < n_rulesets; i++) {
        if (!rulesets[i] || rulesets[i]->ctx != ctx) {
            delete S;
            return fail(SE_ERR_ARG, "rule set missing or from another context");
        }
        S->rs.push_back(rulesets[i]);
        FillSpec f{};
        f.n = 1;
        f.ids[0] = 0;
        S->fills.push_back(f);
    }
    generate(S);
    *out = S;
    return SE_OK;
}

void se_serve_destroy(se_serve* S) {
    if (!S) return;
    for (auto& r : S->reqs) delete r.w;
    delete S;
}

int se_serve_set_fill(se_serve* S, int rs, const uint32_t* ids, const uint32_t* thresholds, int n) {
    if (!S || rs < 0 || rs >= (int)S->rs.size() || !ids || !thresholds || n < 1 || n > 16)
        return fail(SE_ERR_ARG, "bad fill");
    FillSpec f{};
    f.n = n;
    for (int i = 0; i < n; i++) {
        f.ids[i] = ids[i];
        f.thr[i] = thresholds[i];
    }
    S->fills[rs] = f;
    return SE_OK;
}

uint64_t se_serve_request_count(const se_serve* S) { return S ? S->reqs.size() : 0; }

uint64_t se_fnv1a64(const void* data, size_t bytes) {
    uint64_t h = 1469598103934665603ull;
    const unsigned char* b = (const unsigned char*)data;
    for (size_t i = 0; i < bytes; i++) h = (h ^ b[i]) * 1099511628211ull;
    return h;
}

int se_serve_request_info(const se_serve* S, uint64_t i, se_serve_request* out) {
    if (!S || !out || i >= S->reqs.size()) return fail(SE_ERR_ARG, "bad request index");
    const Request& r = S->reqs[i];
    out->arrival_s = r.arrival;
    out->done_s = r.done_t;
    out->side = r.side;
    out->steps = r.steps;
    out->ruleset = r.rs;
    out->nmods = (uint32_t)r.mods.size();
    out->fill_seed = r.fill_seed;
    out->state = r.state;
    out->verified = r.verified;
    out->hash = r.hash;
    return SE_OK;
}

int se_serve_request_mod(const se_serve* S, uint64_t i, uint32_t j, uint32_t* step, se_modification* m) {
    if (!S || i >= S->reqs.size() || j >= S->reqs[i].mods.size()) return fail(SE_ERR_ARG, "bad index");
    if (step) *step = S->reqs[i].mods[j].step;
    if (m) *m = S->reqs[i].mods[j].m;
    return SE_OK;
}

int se_serve_run(se_serve* S, se_serve_stats* stats) {
    if (!S || !stats) return fail(SE_ERR_ARG, "null argument");
    DeviceGuard dg(S->ctx->device);
    Arena A;
    int rc = A.init(S->cfg.hbm_budget);
    if (rc) return rc;
    std::vector<Slot> slots(S->cfg.max_inflight);
    for (auto& sl : slots) HIPCHECK(hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming));
    std::deque<size_t> queue;  // arrived, waiting for admission
    std::vector<size_t> active;
    std::deque<int> order;  // busy slots in submission order
    size_t next = 0;
    uint64_t done = 0, rounds = 0, rejected = 0;
    double cells_done = 0;
    HIPCHECK(hipStreamSynchronize(S->ctx->stream));
    auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        double t = now_s(t0);
        while (next < S->reqs.size() && S->reqs[next].arrival <= t) queue.push_back(next++);
        // admission: FIFO while the arena has room; a request larger than the whole
        // budget is rejected
        while (!queue.empty()) {
            Request& r = S->reqs[queue.front()];
            uint32_t pitch = (r.side + 63u) & ~63u;
            size_t need = 2 * (size_t)pitch * r.side * 4 + sizeof(se_modification) * SE_MAX_MODIFICATIONS;
            if (need > A.size) {
                r.state = 3;
                rejected++;
                queue.pop_front();
                continue;
            }
            int a = admit(S, r, A);
            if (a == 1) break;
            if (a) return a;
            active.push_back(queue.front());
            queue.pop_front();
        }
        // retire completed rounds
        while (!order.empty()) {
            Slot& sl = slots[order.front()];
            hipError_t q = hipEventQuery(sl.ev);
            if (q == hipErrorNotReady) break;
            if (q != hipSuccess) return fail(SE_ERR_HIP, std::string("round event: ") + hipGetErrorString(q));
            rc = retire(S, sl, A, t0, &done, &cells_done);
            if (rc) return rc;
            order.pop_front();
        }
        active.erase(std::remove_if(active.begin(), active.end(), [&](size_t i) { return S->reqs[i].state != 1; }),
                     active.end());
        // enqueue a round when a slot is free
        bool progressed = false;
        if ((int)order.size() < S->cfg.max_inflight) {
            int free_slot = -1;
            for (int i = 0; i < (int)slots.size(); i++)
                if (!slots[i].busy) { free_slot = i; break; }
            if (free_slot >= 0) {
                bool any = false;
                rc = enqueue_round(S, slots[free_slot], active, &any);
                if (rc) return rc;
                if (any) {
                    order.push_back(free_slot);
                    rounds++;
                    progressed = true;
                }
            }
        }
        bool all_arrived = next >= S->reqs.size();
        if (all_arrived && queue.empty() && order.empty()) {
            bool pending = false;
            for (size_t i : active)
                if (S->reqs[i].scheduled < S->reqs[i].steps) pending = true;
            if (!pending) break;
        }
        if (S->cfg.drain_timeout_s > 0 && all_arrived && now_s(t0) > S->cfg.duration_s + S->cfg.drain_timeout_s)
            break;
        if (!progressed) {
            // Nothing to submit.  With every slot busy, block on the oldest round;
            // otherwise poll briefly so that an arrival can start the next round.
            double wait = next < S->reqs.size() ? S->reqs[next].arrival - now_s(t0) : 1e9;
            if ((int)order.size() >= S->cfg.max_inflight) {
                HIPCHECK(hipEventSynchronize(slots[order.front()].ev));
            } else if (wait > 0) {
                std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.0002)));
            }
        }
    }
    HIPCHECK(hipStreamSynchronize(S->ctx->stream));
    double wall = now_s(t0);
    // statistics over completed requests
    std::vector<double> lat, lat_all;
    double offered = 0, processed = 0;
    for (auto& r : S->reqs) {
        offered += (double)r.side * r.side * r.steps;
        processed += (double)r.side * r.side * r.scheduled;  // every enqueued round has completed here
        if (r.state == 2) lat.push_back((r.done_t - r.arrival) * 1e3);
        // censored view: an arrived request not done by the end counts as done at the end
        if (r.arrival <= wall && r.state != 3) lat_all.push_back(((r.state == 2 ? r.done_t : wall) - r.arrival) * 1e3);
    }
    std::sort(lat_all.begin(), lat_all.end());
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double p) {
        if (lat.empty()) return 0.0;
        size_t k = (size_t)std::min<double>(lat.size() - 1, std::floor(p * (lat.size() - 1) + 0.5));
        return lat[k];
    };
    std::memset(stats, 0, sizeof *stats);
    stats->requests = S->reqs.size();
    stats->completed = done;
    stats->rejected = rejected;
    stats->unfinished = S->reqs.size() - done - rejected;
    stats->p50_ms = pct(0.50);
    stats->p99_ms = pct(0.99);
    stats->max_ms = lat.empty() ? 0 : lat.back();
    double sum = 0;
    for (double x : lat) sum += x;
    stats->mean_ms = lat.empty() ? 0 : sum / lat.size();
    stats->wall_s = wall;
    stats->cell_updates = cells_done;
    stats->gcups = cells_done / wall / 1e9;
    stats->processed_gcups = processed / wall / 1e9;
    auto pct_all = [&](double p) {
        if (lat_all.empty()) return 0.0;
        size_t k = (size_t)std::min<double>(lat_all.size() - 1, std::floor(p * (lat_all.size() - 1) + 0.5));
        return lat_all[k];
    };
    stats->p50_ms_censored = pct_all(0.50);
    stats->p99_ms_censored = pct_all(0.99);
    stats->offered_gcups = offered / S->cfg.duration_s / 1e9;
    stats->peak_bytes = A.peak;
    stats->rounds = rounds;
    for (auto& sl : slots) {
        (void)hipEventDestroy(sl.ev);
        if (sl.d_recs) (void)hipFree(sl.d_recs);
        if (sl.d_mods) (void)hipFree(sl.d_mods);
        if (sl.d_pre) (void)hipFree(sl.d_pre);
        if (sl.p_recs) (void)hipHostFree(sl.p_recs);
        if (sl.p_mods) (void)hipHostFree(sl.p_mods);
        if (sl.p_pre) (void)hipHostFree(sl.p_pre);
    }
    for (auto& r : S->reqs)
        if (r.w) {
            delete r.w;
            r.w = nullptr;
        }
    return SE_OK;
}

}  // extern "C"
