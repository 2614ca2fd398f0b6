// Serving loop with continuous batching (se_serve_* in include/sandengine_hip.h).
//
// Requests (a world, a rule set, a step count and a list of modifications at
// given steps) arrive by a seeded Poisson process.  The loop:
//   * admits arrived requests in FIFO order while their state fits the HBM
//     budget: both grids come from one pre-allocated arena (no hipMalloc per
//     request) and the seeded fill is launched asynchronously;
//   * with a latency objective (slo_ms), admits a request only when the round
//     model predicts that it and every admitted request finish within their
//     deadlines, lets it wait while it still could, and sheds it once it cannot
//     (the round model: a fixed part plus the round's cell-updates over a
//     capacity, both learned from the rounds' GPU time);
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
#include <condition_variable>
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
    double pred_ms = 0;  // latency predicted at admission (deadline admission)
};

struct Slot {  // one in-flight round
    hipEvent_t ev0 = nullptr, ev = nullptr;  // round start and end (timing events)
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
    // final grids of verified requests, copied back with the round
    uint32_t* p_ver = nullptr;
    size_t cap_ver = 0;
    std::vector<std::pair<size_t, size_t>> ver;  // (request, cell offset in p_ver)
    hipEvent_t ev_ver = nullptr;  // after the copies into p_ver (on the copy stream)
    bool has_ver = false;
    int hashing = 0;     // grids of p_ver not yet hashed (guarded by the hasher's mutex)
    double cells = 0;    // the round's work in the round model's unit (world_work)
    Slot() = default;
    Slot(const Slot&) = delete;
    Slot& operator=(const Slot&) = delete;
    ~Slot() {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev) (void)hipEventDestroy(ev);
        if (ev_ver) (void)hipEventDestroy(ev_ver);
        if (d_recs) (void)hipFree(d_recs);
        if (d_mods) (void)hipFree(d_mods);
        if (d_pre) (void)hipFree(d_pre);
        if (p_recs) (void)hipHostFree(p_recs);
        if (p_mods) (void)hipHostFree(p_mods);
        if (p_pre) (void)hipHostFree(p_pre);
        if (p_ver) (void)hipHostFree(p_ver);
    }
};

// Round time = fixed_ms + cells / cap, fitted to the rounds' GPU time by least
// squares with exponential forgetting (each round's weight decays by 0.98 per
// later round).  Until the rounds' sizes spread enough to separate the two
// terms, and whenever the fit leaves its plausible range, the fixed part stays
// and the capacity follows the latest rounds alone.
struct RoundModel {
    double fixed_ms = 0.05;
    double cap = 1.5e12;  // cell-updates per second
    double s0 = 0, s1 = 0, s2 = 0, sy = 0, sxy = 0;  // weighted sums over (x = Gcell, y = ms)
    double predict_ms(double cells) const { return fixed_ms + cells / cap * 1e3; }
    void observe(double cells, double ms) {
        if (!(ms > 0) || !(cells >= 0)) return;
        const double x = cells * 1e-9, lam = 0.98;
        s0 = lam * s0 + 1;
        s1 = lam * s1 + x;
        s2 = lam * s2 + x * x;
        sy = lam * sy + ms;
        sxy = lam * sxy + x * ms;
        const double det = s0 * s2 - s1 * s1;
        if (s0 > 8 && det > 0.05 * s0 * s2) {
            const double b = (s0 * sxy - s1 * sy) / det, a = (sy - b * s1) / s0;  // ms per Gcell, ms
            if (b > 0.05 && b < 100 && a > 0.005 && a < 1.0) {
                fixed_ms = a;
                cap = 1e12 / b;
                return;
            }
        }
        if (ms > 2 * fixed_ms && cells > 0)
            cap = std::min(2e13, std::max(1e10, 0.7 * cap + 0.3 * cells / ((ms - fixed_ms) * 1e-3)));
    }
};

// Deadline admission plans for this share of the latency objective (the rest
// absorbs host polling and the model's error).
constexpr double kSloHeadroom = 0.85;

// Stream synchronisation at scope exit: rounds still on the device finish
// before the slots' buffers and the arena are freed (early error returns too).
struct StreamDrain {
    hipStream_t s;
    ~StreamDrain() { (void)hipStreamSynchronize(s); }
};

// The stream that copies verified grids back, beside the rounds (default
// priority); drained and destroyed at scope exit.
struct CopyStream {
    hipStream_t s = nullptr;
    ~CopyStream() {
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    }
};

uint64_t fnv64(const uint32_t* p, size_t n);

// Hashes verified requests' final grids on a thread of its own, so that the
// serving loop does not stall on them.  A slot's staging buffer is rewritten
// only after its previous grids are hashed (wait_idle).
struct Hasher {
    struct Job {
        uint64_t* hash;
        int* verified;
        const uint32_t* cells;
        size_t n;
        Slot* slot;
    };
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Job> q;
    bool stop = false;
    std::thread th;
    Hasher() : th([this] { loop(); }) {}
    ~Hasher() {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv.notify_all();
        th.join();
    }
    void push(const Job& j) {
        {
            std::lock_guard<std::mutex> lk(mu);
            j.slot->hashing++;
            q.push_back(j);
        }
        cv.notify_all();
    }
    void wait_idle(Slot& sl) {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return sl.hashing == 0; });
    }
    void loop() {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            cv.wait(lk, [&] { return stop || !q.empty(); });
            if (q.empty()) return;  // stop requested and nothing left
            Job j = q.front();
            q.pop_front();
            lk.unlock();
            uint64_t h = fnv64(j.cells, j.n);
            lk.lock();
            *j.hash = h;
            *j.verified = 1;
            j.slot->hashing--;
            cv.notify_all();
        }
    }
};

int grow(void** dev, void** pinned, size_t* cap, size_t need, size_t elem) {
    if (need <= *cap) return SE_OK;
    size_t n = std::max(need, *cap * 2 + 64);
    if (*dev) (void)hipFree(*dev);
    if (*pinned) (void)hipHostFree(*pinned);
    *dev = nullptr;
    *pinned = nullptr;
    *cap = 0;
    HIPCHECK(hipMalloc(dev, n * elem));
    HIPCHECK(hipHostMalloc(pinned, n * elem, hipHostMallocDefault));
    *cap = n;
    return SE_OK;
}

int grow_pinned(uint32_t** pinned, size_t* cap, size_t cells) {
    if (cells <= *cap) return SE_OK;
    size_t n = std::max(cells, *cap * 2);
    if (*pinned) (void)hipHostFree(*pinned);
    *pinned = nullptr;
    *cap = 0;
    HIPCHECK(hipHostMalloc((void**)pinned, n * sizeof(uint32_t), hipHostMallocDefault));
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
    bool ran = false;
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

int lds_index(const se_module* m) { return se_rt::default_lds_index(m); }

// Steps per round of a request's world before cuts at modification steps: at
// most the tile halo for worlds spanning several LDS tiles, quantum_small for
// worlds inside one tile (batched kernel); `quantum` for the plain kernel.
uint32_t quantum_cap(const se_serve* S, const Request& r) {
    if (S->cfg.kernel != 1) return std::max<uint32_t>(1, S->cfg.quantum);
    const se_module* m = S->rs[r.rs]->mod.get();
    const se::LdsVariant& v = m->lds[lds_index(m)].first;
    bool single = r.side <= (uint32_t)v.tw && r.side <= (uint32_t)v.th;
    return std::max<uint32_t>(1, single ? S->cfg.quantum_small : std::min<uint32_t>(S->cfg.quantum, v.hb));
}

// Steps the world takes in this round, starting at step r.scheduled.
uint32_t quantum(const se_serve* S, const Request& r) {
    uint32_t q = std::min(r.steps - r.scheduled, quantum_cap(S, r));
    // the next modification step after this round's first step ends the round
    for (size_t j = r.next_mod; j < r.mods.size(); j++)
        if (r.mods[j].step > r.scheduled) {
            q = std::min(q, r.mods[j].step - r.scheduled);
            break;
        }
    return q;
}

// Rounds a request still needs: its remaining steps over the quantum, plus one
// for every later modification step (each may cut a round short).
double rounds_left(const se_serve* S, const Request& r) {
    uint32_t rem = r.steps - r.scheduled, q = quantum_cap(S, r);
    double n = (double)((rem + q - 1) / q);
    uint32_t last = r.scheduled;
    for (size_t j = r.next_mod; j < r.mods.size(); j++)
        if (r.mods[j].step > last) {
            n += 1;
            last = r.mods[j].step;
        }
    return n;
}

// Work a world adds to a round of q steps, in the round model's unit: the
// batched kernel's cost follows its LDS tiles (a partly filled edge tile costs
// a full one), so the cells of its tiles times q; the plain kernel's follows
// the cells.
double world_work(const se_serve* S, const Request& r, uint32_t q) {
    if (S->cfg.kernel != 1) return (double)r.side * r.side * q;
    const se_module* m = S->rs[r.rs]->mod.get();
    const se::LdsVariant& v = m->lds[lds_index(m)].first;
    const double tx = (double)((r.side + v.tw - 1) / v.tw), ty = (double)((r.side + v.th - 1) / v.th);
    return tx * ty * v.tw * v.th * q;
}

// The request's work in each of its rounds.
double round_cells(const se_serve* S, const Request& r) {
    return world_work(S, r, std::min(r.steps - r.scheduled, quantum_cap(S, r)));
}

// After the round: final grids of verified requests finishing in it are copied
// to pinned host memory on the copy stream, which waits for the round's end
// event; the next rounds do not wait for the copies.  The requests' arena space
// is released once the copies are done too (hashed at retirement).
int stage_verify(se_serve* S, Slot& sl, Hasher& hs, hipStream_t vst) {
    sl.ver.clear();
    sl.has_ver = false;
    if (!S->cfg.verify_every) return SE_OK;
    size_t need = 0;
    for (size_t i : sl.finishing)
        if (i % S->cfg.verify_every == 0) need += (size_t)S->reqs[i].side * S->reqs[i].side;
    if (!need) return SE_OK;
    hs.wait_idle(sl);  // the staging buffer still holds grids being hashed
    int rc = grow_pinned(&sl.p_ver, &sl.cap_ver, need);
    if (rc) return rc;
    HIPCHECK(hipStreamWaitEvent(vst, sl.ev, 0));
    size_t off = 0;
    for (size_t i : sl.finishing) {
        if (i % S->cfg.verify_every) continue;
        Request& r = S->reqs[i];
        HIPCHECK(hipMemcpy2DAsync(sl.p_ver + off, (size_t)r.side * 4, r.w->buf[r.w->cur], (size_t)r.w->pitch * 4,
                                  (size_t)r.side * 4, r.side, hipMemcpyDeviceToHost, vst));
        sl.ver.push_back({i, off});
        off += (size_t)r.side * r.side;
    }
    HIPCHECK(hipEventRecord(sl.ev_ver, vst));
    sl.has_ver = true;
    return SE_OK;
}

// Builds and enqueues one round over every admitted world with steps left.
int enqueue_round(se_serve* S, Slot& sl, Hasher& hs, hipStream_t vst, std::vector<size_t>& active, bool* any) {
    *any = false;
    sl.finishing.clear();
    sl.h_recs.clear();
    sl.h_mods.clear();
    sl.cells = 0;
    hipStream_t st = S->ctx->stream;
    // group by module (rule set)
    std::map<se_module*, std::vector<size_t>> groups;
    for (size_t i : active) {
        Request& r = S->reqs[i];
        if (r.scheduled < r.steps) groups[r.w->mod.get()].push_back(i);
    }
    if (groups.empty()) return SE_OK;
    *any = true;
    HIPCHECK(hipEventRecord(sl.ev0, st));
    if (S->cfg.kernel == 0) {  // plain kernel, world by world
        for (auto& g : groups)
            for (size_t i : g.second) {
                Request& r = S->reqs[i];
                uint32_t q = quantum(S, r);
                while (r.next_mod < r.mods.size() && r.mods[r.next_mod].step == r.scheduled)
                    r.w->pending.push_back(r.mods[r.next_mod++].m);
                for (uint32_t s = 0; s < q; s++) {
                    int rc = se_rt::plain_step(r.w);
                    if (rc) return rc;
                }
                sl.cells += world_work(S, r, q);
                r.scheduled += q;
                if (r.scheduled == r.steps) sl.finishing.push_back(i);
            }
        HIPCHECK(hipEventRecord(sl.ev, st));
        int rc = stage_verify(S, sl, hs, vst);
        if (rc) return rc;
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
            uint32_t q = quantum(S, r);
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
            sl.cells += world_work(S, r, q);
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
    rc = stage_verify(S, sl, hs, vst);
    if (rc) return rc;
    sl.busy = true;
    return SE_OK;
}

int retire(se_serve* S, Slot& sl, Hasher& hs, Arena& A, RoundModel& M, std::chrono::steady_clock::time_point t0,
           uint64_t* done, double* cells_done) {
    double t = now_s(t0);
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, sl.ev0, sl.ev));
    M.observe(sl.cells, ms);
    for (auto& v : sl.ver) {
        Request& r = S->reqs[v.first];
        hs.push({&r.hash, &r.verified, sl.p_ver + v.second, (size_t)r.side * r.side, &sl});
    }
    sl.ver.clear();
    for (size_t i : sl.finishing) {
        Request& r = S->reqs[i];
        r.done_t = t;
        r.state = 2;
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

size_t request_bytes(const Request& r) {
    uint32_t pitch = (r.side + 63u) & ~63u;
    return 2 * (size_t)pitch * r.side * 4 + sizeof(se_modification) * SE_MAX_MODIFICATIONS;
}

}  // namespace

extern "C" {

int se_serve_create(se_ctx* ctx, const se_ruleset* const* rulesets, int n_rulesets, const se_serve_config* cfg,
                    se_serve** out) {
    if (!ctx || !rulesets || n_rulesets < 1 || n_rulesets > 16 || !cfg || !out) return fail(SE_ERR_ARG, "bad arguments");
    if (!(cfg->rate > 0) || !(cfg->duration_s > 0) || cfg->side_min < 2 || cfg->side_max < cfg->side_min ||
        cfg->side_max > 16384 || cfg->steps_min < 1 || cfg->steps_max < cfg->steps_min || cfg->mods_max > 256 ||
        cfg->steps_max > (uint32_t)INT32_MAX - 1 || cfg->hbm_budget < (1u << 20) ||
        (cfg->kernel != 0 && cfg->kernel != 1) || !(cfg->slo_ms >= 0) || !(cfg->capacity_gcups >= 0))
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
    if (S->ran) return fail(SE_ERR_STATE, "a serving run has already taken place on this object");
    S->ran = true;
    DeviceGuard dg(S->ctx->device);
    const double slo = S->cfg.slo_ms;
    Arena A;
    int rc = A.init(S->cfg.hbm_budget);
    if (rc) return rc;
    std::vector<Slot> slots(S->cfg.max_inflight);
    StreamDrain drain{S->ctx->stream};  // destroyed before the slots and the arena
    CopyStream vs;                      // drained before the slots' pinned buffers go
    Hasher hasher;                      // joined before the streams drain and the slots go
    HIPCHECK(hipStreamCreateWithFlags(&vs.s, hipStreamNonBlocking));
    for (auto& sl : slots) {
        HIPCHECK(hipEventCreate(&sl.ev0));
        HIPCHECK(hipEventCreate(&sl.ev));
        HIPCHECK(hipEventCreateWithFlags(&sl.ev_ver, hipEventDisableTiming));
    }
    RoundModel M;
    if (S->cfg.capacity_gcups > 0) M.cap = S->cfg.capacity_gcups * 1e9;
    std::deque<size_t> queue;  // arrived, waiting for admission
    std::vector<size_t> active;
    std::deque<int> order;  // busy slots in submission order
    size_t next = 0;
    uint64_t done = 0, rounds = 0, rejected = 0, shed = 0;
    double cells_done = 0;
    HIPCHECK(hipStreamSynchronize(S->ctx->stream));
    auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        double t = now_s(t0);
        while (next < S->reqs.size() && S->reqs[next].arrival <= t) queue.push_back(next++);
        if (slo <= 0) {
            // admission: FIFO while the arena has room; a request larger than the
            // whole budget is rejected
            while (!queue.empty()) {
                Request& r = S->reqs[queue.front()];
                if (request_bytes(r) > A.size) {
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
        } else if (!queue.empty()) {
            // deadline admission.  Rounds already on the device run first (pend);
            // a request with n rounds left then finishes after n rounds of
            // T(U) = fixed + U / capacity, U the cell-updates of one round.
            // Admitting r raises U by u_r; it must keep every admitted request i
            // within n_i * T(U) <= deadline_i - now - pend, and r itself too.
            double now = now_s(t0) * 1e3;
            const double budget_ms = slo * kSloHeadroom;
            double U = 0, pend = 0, tmax = 1e300;
            for (int k : order) pend += M.predict_ms(slots[k].cells);
            for (size_t i : active) {
                const Request& r = S->reqs[i];
                if (r.state != 1 || r.scheduled >= r.steps) continue;
                U += round_cells(S, r);
                tmax = std::min(tmax, (r.arrival * 1e3 + budget_ms - now - pend) / rounds_left(S, r));
            }
            for (auto it = queue.begin(); it != queue.end();) {
                Request& r = S->reqs[*it];
                if (request_bytes(r) > A.size) {
                    r.state = 3;
                    rejected++;
                    it = queue.erase(it);
                    continue;
                }
                double n = rounds_left(S, r), u = round_cells(S, r);
                double slack = r.arrival * 1e3 + budget_ms - now - pend;
                if (n * M.predict_ms(u) > slack) {  // cannot make it even alone
                    r.state = 4;
                    shed++;
                    it = queue.erase(it);
                    continue;
                }
                double tn = M.predict_ms(U + u);
                if (n * tn > slack || tn > tmax) {  // would be late, or make another late: wait
                    ++it;
                    continue;
                }
                int a = admit(S, r, A);
                if (a == 1) {  // no room in the arena now
                    ++it;
                    continue;
                }
                if (a) return a;
                active.push_back(*it);
                r.pred_ms = now + pend + n * tn - r.arrival * 1e3;
                U += u;
                tmax = std::min(tmax, slack / n);
                it = queue.erase(it);
            }
        }
        // retire completed rounds
        while (!order.empty()) {
            Slot& sl = slots[order.front()];
            hipError_t q = hipEventQuery(sl.ev);
            if (q == hipSuccess && sl.has_ver) q = hipEventQuery(sl.ev_ver);
            if (q == hipErrorNotReady) break;
            if (q != hipSuccess) return fail(SE_ERR_HIP, std::string("round event: ") + hipGetErrorString(q));
            rc = retire(S, sl, hasher, A, M, t0, &done, &cells_done);
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
                rc = enqueue_round(S, slots[free_slot], hasher, vs.s, active, &any);
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
                const Slot& sl = slots[order.front()];
                HIPCHECK(hipEventSynchronize(sl.has_ver ? sl.ev_ver : sl.ev));
            } else if (wait > 0) {
                std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.0002)));
            }
        }
    }
    HIPCHECK(hipStreamSynchronize(S->ctx->stream));
    HIPCHECK(hipStreamSynchronize(vs.s));
    double wall = now_s(t0);
    for (auto& sl : slots) hasher.wait_idle(sl);  // every verified grid hashed
    // statistics over completed requests
    std::vector<double> lat, lat_all, ratio;
    double offered = 0, processed = 0, shed_cells = 0;
    uint64_t verified = 0, met = 0;
    for (auto& r : S->reqs) {
        double work = (double)r.side * r.side * r.steps;
        offered += work;
        processed += (double)r.side * r.side * r.scheduled;  // every enqueued round has completed here
        if (r.state == 4) shed_cells += work;
        if (r.state == 2) {
            double l = (r.done_t - r.arrival) * 1e3;
            lat.push_back(l);
            if (slo <= 0 || l <= slo) met++;
            if (r.pred_ms > 0) ratio.push_back(l / r.pred_ms);
        }
        verified += (uint64_t)r.verified;
        // censored view: an arrived request not done by the end counts as done at the end
        if (r.arrival <= wall && r.state != 3 && r.state != 4)
            lat_all.push_back(((r.state == 2 ? r.done_t : wall) - r.arrival) * 1e3);
    }
    std::sort(lat_all.begin(), lat_all.end());
    std::sort(lat.begin(), lat.end());
    std::sort(ratio.begin(), ratio.end());
    auto pct = [](const std::vector<double>& v, double p) {
        if (v.empty()) return 0.0;
        size_t k = (size_t)std::min<double>(v.size() - 1, std::floor(p * (v.size() - 1) + 0.5));
        return v[k];
    };
    std::memset(stats, 0, sizeof *stats);
    stats->requests = S->reqs.size();
    stats->completed = done;
    stats->rejected = rejected;
    stats->shed = shed;
    stats->unfinished = S->reqs.size() - done - rejected - shed;
    stats->p50_ms = pct(lat, 0.50);
    stats->p99_ms = pct(lat, 0.99);
    stats->max_ms = lat.empty() ? 0 : lat.back();
    double sum = 0;
    for (double x : lat) sum += x;
    stats->mean_ms = lat.empty() ? 0 : sum / lat.size();
    stats->wall_s = wall;
    stats->cell_updates = cells_done;
    stats->gcups = cells_done / wall / 1e9;
    stats->processed_gcups = processed / wall / 1e9;
    stats->p50_ms_censored = pct(lat_all, 0.50);
    stats->p99_ms_censored = pct(lat_all, 0.99);
    stats->offered_gcups = offered / S->cfg.duration_s / 1e9;
    stats->peak_bytes = A.peak;
    stats->rounds = rounds;
    stats->verified = verified;
    stats->shed_cell_frac = offered > 0 ? shed_cells / offered : 0;
    stats->slo_met_frac = lat.empty() ? 1.0 : (double)met / lat.size();
    stats->capacity_gcups_est = M.cap / 1e9;
    stats->round_fixed_ms = M.fixed_ms;
    stats->pred_ratio_p50 = pct(ratio, 0.50);
    stats->pred_ratio_p99 = pct(ratio, 0.99);
    for (auto& r : S->reqs)
        if (r.w) {
            delete r.w;
            r.w = nullptr;
        }
    return SE_OK;
}

}  // extern "C"
