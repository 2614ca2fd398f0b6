// Device runtime and C ABI (include/sandengine_hip.h).
//
// A se_world is the MI355X replacement for sandengine-core's Simulation
// (sandengine-core/src/simulation.rs:97-253): the RGBA32F ping-pong textures
// become two u32 grids in HBM with a 256-byte-aligned row pitch, the
// 256-entry modification UBO becomes a small device array uploaded only for
// steps that carry modifications, and the compute-shader dispatch becomes a
// hiprtc-compiled kernel launch on the world's stream.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "se_internal.hpp"

namespace {
thread_local std::string g_err;
}  // namespace

namespace se_rt {
int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
}  // namespace se_rt

namespace {
using se_rt::fail;
using se_rt::launch;
using se_rt::launch_ex;
using se_rt::DeviceGuard;

char* dupstr(const std::string& s) {
    char* p = (char*)std::malloc(s.size() + 1);
    if (p) std::memcpy(p, s.c_str(), s.size() + 1);
    return p;
}

constexpr int kDefaultTile = 112, kDefaultHalo = 8;  // default single-world tile (see load_module)
constexpr unsigned kMaxLdsBytes = 160 * 1024;        // LDS per workgroup on gfx950

const std::vector<se::LdsVariant>& lds_variants() {
    static const std::vector<se::LdsVariant> v = {
        {112, 112, 8, 256}, {240, 112, 8, 512}, {224, 112, 8, 512}, {96, 112, 8, 256}, {224, 224, 16, 512},
        {96, 96, 8, 256},   {112, 112, 8, 512}, {104, 104, 12, 256}};
    return v;
}

// hiprtc options: gfx950, IEEE f32 with no contraction, packed-f32 forms off, and
// no merging of byte LDS accesses into misaligned wider ones (slow on LDS)
// Static LDS of an LDS kernel variant: tile (rounded to 16 bytes), the table,
// a 16-byte slot for the tile flag word.
unsigned lds_bytes_of(const se::LdsVariant& v, uint64_t table_words) {
    return (unsigned)(((v.tw + 2 * v.hb) * (v.th + 2 * v.hb) + 15) & ~15) + (unsigned)(table_words * 4) + 16u;
}

// The variants generated for a rule set whose table has table_words u32 words:
// those whose LDS fits a workgroup.
std::vector<se::LdsVariant> variants_for(uint64_t table_words) {
    std::vector<se::LdsVariant> out;
    for (auto& v : lds_variants())
        if (lds_bytes_of(v, table_words) <= kMaxLdsBytes) out.push_back(v);
    return out;
}

uint64_t table_words_of(const se::HipGenInfo& info, uint64_t ct_x_words) {
    const uint64_t n = (uint64_t)info.nmat, D = (uint64_t)info.ct_D;
    if (info.table) return n * n * n * n * (uint64_t)info.table_R;
    if (info.ctable) return (D * D * D * D + 15) / 16 * 4 + ct_x_words;
    return 0;
}

const char* kJitOpts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
                          "-Wno-parentheses-equality", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                          "-Xclang", "-target-feature", "-Xclang", "-unaligned-access-mode"};

int jit_compile(const std::string& src, std::vector<char>& code) {
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "se_rules.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
        return fail(SE_ERR_JIT, "hiprtcCreateProgram failed");
    hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof kJitOpts / sizeof kJitOpts[0]), kJitOpts);
    if (r != HIPRTC_SUCCESS) {
        size_t ls = 0;
        hiprtcGetProgramLogSize(prog, &ls);
        std::string log(ls + 1, '\0');
        hiprtcGetProgramLog(prog, &log[0]);
        hiprtcDestroyProgram(&prog);
        return fail(SE_ERR_JIT, "hiprtc compile failed: " + log);
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    code.resize(cs);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    return SE_OK;
}

int parse_yaml(const char* yaml, size_t len, se::ParsingResult& pr, std::string* kind = nullptr) {
    if (!yaml) return fail(SE_ERR_ARG, "null rule text");
    try {
        pr = se::parse_string(std::string(yaml, len));
    } catch (const se::LangError& e) {
        if (kind) *kind = e.kind;
        return fail(SE_ERR_PARSE, e.what());
    } catch (const std::exception& e) {
        return fail(SE_ERR_PARSE, e.what());
    }
    return SE_OK;
}

int gen_source(const se::ParsingResult& pr, std::string& src, se::HipGenInfo* info = nullptr,
               bool allow_ctable = true, uint64_t ct_x_words = 0, bool with_lds = true) {
    try {
        se::HipGenInfo local;
        se::HipGenInfo& gi = info ? *info : local;
        se::hip_rules_source(pr, &gi, allow_ctable, ct_x_words);
        src = se::hip_program_source(pr, with_lds ? variants_for(table_words_of(gi, ct_x_words))
                                                  : std::vector<se::LdsVariant>(),
                                     allow_ctable, ct_x_words);
    } catch (const se::LangError& e) {
        return fail(SE_ERR_CODEGEN, e.what());
    } catch (const std::exception& e) {
        return fail(SE_ERR_CODEGEN, e.what());
    }
    return SE_OK;
}

uint64_t fnv1a(const std::string& s) {
    uint64_t h = 1469598103934665603ull;
    for (unsigned char c : s) h = (h ^ c) * 1099511628211ull;
    return h;
}

// The coded table's LDS form from its raw enumeration (se_build_ctable: for
// each tuple of codes < D, class r < R of hash lanes 1..3 and mirror bit m, the
// block result as ids in bytes).  Codes are (id - 3) mod nmat.  Layout, u32 words:
//   T1: D^4 u8 entries, the tuple's outcome vector v (its outcome for every
//       class), padded to 16 bytes;
//   X:  per (v, r, m) an outcome {sel, consts}: new block = v_perm(consts, block,
//       sel), where selector byte j < 4 takes byte j of the block's codes and
//       4 + j constant code byte j.
// An outcome encodes the unmirrored result relative to the tuple (a cell kept in
// place, else the first input cell holding the same id, else a constant).  The
// mirrored result must be the unmirrored result of the swapped tuple (c1, c0,
// c3, c2), swapped back; its entry applies that to the unswapped block.
// Returns false when the rule set does not fit: the mirror is not such a
// permutation, a result holds NULL or WALL (kernels skip the NULL/WALL test in
// tiles free of both), or more than 256 outcome vectors.
bool compact_ctable(const std::vector<uint32_t>& raw, int D, int R, int nmat, std::vector<uint32_t>& out) {
    const size_t tuples = (size_t)D * D * D * D;
    const size_t t1_words = (tuples + 15) / 16 * 4;
    auto enc = [&](uint32_t id) { return (id + (uint32_t)nmat - 3u) % (uint32_t)nmat; };
    auto byte = [](uint32_t v, int j) { return (v >> (8 * j)) & 0xffu; };
    static const int P[4] = {1, 0, 3, 2};  // the mirror swaps c0 <-> c1 and c2 <-> c3
    std::map<std::vector<uint64_t>, int> vecs;
    std::vector<uint64_t> vec(R);
    std::vector<uint32_t> t1(t1_words, 0u), x;
    for (size_t ti = 0; ti < tuples; ti++) {
        uint32_t a[4], id[4];
        size_t rest = ti;
        for (int k = 0; k < 4; k++) {
            a[k] = (uint32_t)(rest % D);
            rest /= D;
            id[k] = (a[k] + 3u) % (uint32_t)nmat;
        }
        const size_t tp = a[1] + (size_t)D * (a[0] + (size_t)D * (a[3] + (size_t)D * a[2]));
        for (int r = 0; r < R; r++) {
            const uint32_t o = raw[(ti * R + r) * 2], om = raw[(ti * R + r) * 2 + 1];
            const uint32_t op = raw[(tp * R + r) * 2];
            for (int j = 0; j < 4; j++) {
                if (byte(om, j) != byte(op, P[j])) return false;
                if (byte(o, j) == 1u || byte(o, j) == 2u) return false;  // results stay free of NULL and WALL
            }
            uint32_t sel = 0, cst = 0;
            for (int j = 0; j < 4; j++) {
                const uint32_t v = byte(o, j);
                int src = id[j] == v ? j : -1;
                for (int i = 0; i < 4 && src < 0; i++)
                    if (id[i] == v) src = i;
                if (src < 0) {
                    src = 4 + j;
                    cst |= enc(v) << (8 * j);
                }
                sel |= (uint32_t)src << (8 * j);
            }
            vec[r] = ((uint64_t)sel << 32) | cst;
        }
        auto it = vecs.find(vec);
        if (it == vecs.end()) {
            if (vecs.size() == 256) return false;
            it = vecs.emplace(vec, (int)vecs.size()).first;
            for (int r = 0; r < R; r++) {
                const uint32_t sel = (uint32_t)(vec[r] >> 32), cst = (uint32_t)vec[r];
                uint32_t selm = 0;
                for (int j = 0; j < 4; j++) {
                    const uint32_t sv = byte(sel, P[j]);
                    selm |= (sv < 4 ? (uint32_t)P[sv] : sv) << (8 * j);
                }
                x.insert(x.end(), {sel, cst, selm, cst});  // (v, r, m = 0) then (v, r, m = 1)
            }
        }
        t1[ti / 4] |= (uint32_t)it->second << (8 * (ti & 3));
    }
    out = t1;
    out.insert(out.end(), x.begin(), x.end());
    return true;
}

constexpr int kCtableUnfit = 1;  // build_ctable: the rule set does not fit the coded table

// The coded table of a rule set: builds a module without LDS kernels, runs its
// se_build_ctable and compacts the result on the host (compact_ctable).
// *stable: bit c set when a block of four cells of code c comes out unchanged
// for every class and mirror bit (see uniform_stable_mask).
int build_ctable(se_ctx* ctx, const se::ParsingResult& pr, const se::HipGenInfo& info,
                 std::vector<uint32_t>& words, uint32_t* stable) {
    std::string src;
    int rc = gen_source(pr, src, nullptr, true, 0, false);
    if (rc) return rc;
    std::vector<char> code;
    rc = jit_compile(src, code);
    if (rc) return rc;
    DeviceGuard dg(ctx->device);
    hipModule_t mod = nullptr;
    HIPCHECK(hipModuleLoadData(&mod, code.data()));
    const size_t D = (size_t)info.ct_D, R = (size_t)info.ct_R, tuples = D * D * D * D;
    std::vector<uint32_t> raw(tuples * R * 2);
    uint32_t* d_raw = nullptr;
    hipFunction_t build;
    if (hipModuleGetFunction(&build, mod, "se_build_ctable") != hipSuccess)
        rc = fail(SE_ERR_JIT, "se_build_ctable missing");
    if (!rc && hipMalloc(&d_raw, raw.size() * 4) != hipSuccess) rc = fail(SE_ERR_OOM, "coded table buffer");
    if (!rc) rc = launch(build, (unsigned)((tuples + 255) / 256), ctx->stream, d_raw);
    if (!rc && hipMemcpyDtoHAsync(raw.data(), d_raw, raw.size() * 4, ctx->stream) != hipSuccess)
        rc = fail(SE_ERR_HIP, "coded table read-back failed");
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(SE_ERR_HIP, "coded table build failed");
    (void)hipFree(d_raw);
    (void)hipModuleUnload(mod);
    if (rc) return rc;
    *stable = 0u;
    for (uint32_t c = 0; c < (uint32_t)D && c < 32u; c++) {
        const size_t ti = (size_t)c * (1 + D + D * D + D * D * D);
        const uint32_t id = (c + 3u) % (uint32_t)info.nmat;
        bool keep = true;
        for (size_t k = 0; k < R * 2; k++) keep = keep && raw[ti * R * 2 + k] == id * 0x01010101u;
        if (keep) *stable |= 1u << c;
    }
    if (!compact_ctable(raw, (int)D, (int)R, info.nmat, words)) return kCtableUnfit;
    return variants_for(words.size()).empty() ? kCtableUnfit : SE_OK;  // no tile fits beside it
}

}  // namespace

namespace {

// Does every entry of a full transition table (entry tuple * R + r holds the
// block result as four id bytes; tuple = c0 + n*(c1 + n*(c2 + n*c3))) leave each
// WALL cell of its tuple WALL?
bool table_keeps_wall(const std::vector<uint32_t>& tbl, size_t n, size_t R) {
    const uint32_t wall = 2u;
    for (size_t tuple = 0; tuple < n * n * n * n; tuple++) {
        uint32_t mask = 0, t = (uint32_t)tuple;
        for (int i = 0; i < 4; i++, t /= (uint32_t)n)
            if (t % n == wall) mask |= 0xffu << (8 * i);
        if (!mask) continue;
        for (size_t r = 0; r < R; r++)
            if ((tbl[tuple * R + r] & mask) != ((wall * 0x01010101u) & mask)) return false;
    }
    return true;
}

// Bit c set when a block of four cells of id c comes out unchanged under every
// hash class of a full transition table (entry tuple * R + r).  A tile whose
// cells all hold such a c is a fixed point of every sub-step (se_lds_body skips
// its sub-steps); with swap-only rules every c qualifies.
uint32_t uniform_stable_mask(const std::vector<uint32_t>& tbl, size_t n, size_t R) {
    uint32_t mask = 0;
    for (size_t c = 0; c < n && c < 32; c++) {
        const size_t tuple = c * (1 + n + n * n + n * n * n);
        bool keep = true;
        for (size_t r = 0; r < R; r++) keep = keep && tbl[tuple * R + r] == (uint32_t)c * 0x01010101u;
        if (keep) mask |= 1u << c;
    }
    return mask;
}

// Writes a u32 module global of a loaded module.
int set_module_u32(hipModule_t mod, const char* name, uint32_t value) {
    void* p = nullptr;
    size_t bytes = 0;
    HIPCHECK(hipModuleGetGlobal(&p, &bytes, mod, name));
    if (bytes != sizeof(uint32_t)) return fail(SE_ERR_JIT, std::string("internal: ") + name + " has the wrong size");
    HIPCHECK(hipMemcpy(p, &value, sizeof value, hipMemcpyHostToDevice));
    return SE_OK;
}

// ctwords: the compacted coded table when the source was generated for one;
// ct_stable its uniform-block mask (build_ctable).
int load_module(se_ctx* ctx, const std::string& src, const se::HipGenInfo& info,
                const std::vector<uint32_t>* ctwords, uint32_t ct_stable, std::shared_ptr<se_module>& out) {
    uint64_t key = fnv1a(src);
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        auto it = ctx->cache.find(key);
        if (it != ctx->cache.end() && it->second.first == src) {
            out = it->second.second;
            return SE_OK;
        }
    }
    std::vector<char> code;
    int rc = jit_compile(src, code);
    if (rc) return rc;
    DeviceGuard dg(ctx->device);
    auto m = std::make_shared<se_module>();
    m->device = ctx->device;
    HIPCHECK(hipModuleLoadData(&m->mod, code.data()));
    for (int i = 0; i < 6; i++) {
        const int R = i % 3 == 0 ? 1 : i % 3 == 1 ? 2 : 4, nt = i / 3;
        std::string name = "se_step_plain_r" + std::to_string(R) + "_nt" + std::to_string(nt);
        HIPCHECK(hipModuleGetFunction(&m->plain_v[i], m->mod, name.c_str()));
    }
    m->plain = m->plain_v[0];
    HIPCHECK(hipModuleGetFunction(&m->fill, m->mod, "se_fill"));
    HIPCHECK(hipModuleGetFunction(&m->fill_seeded, m->mod, "se_fill_seeded"));
    HIPCHECK(hipModuleGetFunction(&m->copy, m->mod, "se_copy"));
    HIPCHECK(hipModuleGetFunction(&m->fill_spec, m->mod, "se_fill_spec"));
    HIPCHECK(hipModuleGetFunction(&m->render, m->mod, "se_render"));
    HIPCHECK(hipModuleGetFunction(&m->noise_fill, m->mod, "se_noise_fill"));
    int max_lds = 65536;
    hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device);
    if (info.table) {
        size_t n = (size_t)info.nmat;
        m->table_entries = n * n * n * n * (size_t)info.table_R;
        m->table_form = 1;
        HIPCHECK(hipMalloc(&m->table, m->table_entries * 4));
        hipFunction_t build;
        HIPCHECK(hipModuleGetFunction(&build, m->mod, "se_build_table"));
        size_t tuples = n * n * n * n;
        int rc = launch(build, (unsigned)((tuples + 255) / 256), ctx->stream, m->table);
        if (rc) return rc;
        HIPCHECK(hipStreamSynchronize(ctx->stream));
        // Edge tiles hold WALL for the cells outside the world.  If no entry
        // changes a WALL cell, the table keeps them by itself and the edge-tile
        // block loop can skip restoring them (se_wall_kept).
        std::vector<uint32_t> tbl(m->table_entries);
        HIPCHECK(hipMemcpy(tbl.data(), m->table, tbl.size() * 4, hipMemcpyDeviceToHost));
        m->wall_kept = table_keeps_wall(tbl, n, (size_t)info.table_R);
        if (m->wall_kept) {
            rc = set_module_u32(m->mod, "se_wall_kept", 1u);
            if (rc) return rc;
        }
        m->stable_mask = uniform_stable_mask(tbl, n, (size_t)info.table_R);
    }
    if (info.ctable) {
        if (!ctwords) return fail(SE_ERR_JIT, "internal: coded-table source without its table");
        const std::vector<uint32_t>& words = *ctwords;
        m->table_entries = words.size();
        m->table_form = 2;
        HIPCHECK(hipMalloc(&m->table, words.size() * 4));
        HIPCHECK(hipMemcpy(m->table, words.data(), words.size() * 4, hipMemcpyHostToDevice));
        m->stable_mask = ct_stable;
    }
    if (m->stable_mask) {
        rc = set_module_u32(m->mod, "se_stable_mask", m->stable_mask);
        if (rc) return rc;
    }
    for (auto& v : variants_for(m->table_entries)) {  // the variants the source was generated with
        hipFunction_t f;
        HIPCHECK(hipModuleGetFunction(&f, m->mod, se::lds_kernel_name(v).c_str()));
        const unsigned bytes = lds_bytes_of(v, m->table_entries);
        if ((int)bytes > max_lds) continue;  // does not fit this device's LDS
        hipFunction_t fb;
        std::string bname = se::lds_kernel_name(v);
        bname.replace(0, std::strlen("se_step_lds_"), "se_batch_lds_");
        HIPCHECK(hipModuleGetFunction(&fb, m->mod, bname.c_str()));
        int per_cu = 0, cus = 0;
        HIPCHECK(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, v.nt, 0));
        HIPCHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
        m->lds.push_back({v, f});
        m->batch.push_back(fb);
        m->lds_bytes.push_back(bytes);
        m->lds_grid.push_back(std::max(8, (std::max(per_cu, 1) * cus) & ~7));
        m->lds_waves.push_back(std::max(per_cu, 1) * v.nt / 64);
    }
    if (m->lds.empty()) return fail(SE_ERR_JIT, "no LDS kernel variant fits the device's LDS");
    // default single-world variant: 112 x 112 tiles, halo 8, with the smallest
    // workgroup that keeps >= 24 wavefronts resident per CU (measured: the full
    // table reaches that with 256 threads, the coded table's larger LDS
    // footprint needs 512), else the one with the most resident wavefronts
    m->default_lds = -1;
    for (size_t i = 0; i < m->lds.size(); i++) {
        const se::LdsVariant& v = m->lds[i].first;
        if (v.tw != kDefaultTile || v.th != kDefaultTile || v.hb != kDefaultHalo) continue;
        const bool enough = m->lds_waves[i] >= 24;
        if (m->default_lds < 0) {
            m->default_lds = (int)i;
            continue;
        }
        const int d = m->default_lds;
        const bool d_enough = m->lds_waves[d] >= 24;
        if (enough && (!d_enough || v.nt < m->lds[d].first.nt)) m->default_lds = (int)i;
        else if (!enough && !d_enough && m->lds_waves[i] > m->lds_waves[d]) m->default_lds = (int)i;
    }
    if (m->default_lds < 0) m->default_lds = 0;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->cache[key] = {src, m};
    out = m;
    return SE_OK;
}

void offset_of(int32_t frame, int& ox, int& oy) {
    int m = frame % 4;  // operations.glsl:25-34
    ox = (m == 1 || m == 3) ? 1 : 0;
    oy = (m == 1 || m == 2) ? 1 : 0;
}

struct ModsArg {
    void* ptr;
    int n;
};

// Uploads the queued modifications (cut at the first size-0 entry, as the
// kernel's loop stops there, falling_sand.glsl:139-141) and clears the queue.
int take_mods(se_world* w, ModsArg& m) {
    m.ptr = w->d_mods;
    m.n = 0;
    if (w->pending.empty()) return SE_OK;
    int n = 0;
    while (n < (int)w->pending.size() && n < SE_MAX_MODIFICATIONS && w->pending[n].mod_size != 0) n++;
    if (n > 0)
        HIPCHECK(hipMemcpyAsync(w->d_mods, w->pending.data(), sizeof(se_modification) * n, hipMemcpyHostToDevice,
                                w->stream));
    w->pending.clear();
    m.n = n;
    return SE_OK;
}

int launch_fill(se_world* w, uint32_t* g, uint32_t v) {
    size_t n = (size_t)w->pitch * w->H;
    unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
    return launch(w->mod->fill, blocks, w->stream, g, n, v);
}

// Colour and light of the step that turned `prev` into `cur` (render enabled).
constexpr int kRenderRows = 8;  // rows per wavefront of se_render (SE_RENDER_ROWS)

int render_step(se_world* w, const uint32_t* prev, const uint32_t* cur) {
    if (!w->color) return SE_OK;
    dim3 grid((w->W + 255) / 256, (w->H + kRenderRows - 1) / kRenderRows);
    void* args[] = {(void*)&prev, (void*)&cur, &w->light[w->lcur], &w->light[w->lcur ^ 1], &w->color,
                    (void*)&w->noise, (void*)&w->W, (void*)&w->H, (void*)&w->pitch};
    HIPCHECK(hipModuleLaunchKernel(w->mod->render, grid.x, grid.y, 1, 256, 1, 1, 0, w->stream, args, nullptr));
    w->lcur ^= 1;
    return SE_OK;
}

}  // namespace

// One Simulation::run with the plain kernel.
int se_rt::plain_step(se_world* w) {
    se_rt::sleep_reset(w);
    w->frame += 1;
    uint32_t* in = w->buf[w->cur];
    uint32_t* out = w->buf[w->cur ^ 1];
    if (w->frame == 1) {  // falling_sand.glsl:128-131
        w->pending.clear();
        int rc = launch_fill(w, out, 0u);
        if (!rc) rc = render_step(w, in, out);
        if (rc) return rc;
        w->cur ^= 1;
        return SE_OK;
    }
    ModsArg m;
    int rc = take_mods(w, m);
    if (rc) return rc;
    int ox, oy;
    offset_of(w->frame, ox, oy);
    int strips = (int)((w->W + 255) / 256);
    int brows = (int)((w->H + (uint32_t)oy + 1) / 2);
    static const int kRows[3] = {1, 2, 4};
    const int R = kRows[w->plain_variant % 3];
    size_t waves = (size_t)strips * ((brows + R - 1) / R);
    unsigned blocks = (unsigned)((waves + 3) / 4);
    int rc2 = launch(w->mod->plain_v[w->plain_variant], blocks, w->stream, (const uint32_t*)in, out, (int)w->W,
                     (int)w->H,
                     (int)w->pitch, (int)w->frame, ox, oy, strips, brows, (const void*)m.ptr, m.n);
    if (!rc2) rc2 = render_step(w, in, out);
    if (rc2) return rc2;
    w->cur ^= 1;
    return SE_OK;
}

namespace {
using se_rt::plain_step;

// LDS instantiation for a launch of up to `sub` sub-steps (0: the variant's halo):
// the world's tuned choice if set, else the default tile (fastest measured on a
// 16384^2 world) when it allows `sub`, else the variant with the smallest halo
// >= sub.

}  // namespace

int se_rt::default_lds_index(const se_module* m) { return m->default_lds; }

namespace {

const std::pair<se::LdsVariant, hipFunction_t>* pick_lds(se_world* w, uint32_t sub) {
    auto& L = w->mod->lds;
    if (w->lds_variant >= 0) {
        const auto& p = L[w->lds_variant];
        return (sub == 0 || (uint32_t)p.first.hb >= sub) ? &p : nullptr;
    }
    {
        const auto& p = L[w->mod->default_lds];
        if (sub == 0 || sub <= (uint32_t)p.first.hb) return &p;
    }
    const std::pair<se::LdsVariant, hipFunction_t>* best = nullptr;
    for (auto& p : L) {
        if (sub != 0 && (uint32_t)p.first.hb < sub) continue;
        if (!best || p.first.hb < best->first.hb) best = &p;
    }
    return best;
}

// Advances `n` <= hb frames with one LDS launch (mods, if any, at its first).
// With tile sleeping on, the launch records which tiles ended settled and, when
// the previous launch of the same variant left such a record, lets tiles that
// stayed settled skip their loads and stores (se_tile_sleeps).
int lds_launch(se_world* w, const std::pair<se::LdsVariant, hipFunction_t>& v, int n) {
    ModsArg m;
    int rc = take_mods(w, m);
    if (rc) return rc;
    const se::LdsVariant& L = v.first;
    int tiles_x = (int)((w->W + L.tw - 1) / L.tw), tiles_y = (int)((w->H + L.th - 1) / L.th);
    const int vi = (int)(&v - w->mod->lds.data());
    const unsigned char* s_in = nullptr;
    unsigned char* s_out = nullptr;
    if (w->sleep_enabled) {
        const size_t ntiles = (size_t)tiles_x * tiles_y;
        if (ntiles > w->sleep_cap) {
            hipFree(w->d_sleep[0]);
            hipFree(w->d_sleep[1]);
            w->d_sleep[0] = w->d_sleep[1] = nullptr;
            w->sleep_cap = 0;
            w->sleep_variant = -1;
            HIPCHECK(hipMalloc(&w->d_sleep[0], ntiles));
            HIPCHECK(hipMalloc(&w->d_sleep[1], ntiles));
            w->sleep_cap = ntiles;
        }
        if (w->sleep_variant == vi) s_in = w->d_sleep[w->sleep_p];
        s_out = w->d_sleep[w->sleep_p ^ 1];
    }
    rc = launch_ex(v.second, (unsigned)(tiles_x * tiles_y), (unsigned)L.nt, 0u, w->stream,
                   (const uint32_t*)w->buf[w->cur], w->buf[w->cur ^ 1], (int)w->W, (int)w->H, (int)w->pitch,
                   (int)(w->frame + 1), n, tiles_x, (const void*)m.ptr, m.n, (const uint32_t*)w->mod->table, s_in,
                   s_out);
    if (rc) return rc;
    if (s_out) {
        w->sleep_p ^= 1;
        w->sleep_variant = vi;
    }
    w->frame += n;
    w->cur ^= 1;
    return SE_OK;
}

int blocked_steps(se_world* w, uint32_t nsteps, uint32_t sub) {
    if (w->color) {  // render side-outputs need every step's grid: one step per launch
        for (uint32_t s = 0; s < nsteps; s++) {
            int rc = plain_step(w);
            if (rc) return rc;
        }
        return SE_OK;
    }
    auto* v = pick_lds(w, sub);
    if (!v) return fail(SE_ERR_ARG, "no LDS variant supports " + std::to_string(sub) + " sub-steps");
    uint32_t k = sub ? sub : (uint32_t)v->first.hb;
    // a world inside one tile keeps its WALL border in LDS: no halo limit on k
    const bool single = w->W <= (uint32_t)v->first.tw && w->H <= (uint32_t)v->first.th;
    if (single && sub == 0) k = 1u << 20;
    uint32_t done = 0;
    while (done < nsteps) {
        if (w->frame + 1 == 1) {  // the clearing frame stays a step of its own
            int rc = plain_step(w);
            if (rc) return rc;
            done++;
            continue;
        }
        uint32_t n = std::min<uint32_t>(k, nsteps - done);
        int rc = lds_launch(w, *v, (int)n);
        if (rc) return rc;
        done += n;
    }
    return SE_OK;
}

}  // namespace

// ------------------------------------------------------------------ C ABI
extern "C" {

const char* se_last_error(void) { return g_err.c_str(); }
const char* se_version(void) { return "sandengine-hip 0.1 (gfx950)"; }
void se_free(void* p) { std::free(p); }

int se_parse_ir_json(const char* yaml, size_t len, char** json_out, char** kind_out) {
    if (json_out) *json_out = nullptr;
    if (kind_out) *kind_out = nullptr;
    se::ParsingResult pr;
    std::string kind;
    int rc = parse_yaml(yaml, len, pr, &kind);
    if (rc) {
        if (kind_out) *kind_out = dupstr(kind);
        return rc;
    }
    if (json_out) *json_out = dupstr(se::ir_json(pr));
    return SE_OK;
}

int se_emit_glsl(const char* yaml, size_t len, char** materials_out, char** rules_out) {
    se::ParsingResult pr;
    int rc = parse_yaml(yaml, len, pr);
    if (rc) return rc;
    auto g = se::create_glsl(pr);
    if (materials_out) *materials_out = dupstr(g.first);
    if (rules_out) *rules_out = dupstr(g.second);
    return SE_OK;
}

int se_emit_hip(const char* yaml, size_t len, char** src_out) {
    se::ParsingResult pr;
    int rc = parse_yaml(yaml, len, pr);
    if (rc) return rc;
    std::string src;
    rc = gen_source(pr, src);
    if (rc) return rc;
    if (src_out) *src_out = dupstr(src);
    return SE_OK;
}

int se_jit_check(const char* yaml, size_t len, size_t* code_bytes) {
    se::ParsingResult pr;
    int rc = parse_yaml(yaml, len, pr);
    if (rc) return rc;
    std::string src;
    rc = gen_source(pr, src);
    if (rc) return rc;
    std::vector<char> code;
    rc = jit_compile(src, code);
    if (rc) return rc;
    if (code_bytes) *code_bytes = code.size();
    return SE_OK;
}

int se_ctx_create(int device, se_ctx** out) {
    if (!out) return fail(SE_ERR_ARG, "null out");
    *out = nullptr;
    int n = 0;
    HIPCHECK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(SE_ERR_ARG, "device " + std::to_string(device) + " not present");
    DeviceGuard dg(device);
    auto* c = new se_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(SE_ERR_HIP, "hipStreamCreate failed");
    }
    // set-up of se_grids_step_noalloc: the device's scratch array and its pinned
    // staging.  Without them that entry point reports SE_ERR_STATE; the rest of
    // the library does not need them.
    if (se_rt::noalloc_scratch(device, &c->scratch) == SE_OK) {
        if (hipHostMalloc(&c->staging, SE_NOALLOC_SCRATCH_BYTES, hipHostMallocDefault) != hipSuccess ||
            hipEventCreateWithFlags(&c->staging_done, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            if (c->staging) (void)hipHostFree(c->staging);
            c->staging = nullptr;
            c->scratch = nullptr;
        }
    }
    *out = c;
    return SE_OK;
}

void se_ctx_destroy(se_ctx* ctx) {
    if (!ctx) return;
    {
        DeviceGuard dg(ctx->device);
        hipStreamSynchronize(ctx->stream);
        hipStreamDestroy(ctx->stream);
        if (ctx->staging_done) {
            hipEventSynchronize(ctx->staging_done);
            hipEventDestroy(ctx->staging_done);
        }
        if (ctx->staging) hipHostFree(ctx->staging);
    }
    delete ctx;
}

int se_ctx_mem_info(se_ctx* ctx, size_t* total, size_t* free_bytes) {
    if (!ctx) return fail(SE_ERR_ARG, "null ctx");
    DeviceGuard dg(ctx->device);
    size_t f = 0, t = 0;
    HIPCHECK(hipMemGetInfo(&f, &t));
    if (total) *total = t;
    if (free_bytes) *free_bytes = f;
    return SE_OK;
}

int se_ruleset_from_yaml(se_ctx* ctx, const char* yaml, size_t len, se_ruleset** out) {
    if (!ctx || !out) return fail(SE_ERR_ARG, "null argument");
    *out = nullptr;
    auto rs = std::make_unique<se_ruleset>();
    int rc = parse_yaml(yaml, len, rs->pr);
    if (rc) return rc;
    std::string src;
    se::HipGenInfo info;
    rc = gen_source(rs->pr, src, &info);
    if (rc) return rc;
    if (info.ctable) {
        // the coded table is enumerated by a first build; the LDS kernels are then
        // generated with its size (static LDS), or without it when it does not fit
        std::vector<uint32_t> words;
        uint32_t stable = 0;
        rc = build_ctable(ctx, rs->pr, info, words, &stable);
        if (rc == kCtableUnfit) {
            rc = gen_source(rs->pr, src, &info, false);
            if (!rc) rc = load_module(ctx, src, info, nullptr, 0u, rs->mod);
        } else if (!rc) {
            const uint64_t t1_words = ((uint64_t)info.ct_D * info.ct_D * info.ct_D * info.ct_D + 15) / 16 * 4;
            rc = gen_source(rs->pr, src, &info, true, words.size() - t1_words);
            if (!rc) rc = load_module(ctx, src, info, &words, stable, rs->mod);
        }
    } else {
        rc = load_module(ctx, src, info, nullptr, 0u, rs->mod);
    }
    if (rc) return rc;
    rs->ctx = ctx;
    rs->nmat = (int)rs->pr.materials.size();
    *out = rs.release();
    return SE_OK;
}

void se_ruleset_destroy(se_ruleset* rs) { delete rs; }

int se_ruleset_material_id(const se_ruleset* rs, const char* name, int32_t* id) {
    if (!rs || !name || !id) return fail(SE_ERR_ARG, "null argument");
    for (auto& m : rs->pr.materials)
        if (m.name == name) {
            *id = (int32_t)m.id;
            return SE_OK;
        }
    return fail(SE_ERR_ARG, std::string("no material named '") + name + "'");
}

int se_ruleset_kernel_form(const se_ruleset* rs, int32_t* form) {
    if (!rs || !form) return fail(SE_ERR_ARG, "null argument");
    *form = rs->mod->table_form;
    return SE_OK;
}

int se_ruleset_material_count(const se_ruleset* rs, int32_t* n) {
    if (!rs || !n) return fail(SE_ERR_ARG, "null argument");
    *n = rs->nmat;
    return SE_OK;
}

int se_world_create(se_ctx* ctx, const se_ruleset* rs, uint32_t w, uint32_t h, se_world** out) {
    if (!ctx || !rs || !out) return fail(SE_ERR_ARG, "null argument");
    *out = nullptr;
    if (w == 0 || h == 0 || w > (1u << 20) || h > (1u << 20))
        return fail(SE_ERR_ARG, "world size must be in [1, 2^20] per side");
    if (rs->ctx != ctx) return fail(SE_ERR_ARG, "rule set belongs to another context");
    DeviceGuard dg(ctx->device);
    auto* W = new se_world();
    W->ctx = ctx;
    W->rs = rs;
    W->mod = rs->mod;
    W->W = w;
    W->H = h;
    W->pitch = (w + 63u) & ~63u;
    W->stream = ctx->stream;
    size_t bytes = (size_t)W->pitch * h * sizeof(uint32_t);
    if (hipMalloc(&W->buf[0], bytes) != hipSuccess || hipMalloc(&W->buf[1], bytes) != hipSuccess ||
        hipMalloc(&W->d_mods, sizeof(se_modification) * SE_MAX_MODIFICATIONS) != hipSuccess) {
        hipFree(W->buf[0]);
        hipFree(W->buf[1]);
        hipFree(W->d_mods);
        delete W;
        hipGetLastError();
        return fail(SE_ERR_OOM, "hipMalloc of " + std::to_string(2 * bytes) + " bytes failed");
    }
    int rc = launch_fill(W, W->buf[0], 0u);
    if (!rc) rc = launch_fill(W, W->buf[1], 0u);
    if (rc) {
        se_world_destroy(W);
        return rc;
    }
    *out = W;
    return SE_OK;
}

void se_world_destroy(se_world* w) {
    if (!w) return;
    {
        DeviceGuard dg(w->ctx->device);
        hipStreamSynchronize(w->stream);
        if (w->owns) {
            hipFree(w->buf[0]);
            hipFree(w->buf[1]);
            hipFree(w->d_mods);
        }
        hipFree(w->light[0]);
        hipFree(w->light[1]);
        hipFree(w->color);
        hipFree(w->noise);
        hipFree(w->d_sleep[0]);
        hipFree(w->d_sleep[1]);
    }
    delete w;
}

int se_world_upload(se_world* w, const uint32_t* cells, size_t n) {
    if (!w || !cells) return fail(SE_ERR_ARG, "null argument");
    if (n != (size_t)w->W * w->H) return fail(SE_ERR_ARG, "upload size mismatch");
    DeviceGuard dg(w->ctx->device);
    se_rt::sleep_reset(w);
    HIPCHECK(hipMemcpy2DAsync(w->buf[w->cur], (size_t)w->pitch * 4, cells, (size_t)w->W * 4, (size_t)w->W * 4, w->H,
                              hipMemcpyHostToDevice, w->stream));
    HIPCHECK(hipStreamSynchronize(w->stream));
    return SE_OK;
}

int se_world_download(se_world* w, uint32_t* cells, size_t n) {
    if (!w || !cells) return fail(SE_ERR_ARG, "null argument");
    if (n != (size_t)w->W * w->H) return fail(SE_ERR_ARG, "download size mismatch");
    DeviceGuard dg(w->ctx->device);
    HIPCHECK(hipMemcpy2DAsync(cells, (size_t)w->W * 4, w->buf[w->cur], (size_t)w->pitch * 4, (size_t)w->W * 4, w->H,
                              hipMemcpyDeviceToHost, w->stream));
    HIPCHECK(hipStreamSynchronize(w->stream));
    return SE_OK;
}

int se_world_push_modification(se_world* w, const se_modification* m) {
    if (!w || !m) return fail(SE_ERR_ARG, "null argument");
    if (w->pending.size() >= SE_MAX_MODIFICATIONS)
        return fail(SE_ERR_FULL, "more than 256 modifications queued for one step");
    w->pending.push_back(*m);
    return SE_OK;
}

int se_world_step(se_world* w, uint32_t nsteps) {
    if (!w) return fail(SE_ERR_ARG, "null world");
    if ((int64_t)w->frame + nsteps > INT32_MAX) return fail(SE_ERR_ARG, "frame counter would overflow");
    DeviceGuard dg(w->ctx->device);
    for (uint32_t s = 0; s < nsteps; s++) {
        int rc = plain_step(w);
        if (rc) return rc;
    }
    return SE_OK;
}

int se_world_step_blocked(se_world* w, uint32_t nsteps, uint32_t sub) {
    if (!w) return fail(SE_ERR_ARG, "null world");
    if ((int64_t)w->frame + nsteps > INT32_MAX) return fail(SE_ERR_ARG, "frame counter would overflow");
    DeviceGuard dg(w->ctx->device);
    return blocked_steps(w, nsteps, sub);
}

int se_world_set_variant(se_world* w, int32_t plain_variant, int32_t lds_variant) {
    if (!w) return fail(SE_ERR_ARG, "null world");
    if (plain_variant < 0 || plain_variant >= 6) return fail(SE_ERR_ARG, "plain variant out of range");
    if (lds_variant < -1 || lds_variant >= (int)w->mod->lds.size()) return fail(SE_ERR_ARG, "lds variant out of range");
    w->plain_variant = plain_variant;
    w->lds_variant = lds_variant;
    se_rt::sleep_reset(w);
    return SE_OK;
}

int se_world_set_sleep(se_world* w, int enable) {
    if (!w) return fail(SE_ERR_ARG, "null world");
    w->sleep_enabled = enable != 0;
    se_rt::sleep_reset(w);
    return SE_OK;
}

int se_variant_info(const se_world* w, int32_t kind, int32_t index, int32_t* a, int32_t* b, int32_t* c) {
    if (!w) return fail(SE_ERR_ARG, "null world");
    auto put = [&](int32_t x, int32_t y, int32_t z) {
        if (a) *a = x;
        if (b) *b = y;
        if (c) *c = z;
        return SE_OK;
    };
    if (kind == 0) {
        if (index < 0 || index >= 6) return fail(SE_ERR_ARG, "plain variant out of range");
        static const int kRows[3] = {1, 2, 4};
        return put(kRows[index % 3], index / 3, 0);
    }
    if (kind == 2) {  // the LDS variant se_world_step_blocked(w, n, index) launches
        auto* p = pick_lds(const_cast<se_world*>(w), (uint32_t)std::max(index, 0));
        if (!p) return fail(SE_ERR_ARG, "no LDS variant supports " + std::to_string(index) + " sub-steps");
        return put((int32_t)(p - &w->mod->lds[0]), 0, 0);
    }
    if (index < 0 || index >= (int)w->mod->lds.size()) return fail(SE_ERR_ARG, "lds variant out of range");
    const se::LdsVariant& v = w->mod->lds[index].first;
    if (kind == 1) return put(v.tw, v.th, v.hb * 1000 + v.nt);
    if (kind == 3) return put(w->mod->lds_grid[index], (int32_t)w->mod->lds_bytes[index], 0);
    return fail(SE_ERR_ARG, "unknown variant kind");
}

int se_world_set_render(se_world* w, int enable) {
    if (!w) return fail(SE_ERR_ARG, "null world");
    DeviceGuard dg(w->ctx->device);
    HIPCHECK(hipStreamSynchronize(w->stream));
    auto release = [&]() {
        hipFree(w->light[0]);
        hipFree(w->light[1]);
        hipFree(w->color);
        hipFree(w->noise);
        w->light[0] = w->light[1] = w->color = nullptr;
        w->noise = nullptr;
    };
    if (!enable) {
        release();
        return SE_OK;
    }
    if (w->color) return SE_OK;
    const size_t bytes = (size_t)w->pitch * w->H * 16;
    if (hipMalloc(&w->light[0], bytes) != hipSuccess || hipMalloc(&w->light[1], bytes) != hipSuccess ||
        hipMalloc(&w->color, bytes) != hipSuccess || hipMalloc(&w->noise, bytes / 4) != hipSuccess) {
        hipGetLastError();
        release();
        return fail(SE_ERR_OOM, "hipMalloc of " + std::to_string(13 * bytes / 4) + " bytes for render outputs failed");
    }
    for (void* p : {w->light[0], w->light[1], w->color}) HIPCHECK(hipMemsetAsync(p, 0, bytes, w->stream));
    {
        dim3 grid((w->W + 63) / 64, (w->H + 3) / 4);
        void* args[] = {(void*)&w->noise, (void*)&w->W, (void*)&w->H, (void*)&w->pitch};
        HIPCHECK(hipModuleLaunchKernel(w->mod->noise_fill, grid.x, grid.y, 1, 256, 1, 1, 0, w->stream, args, nullptr));
    }
    HIPCHECK(hipStreamSynchronize(w->stream));
    w->lcur = 0;
    return SE_OK;
}

int se_world_render_download(se_world* w, float* color, float* light) {
    if (!w) return fail(SE_ERR_ARG, "null world");
    if (!w->color) return fail(SE_ERR_STATE, "render outputs are not enabled for this world");
    DeviceGuard dg(w->ctx->device);
    if (color)
        HIPCHECK(hipMemcpy2DAsync(color, (size_t)w->W * 16, w->color, (size_t)w->pitch * 16, (size_t)w->W * 16, w->H,
                                  hipMemcpyDeviceToHost, w->stream));
    if (light)
        HIPCHECK(hipMemcpy2DAsync(light, (size_t)w->W * 16, w->light[w->lcur], (size_t)w->pitch * 16,
                                  (size_t)w->W * 16, w->H, hipMemcpyDeviceToHost, w->stream));
    HIPCHECK(hipStreamSynchronize(w->stream));
    return SE_OK;
}

int se_world_render_timed(se_world* w, uint32_t n, float* gpu_ms) {
    if (!w || !gpu_ms) return fail(SE_ERR_ARG, "null argument");
    if (!w->color) return fail(SE_ERR_STATE, "render outputs are not enabled for this world");
    if (w->frame < 1) return fail(SE_ERR_STATE, "no step has been taken");
    DeviceGuard dg(w->ctx->device);
    hipEvent_t a, b;
    HIPCHECK(hipEventCreate(&a));
    HIPCHECK(hipEventCreate(&b));
    hipEventRecord(a, w->stream);
    int rc = SE_OK;
    for (uint32_t i = 0; i < n && !rc; i++) rc = render_step(w, w->buf[w->cur ^ 1], w->buf[w->cur]);
    hipEventRecord(b, w->stream);
    hipError_t e = hipEventSynchronize(b);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    hipEventDestroy(a);
    hipEventDestroy(b);
    if (rc) return rc;
    if (e != hipSuccess) return fail(SE_ERR_HIP, hipGetErrorString(e));
    *gpu_ms = ms;
    return SE_OK;
}

int se_world_render_ptrs(se_world* w, void** color, void** light, uint32_t* pitch) {
    if (!w) return fail(SE_ERR_ARG, "null world");
    if (!w->color) return fail(SE_ERR_STATE, "render outputs are not enabled for this world");
    if (color) *color = w->color;
    if (light) *light = w->light[w->lcur];
    if (pitch) *pitch = w->pitch;
    return SE_OK;
}

// ------------------------------------------------------------------ save / load
// File: a 40-byte header (magic "SEWORLD1", version 1, W, H, frame, number of
// queued modifications, a reserved word, FNV-1a 64 of the rule set's IR), the
// grid as W*H little-endian u32 (row-major), then the queued 32-byte
// modifications.  The grid moves in chunks of whole rows through a host buffer
// of at most 16 MiB.
namespace {
struct WorldFileHeader {
    char magic[8];
    uint32_t version, W, H;
    int32_t frame;
    uint32_t npending, reserved;
    uint64_t ruleset_hash;
};
static_assert(sizeof(WorldFileHeader) == 40, "world file header layout");
const char kWorldMagic[8] = {'S', 'E', 'W', 'O', 'R', 'L', 'D', '1'};

uint64_t ruleset_hash(const se_ruleset* rs) {
    const std::string j = se::ir_json(rs->pr);
    return se_fnv1a64(j.data(), j.size());
}

size_t chunk_rows(uint32_t W) { return std::max<size_t>(1, (16u << 20) / ((size_t)W * 4)); }

struct File {
    FILE* f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};
}  // namespace

int se_world_save(se_world* w, const char* path) {
    if (!w || !path) return fail(SE_ERR_ARG, "null argument");
    File F;
    F.f = std::fopen(path, "wb");
    if (!F.f) return fail(SE_ERR_ARG, std::string("cannot open '") + path + "' for writing");
    WorldFileHeader h{};
    std::memcpy(h.magic, kWorldMagic, 8);
    h.version = 1;
    h.W = w->W;
    h.H = w->H;
    h.frame = w->frame;
    h.npending = (uint32_t)w->pending.size();
    h.ruleset_hash = ruleset_hash(w->rs);
    if (std::fwrite(&h, sizeof h, 1, F.f) != 1) return fail(SE_ERR_ARG, "write failed");
    DeviceGuard dg(w->ctx->device);
    const size_t rows = chunk_rows(w->W);
    std::vector<uint32_t> buf(rows * w->W);
    for (size_t y = 0; y < w->H; y += rows) {
        const size_t n = std::min<size_t>(rows, w->H - y);
        HIPCHECK(hipMemcpy2DAsync(buf.data(), (size_t)w->W * 4, w->buf[w->cur] + y * w->pitch, (size_t)w->pitch * 4,
                                  (size_t)w->W * 4, n, hipMemcpyDeviceToHost, w->stream));
        HIPCHECK(hipStreamSynchronize(w->stream));
        if (std::fwrite(buf.data(), 4, n * w->W, F.f) != n * w->W) return fail(SE_ERR_ARG, "write failed");
    }
    if (!w->pending.empty() &&
        std::fwrite(w->pending.data(), sizeof(se_modification), w->pending.size(), F.f) != w->pending.size())
        return fail(SE_ERR_ARG, "write failed");
    if (std::fclose(F.f) != 0) {
        F.f = nullptr;
        return fail(SE_ERR_ARG, "write failed");
    }
    F.f = nullptr;
    return SE_OK;
}

int se_world_load(se_ctx* ctx, const se_ruleset* rs, const char* path, se_world** out) {
    if (!ctx || !rs || !path || !out) return fail(SE_ERR_ARG, "null argument");
    *out = nullptr;
    File F;
    F.f = std::fopen(path, "rb");
    if (!F.f) return fail(SE_ERR_ARG, std::string("cannot open '") + path + "'");
    WorldFileHeader h;
    if (std::fread(&h, sizeof h, 1, F.f) != 1 || std::memcmp(h.magic, kWorldMagic, 8) != 0 || h.version != 1)
        return fail(SE_ERR_ARG, std::string("'") + path + "' is not a saved world");
    if (h.ruleset_hash != ruleset_hash(rs)) return fail(SE_ERR_ARG, "the world was saved with a different rule set");
    if (h.frame < 0 || h.npending > SE_MAX_MODIFICATIONS) return fail(SE_ERR_ARG, "corrupt world header");
    se_world* w = nullptr;
    int rc = se_world_create(ctx, rs, h.W, h.H, &w);
    if (rc) return rc;
    std::unique_ptr<se_world, void (*)(se_world*)> guard(w, se_world_destroy);
    DeviceGuard dg(ctx->device);
    const size_t rows = chunk_rows(w->W);
    std::vector<uint32_t> buf(rows * w->W);
    for (size_t y = 0; y < w->H; y += rows) {
        const size_t n = std::min<size_t>(rows, w->H - y);
        if (std::fread(buf.data(), 4, n * w->W, F.f) != n * w->W) return fail(SE_ERR_ARG, "truncated world file");
        HIPCHECK(hipMemcpy2DAsync(w->buf[w->cur] + y * w->pitch, (size_t)w->pitch * 4, buf.data(), (size_t)w->W * 4,
                                  (size_t)w->W * 4, n, hipMemcpyHostToDevice, w->stream));
        HIPCHECK(hipStreamSynchronize(w->stream));
    }
    w->pending.resize(h.npending);
    if (h.npending && std::fread(w->pending.data(), sizeof(se_modification), h.npending, F.f) != h.npending)
        return fail(SE_ERR_ARG, "truncated world file");
    w->frame = h.frame;
    *out = guard.release();
    return SE_OK;
}

int se_world_get_frame(const se_world* w, int32_t* frame) {
    if (!w || !frame) return fail(SE_ERR_ARG, "null argument");
    *frame = w->frame;
    return SE_OK;
}

int se_world_set_frame(se_world* w, int32_t frame) {
    if (!w) return fail(SE_ERR_ARG, "null world");
    if (frame < 0) return fail(SE_ERR_ARG, "negative frame");
    w->frame = frame;
    return SE_OK;
}

int se_world_sync(se_world* w) {
    if (!w) return fail(SE_ERR_ARG, "null world");
    DeviceGuard dg(w->ctx->device);
    HIPCHECK(hipStreamSynchronize(w->stream));
    return SE_OK;
}

int se_world_device_ptr(se_world* w, void** ptr, uint32_t* pitch) {
    if (!w) return fail(SE_ERR_ARG, "null world");
    se_rt::sleep_reset(w);  // the caller may write through the pointer
    if (ptr) *ptr = w->buf[w->cur];
    if (pitch) *pitch = w->pitch;
    return SE_OK;
}

int se_world_fill_seeded(se_world* w, uint64_t seed, const uint32_t* ids, const uint32_t* thresholds, int n,
                         uint32_t stone_rows, uint32_t stone_id) {
    if (!w || !ids || !thresholds || n < 1 || n > 64) return fail(SE_ERR_ARG, "bad fill arguments");
    DeviceGuard dg(w->ctx->device);
    se_rt::sleep_reset(w);
    // ids and thresholds are staged in the world's modification buffer (8 KiB)
    void* d = w->d_mods;
    uint32_t host[128] = {0};
    std::memcpy(host, ids, n * 4);
    std::memcpy(host + 64, thresholds, n * 4);
    HIPCHECK(hipMemcpyAsync(d, host, sizeof host, hipMemcpyHostToDevice, w->stream));
    size_t cells = (size_t)w->W * w->H;
    unsigned blocks = (unsigned)std::min<size_t>((cells + 255) / 256, 1u << 16);
    int rc = launch(w->mod->fill_seeded, blocks, w->stream, w->buf[w->cur], (int)w->W, (int)w->H, (int)w->pitch,
                    (unsigned long long)seed, (const uint32_t*)d, (const uint32_t*)d + 64, n, (int)stone_rows,
                    stone_id);
    hipError_t e2 = hipStreamSynchronize(w->stream);
    if (rc) return rc;
    if (e2 != hipSuccess) return fail(SE_ERR_HIP, std::string("fill sync: ") + hipGetErrorString(e2));
    return SE_OK;
}

int se_world_copy_timed(se_world* w, uint32_t reps, float* gpu_ms) {
    if (!w || !gpu_ms) return fail(SE_ERR_ARG, "null argument");
    DeviceGuard dg(w->ctx->device);
    se_rt::sleep_reset(w);
    size_t n = (size_t)w->pitch * w->H;
    if (n % 4) return fail(SE_ERR_ARG, "grid size not a multiple of 4 cells");
    size_t n4 = n / 4;
    unsigned blocks = (unsigned)std::min<size_t>((n4 + 255) / 256, 1u << 20);
    hipEvent_t a, b;
    HIPCHECK(hipEventCreate(&a));
    HIPCHECK(hipEventCreate(&b));
    hipEventRecord(a, w->stream);
    int rc = SE_OK;
    for (uint32_t i = 0; i < reps && !rc; i++) {
        rc = launch(w->mod->copy, blocks, w->stream, (const uint32_t*)w->buf[w->cur], w->buf[w->cur ^ 1], n4);
        w->cur ^= 1;
    }
    hipEventRecord(b, w->stream);
    hipError_t e = hipEventSynchronize(b);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    hipEventDestroy(a);
    hipEventDestroy(b);
    if (rc) return rc;
    if (e != hipSuccess) return fail(SE_ERR_HIP, hipGetErrorString(e));
    *gpu_ms = ms;
    return SE_OK;
}

int se_world_step_timed(se_world* w, uint32_t nsteps, int mode, uint32_t sub, float* gpu_ms) {
    if (!w || !gpu_ms) return fail(SE_ERR_ARG, "null argument");
    DeviceGuard dg(w->ctx->device);
    hipEvent_t a, b;
    HIPCHECK(hipEventCreate(&a));
    HIPCHECK(hipEventCreate(&b));
    hipEventRecord(a, w->stream);
    int rc = mode == 0 ? se_world_step(w, nsteps) : blocked_steps(w, nsteps, sub);
    hipEventRecord(b, w->stream);
    hipError_t e = hipEventSynchronize(b);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    hipEventDestroy(a);
    hipEventDestroy(b);
    if (rc) return rc;
    if (e != hipSuccess) return fail(SE_ERR_HIP, hipGetErrorString(e));
    *gpu_ms = ms;
    return SE_OK;
}

// Picks the LDS variant for the world's next `nsteps` blocked steps by timing
// them: from a snapshot of the grid, frame and queued modifications (restored
// after every trial and at the end), each variant advances min(nsteps, 48)
// steps with sub-steps = its halo (a longer run is timed over whole launches:
// 48 is a multiple of every generated halo), `rounds` times; the fastest per
// step becomes the world's variant, as with se_world_set_variant.  Which variant
// wins depends on the step count: a deeper halo needs fewer launches for a short
// run (20 steps: two of 10 instead of 8 + 8 + 4) but costs more per step.
int se_world_autotune(se_world* w, uint32_t nsteps, uint32_t rounds, int32_t* best_variant, uint32_t* best_sub,
                      float* best_ms_per_step) {
    if (!w || nsteps == 0 || rounds == 0) return fail(SE_ERR_ARG, "bad arguments");
    if (w->color) return fail(SE_ERR_ARG, "autotune: render side-outputs are on (blocked steps run the plain kernel)");
    const uint32_t trial = std::min<uint32_t>(nsteps, 48u);
    if ((int64_t)w->frame + trial > INT32_MAX) return fail(SE_ERR_ARG, "frame counter would overflow");
    DeviceGuard dg(w->ctx->device);
    struct Resources {
        void* snap = nullptr;
        hipEvent_t ev[2] = {nullptr, nullptr};
        ~Resources() {
            if (snap) (void)hipFree(snap);
            for (auto e : ev)
                if (e) (void)hipEventDestroy(e);
        }
    } R;
    const size_t bytes = (size_t)w->pitch * w->H * 4;
    if (hipMalloc(&R.snap, bytes) != hipSuccess) return fail(SE_ERR_OOM, "autotune snapshot allocation failed");
    HIPCHECK(hipEventCreate(&R.ev[0]));
    HIPCHECK(hipEventCreate(&R.ev[1]));
    HIPCHECK(hipMemcpyAsync(R.snap, w->buf[w->cur], bytes, hipMemcpyDeviceToDevice, w->stream));
    const int32_t frame0 = w->frame;
    const int cur0 = w->cur, variant0 = w->lds_variant;
    const std::vector<se_modification> pending0 = w->pending;
    auto restore = [&]() -> int {
        w->frame = frame0;
        w->cur = cur0;
        w->pending = pending0;
        HIPCHECK(hipMemcpyAsync(w->buf[w->cur], R.snap, bytes, hipMemcpyDeviceToDevice, w->stream));
        HIPCHECK(hipStreamSynchronize(w->stream));
        se_rt::sleep_reset(w);
        return SE_OK;
    };
    double best = 1e300;
    int bv = -1;
    for (uint32_t r = 0; r < rounds; r++) {
        for (int vi = 0; vi < (int)w->mod->lds.size(); vi++) {
            w->lds_variant = vi;
            HIPCHECK(hipEventRecord(R.ev[0], w->stream));
            int rc = blocked_steps(w, trial, 0);
            hipError_t e = hipEventRecord(R.ev[1], w->stream);
            if (e == hipSuccess) e = hipEventSynchronize(R.ev[1]);
            float ms = 0;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, R.ev[0], R.ev[1]);
            const int rr = restore();
            if (rc || rr) {
                w->lds_variant = variant0;
                return rc ? rc : rr;
            }
            if (e != hipSuccess) {
                w->lds_variant = variant0;
                return fail(SE_ERR_HIP, hipGetErrorString(e));
            }
            if (ms / trial < best) {
                best = ms / trial;
                bv = vi;
            }
        }
    }
    w->lds_variant = bv;
    if (best_variant) *best_variant = bv;
    if (best_sub) *best_sub = (uint32_t)w->mod->lds[bv].first.hb;
    if (best_ms_per_step) *best_ms_per_step = (float)best;
    return SE_OK;
}

}  // extern "C"
