// Internal types and helpers shared by the runtime translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sandengine_hip.h"
#include "se_hipgen.hpp"
#include "se_lang.hpp"

namespace se_rt {

// Records the message for se_last_error() and returns `code`.
int fail(int code, const std::string& msg);

#define HIPCHECK(expr)                                                                                 \
    do {                                                                                               \
        hipError_t _e = (expr);                                                                        \
        if (_e != hipSuccess) return se_rt::fail(SE_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

// Launches `f` on a 1-D grid of 256-thread workgroups; each argument is passed
// by address (kernelParams), so its C++ type must match the kernel parameter.
template <typename... A>
int launch(hipFunction_t f, unsigned grid, hipStream_t s, A... args) {
    void* p[] = {(void*)&args...};
    hipError_t e = hipModuleLaunchKernel(f, grid, 1, 1, 256, 1, 1, 0, s, p, nullptr);
    if (e != hipSuccess) return se_rt::fail(SE_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return SE_OK;
}

template <typename... A>
int launch_ex(hipFunction_t f, unsigned grid, unsigned block, unsigned shmem, hipStream_t s, A... args) {
    void* p[] = {(void*)&args...};
    hipError_t e = hipModuleLaunchKernel(f, grid, 1, 1, block, 1, 1, shmem, s, p, nullptr);
    if (e != hipSuccess) return se_rt::fail(SE_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return SE_OK;
}

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        hipGetDevice(&prev);
        if (prev != dev) hipSetDevice(dev);
    }
    ~DeviceGuard() {
        int now;
        hipGetDevice(&now);
        if (prev >= 0 && now != prev) hipSetDevice(prev);
    }
};


}  // namespace se_rt

// ------------------------------------------------------------------ internal types
struct se_module {
    hipModule_t mod = nullptr;
    uint32_t* table = nullptr;       // transition table (SE_TBL), device memory
    size_t table_entries = 0;        // 0 when the rule set runs the generated chain
    int table_form = 0;              // 0 generated chain, 1 full table, 2 coded table
    bool wall_kept = false;          // full table: no entry changes a WALL cell
    uint32_t stable_mask = 0;        // bit c: a block of four cells of LDS code c never changes
    std::vector<unsigned> lds_bytes; // dynamic LDS per LDS variant
    std::vector<int> lds_grid;       // resident workgroups per LDS variant at full occupancy (multiple of 8)
    std::vector<int> lds_waves;      // resident wavefronts per CU per LDS variant
    int default_lds = 0;             // LDS variant of se_world_step_blocked when none is set
    hipFunction_t plain = nullptr, fill = nullptr, fill_seeded = nullptr, copy = nullptr, fill_spec = nullptr;
    hipFunction_t render = nullptr;  // se_render: colour and light side-outputs of a step
    hipFunction_t noise_fill = nullptr;  // se_noise_fill: the colour noise image
    hipFunction_t plain_v[6] = {};  // se_step_plain_r{1,2,4}_nt{0,1}
    std::vector<std::pair<se::LdsVariant, hipFunction_t>> lds;
    std::vector<hipFunction_t> batch;  // se_batch_lds_* parallel to `lds`
    int device = 0;
    ~se_module() {
        if (mod) {
            int prev;
            hipGetDevice(&prev);
            hipSetDevice(device);
            hipFree(table);
            hipModuleUnload(mod);
            hipSetDevice(prev);
        }
    }
};

struct se_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    std::map<uint64_t, std::pair<std::string, std::shared_ptr<se_module>>> cache;
    // se_grids_step_noalloc: the device's scratch array, a pinned staging buffer
    // of its size, and the event after the last copy out of that buffer
    void* scratch = nullptr;
    void* staging = nullptr;
    hipEvent_t staging_done = nullptr;
    std::mutex staging_mu;
};

struct se_ruleset {
    se_ctx* ctx = nullptr;
    se::ParsingResult pr;
    std::shared_ptr<se_module> mod;
    int nmat = 0;
};

struct se_world {
    se_ctx* ctx = nullptr;
    const se_ruleset* rs = nullptr;
    std::shared_ptr<se_module> mod;
    uint32_t W = 0, H = 0, pitch = 0;
    uint32_t* buf[2] = {nullptr, nullptr};
    int cur = 0;
    int32_t frame = 0;
    std::vector<se_modification> pending;
    void* d_mods = nullptr;
    hipStream_t stream = nullptr;
    bool owns = true;       // buffers from hipMalloc (false: carved from a serving arena)
    int plain_variant = 3;  // index into se_module::plain_v (default: 1 row pair, non-temporal stores)
    int lds_variant = -1;   // index into se_module::lds, -1 = default choice
    // render side-outputs (se_world_set_render): RGBA f32 per cell, row pitch `pitch`
    void* light[2] = {nullptr, nullptr};  // ping-pong: light[lcur] holds the last step's light
    float* noise = nullptr;                // colour noise per cell (position only), filled once
    void* color = nullptr;
    int lcur = 0;
    // tile sleeping of the LDS kernel (se_tile_sleeps): per tile of variant
    // sleep_variant, the settled code in the last launch or 0xff; valid only while
    // nothing but that kernel has written either grid buffer (sleep_variant -1
    // otherwise)
    unsigned char* d_sleep[2] = {nullptr, nullptr};
    size_t sleep_cap = 0;
    int sleep_variant = -1;
    int sleep_p = 0;
    bool sleep_enabled = true;
};


namespace se_rt {
// One Simulation::run of `w` with the plain kernel (frame += 1, mods, step).
int plain_step(se_world* w);
// The grids of `w` were written by something other than the LDS kernel: its
// tile-sleeping record no longer describes them.
inline void sleep_reset(se_world* w) { w->sleep_variant = -1; }
// Index of the default LDS tile variant of a module (batches and serving use it
// unless tuned).
int default_lds_index(const se_module* m);
// Device address of the no-allocation scratch array on `device` (current).
int noalloc_scratch(int device, void** out);
}  // namespace se_rt
