/*
 * sandengine_hip.h -- C ABI of the MI355X-native sandengine grid-update step.
 *
 * This is the drop-in boundary for the reference's simulation path:
 *   - the rule-language front end  sandengine-lang (parse_string,
 *     create_glsl_from_parser) -- sandengine-lang/src/parser.rs:93,
 *     sandengine-lang/src/lib.rs:17;
 *   - the step driver  sandengine_core::simulation::Simulation
 *     (new / run / modifications / params.frame) --
 *     sandengine-core/src/simulation.rs:97-253.
 * The reference has no C ABI (it is Rust calling GL through glium); each entry
 * point below names the reference interface it replaces.
 *
 * Conventions: every int-returning function returns SE_OK (0) or a negative
 * se_status; nothing aborts or unwinds across the boundary; se_last_error()
 * gives a thread-local message for the last failure.  Host buffers belong to
 * the caller; device buffers and compiled modules belong to the library.  A
 * se_world must not be used from two threads at once; a se_ctx is internally
 * synchronised.  Grids are u32 cell ids, row-major, W*H, y growing down.
 */
#ifndef SANDENGINE_HIP_H
#define SANDENGINE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct se_ctx se_ctx;
typedef struct se_ruleset se_ruleset;
typedef struct se_world se_world;
typedef struct se_batch se_batch;
typedef struct se_serve se_serve;

/* == SimModification, sandengine-core/src/simulation.rs:46-56 (32 bytes).
 * mod_shape: 0 circle, 1 square (MODSHAPE_*, simulation.rs:41-42). */
typedef struct {
    int32_t position[2];
    int32_t mod_shape;
    int32_t mod_size;
    int32_t mod_matID;
    int32_t _pad4[3];
} se_modification;

enum se_status {
    SE_OK = 0,
    SE_ERR_ARG = -1,     /* bad argument or size */
    SE_ERR_PARSE = -2,   /* rule file rejected (ParsingErr category in the message) */
    SE_ERR_CODEGEN = -3, /* rule set uses something the HIP backend cannot express */
    SE_ERR_JIT = -4,     /* hiprtc compile or module load failed */
    SE_ERR_HIP = -5,     /* HIP runtime error */
    SE_ERR_OOM = -6,     /* device or host allocation failed */
    SE_ERR_FULL = -7,    /* admission refused: HBM budget or slot limit */
    SE_ERR_STATE = -8    /* call not valid in the current state */
};

#define SE_MAX_MODIFICATIONS 256 /* MAX_MODIFICATIONS, simulation.rs:43 */

/* Thread-local text of the last error. */
const char* se_last_error(void);
const char* se_version(void);
void se_free(void* p);

/* ---------------------------------------------------------------- language (no GPU needed) */
/* parse_string (sandengine-lang/src/parser.rs:93-152); writes the IR as JSON
 * (malloc'd, free with se_free).  On a rule-file error returns SE_ERR_PARSE and
 * *kind_out (if given) names the ParsingErr category. */
int se_parse_ir_json(const char* yaml, size_t len, char** json_out, char** kind_out);
/* create_glsl_from_parser (sandengine-lang/src/lib.rs:17-147): the text of
 * gen/materials.glsl and gen/rules.glsl, returned instead of written to disk. */
int se_emit_glsl(const char* yaml, size_t len, char** materials_out, char** rules_out);
/* HIP backend: the translation unit hiprtc compiles for this rule set. */
int se_emit_hip(const char* yaml, size_t len, char** src_out);
/* Compiles the HIP backend's output with hiprtc for gfx950 (no device needed);
 * *code_bytes receives the code-object size. */
int se_jit_check(const char* yaml, size_t len, size_t* code_bytes);

/* ---------------------------------------------------------------- context */
/* One per device; owns the JIT cache and the default stream.
 * Replaces the GL context + ComputeShader::from_source (simulation.rs:128-137). */
int se_ctx_create(int device, se_ctx** out);
void se_ctx_destroy(se_ctx* ctx);
/* Device HBM bytes (total, free). */
int se_ctx_mem_info(se_ctx* ctx, size_t* total, size_t* free_bytes);

/* ---------------------------------------------------------------- rule sets */
/* parse + HIP codegen + hiprtc JIT (gfx950), cached by source hash.  Replaces
 * parse_string + create_glsl_from_parser + build_compute_shaders + shader
 * compile (src/lib.rs:9-27, simulation.rs:129-138). */
int se_ruleset_from_yaml(se_ctx* ctx, const char* yaml, size_t len, se_ruleset** out);
void se_ruleset_destroy(se_ruleset* rs);
/* Material id by name (ids follow YAML order after EMPTY 0, NULL 1, WALL 2;
 * sandengine-lang/src/parser/materials.rs:52-91). */
int se_ruleset_material_id(const se_ruleset* rs, const char* name, int32_t* id);
int se_ruleset_material_count(const se_ruleset* rs, int32_t* n);
/* How the LDS kernels evaluate a block for this rule set (diagnostic, no
 * reference counterpart): 0 the generated rule chain, 1 the full transition
 * table (N^4 x hash classes), 2 the coded table (tuples without NULL/WALL, the
 * mirror as a byte permutation, outcome vectors).  Results are identical. */
int se_ruleset_kernel_form(const se_ruleset* rs, int32_t* form);

/* ---------------------------------------------------------------- worlds */
/* Simulation::new (simulation.rs:128-192): frame 0, every cell EMPTY. */
int se_world_create(se_ctx* ctx, const se_ruleset* rs, uint32_t w, uint32_t h, se_world** out);
void se_world_destroy(se_world* w);
/* Host <-> device copies of the current grid (the reference has none; they make
 * seeded runs and checks possible).  n must equal w*h. */
int se_world_upload(se_world* w, const uint32_t* cells, size_t n);
int se_world_download(se_world* w, uint32_t* cells, size_t n);
/* Simulation::modifications.push (simulation.rs:125; sandengine-core/src/lib.rs:59-67):
 * queued for the next step only; more than SE_MAX_MODIFICATIONS queued -> SE_ERR_FULL. */
int se_world_push_modification(se_world* w, const se_modification* m);
/* nsteps x Simulation::run (simulation.rs:195-252): each does frame += 1 then
 * one step with that frame (frame 1 clears); queued modifications apply to the
 * first of them and are then cleared.  Plain kernel: one pass over HBM per step. */
int se_world_step(se_world* w, uint32_t nsteps);
/* Same result as se_world_step, advanced by the LDS temporal-blocked kernel in
 * launches of up to `sub` sub-steps (0 = autotuned choice). */
int se_world_step_blocked(se_world* w, uint32_t nsteps, uint32_t sub);
/* Kernel variant selection (no reference counterpart; the autotuner drives it):
 * plain_variant indexes the plain-kernel instantiations (rows per wavefront x
 * store kind), lds_variant the LDS-tile instantiations (-1: default).
 * se_variant_info(kind 0 plain: a=rows, b=non-temporal; kind 1 lds: a=tile w,
 * b=tile h, c=max sub-steps * 1000 + threads per workgroup; kind 2: a=the LDS
 * variant se_world_step_blocked uses for `index` sub-steps per launch; kind 3:
 * a=workgroups of LDS variant `index` resident at once on the device (the
 * batch's persistent launch shape), b=its LDS bytes per workgroup). */
int se_world_set_variant(se_world* w, int32_t plain_variant, int32_t lds_variant);
/* Tile sleeping of se_world_step_blocked (no reference counterpart; on by
 * default): an LDS tile whose cells, halo included, all hold one material whose
 * uniform 2x2 block no rule changes skips its sub-steps, and when it and its
 * eight neighbours were already so in the previous launch it also skips its
 * loads and stores (its output buffer already holds those cells).  Results are
 * identical either way; any other write to the grids restarts the record. */
int se_world_set_sleep(se_world* w, int enable);
int se_variant_info(const se_world* w, int32_t kind, int32_t index, int32_t* a, int32_t* b, int32_t* c);
/* Render side-outputs of every step: the colour and light images that
 * setCell writes beside the grid (shaders/compute/operations.glsl:99-170;
 * Simulation::output_color / output_light, simulation.rs:109-114).  Enable before
 * the first step to match the reference, whose light starts at zero
 * (simulation.rs:180-181) and then propagates one cell per step.  While
 * enabled, every step runs the plain kernel and then the render kernel (the
 * light needs each step's grid), and the world cannot join a batch.  Images are
 * RGBA f32 per cell, W*H*4 floats row-major on download.  Colour noise uses this
 * platform's sinf, so it matches a GL implementation only to within its sin
 * error; the light is plain f32 arithmetic in the shader's order.  The light
 * returned is the one the last step computed (after Simulation::run's swap the
 * reference holds it in input_light; the texture named output_light, which its
 * renderer samples, is then one step older). */
int se_world_set_render(se_world* w, int enable);
int se_world_render_download(se_world* w, float* color, float* light);
int se_world_render_ptrs(se_world* w, void** color, void** light, uint32_t* pitch);
/* Timing helper, no reference counterpart (the reference renders inside every
 * step): runs the render kernel n more times on the last step's two grids,
 * between HIP events on the world's stream, and returns the elapsed GPU time.
 * The grid is unchanged; the light advances n propagation steps on it and the
 * colour is rewritten.  Requires render outputs enabled and one step taken. */
int se_world_render_timed(se_world* w, uint32_t n, float* gpu_ms);
/* World save/load (the reference lists "World loading/ unloading -- using
 * chunks, save all relevant textures" as a to-do, README.md:85-86): the grid,
 * the frame and the queued modifications, in a binary file written and read in
 * chunks of rows.  RAND is a function of (block origin, frame) only, so a loaded
 * world continues exactly as the saved one would have.  Layout: 40-byte header
 * {char magic[8] = "SEWORLD1"; u32 version = 1, W, H; i32 frame; u32
 * n_modifications, reserved; u64 FNV-1a 64 of the rule set's IR JSON}, then
 * W*H u32 cell ids (row-major, little-endian), then n_modifications records of
 * se_modification.  se_world_load creates the world; a file saved with another
 * rule set, a short file or a bad header is refused (SE_ERR_ARG). */
int se_world_save(se_world* w, const char* path);
int se_world_load(se_ctx* ctx, const se_ruleset* rs, const char* path, se_world** out);
/* Params.frame (simulation.rs:78, 201). Negative frames are rejected. */
int se_world_get_frame(const se_world* w, int32_t* frame);
int se_world_set_frame(se_world* w, int32_t frame);
int se_world_sync(se_world* w);
/* Device view of the current grid: pointer, row pitch in cells. */
int se_world_device_ptr(se_world* w, void** ptr, uint32_t* pitch);
/* Seeded i.i.d. fill on the device (benchmark workloads): cell (x,y) draws
 * r = hi32(splitmix64((seed << 40) ^ (y*w + x))) and takes ids[k] for the first k
 * with r < thresholds[k]; rows y >= h - stone_rows become stone_id. */
int se_world_fill_seeded(se_world* w, uint64_t seed, const uint32_t* ids, const uint32_t* thresholds, int n,
                         uint32_t stone_rows, uint32_t stone_id);
/* Runs nsteps (mode 0 plain, 1 blocked with `sub`) between two HIP events on
 * the world's stream; *gpu_ms receives their elapsed time. */
int se_world_step_timed(se_world* w, uint32_t nsteps, int mode, uint32_t sub, float* gpu_ms);
/* No reference counterpart (Simulation::run dispatches one fixed shader,
 * simulation.rs:175).  Chooses the LDS variant for the world's next `nsteps`
 * blocked steps by timing min(nsteps, 48) of them with every variant, from a
 * snapshot of grid, frame and queued modifications that is restored after each
 * trial; the fastest becomes the world's variant (as se_world_set_variant), and
 * *best_sub receives its halo, the sub-steps se_world_step_blocked(w, n, 0)
 * then runs per launch.  Refused while render side-outputs are on. */
int se_world_autotune(se_world* w, uint32_t nsteps, uint32_t rounds, int32_t* best_variant, uint32_t* best_sub,
                      float* best_ms_per_step);
/* `reps` streaming copies of the grid between the world's two buffers (the
 * same bytes a plain step reads and writes, no rules): the measured HBM
 * reference for the plain kernel.  Leaves the grid contents unchanged. */
int se_world_copy_timed(se_world* w, uint32_t reps, float* gpu_ms);

/* ---------------------------------------------------------------- batched worlds
 * No reference counterpart (each reference Simulation is one world,
 * simulation.rs:97-126); every world in a batch takes exactly the steps
 * se_world_step would give it. */
int se_batch_create(se_ctx* ctx, se_batch** out);
void se_batch_destroy(se_batch* b);
/* The world advances `nsteps` steps in each following se_batch_run. */
int se_batch_add(se_batch* b, se_world* w, uint32_t nsteps);
int se_batch_clear(se_batch* b);
/* mode 0: loop the plain kernel world by world (one launch per step);
 * modes 1-3: the batched LDS kernel -- worlds that fit one tile advance
 * through all their steps in one work item, larger ones in rounds of k
 * sub-steps, rule-set groups and the two kinds overlapping on separate streams.
 * Its launch shape: mode 2 one workgroup per work item, mode 3 persistent (as
 * many workgroups as are resident, striding over the items), mode 1 the
 * autotuned shape (else one workgroup per item).
 * lds_variant -1 / k 0: the autotuned (else default) choice; graph != 0
 * captures the whole submission in a hipGraph and launches that.
 * *gpu_ms: time between events around the submission. */
int se_batch_run(se_batch* b, int mode, int32_t lds_variant, uint32_t k, int graph, float* gpu_ms);
/* Autotuner: for each rule-set group, times every LDS variant with k = halo and
 * k = halo/2, in both launch shapes, on `rounds` rounds of the batch's own
 * worlds (grids and frames are snapshotted and restored) and keeps the fastest
 * for later runs.  The outputs describe the last group tuned. */
int se_batch_autotune(se_batch* b, uint32_t rounds, int32_t* best_variant, uint32_t* best_k,
                      int32_t* best_persistent, float* best_ns_per_kcell);
/* The tuned choice for one rule set's group (after se_batch_autotune):
 * SE_ERR_STATE when that group has not been tuned. */
int se_batch_tuned(const se_batch* b, const se_ruleset* rs, int32_t* variant, uint32_t* k, int32_t* persistent);

/* ---------------------------------------------------------------- serving loop
 * No reference counterpart (the reference steps one interactive world per
 * frame, sandengine-core/src/lib.rs:57-68).  Requests arrive by a seeded
 * Poisson process; each is a square world with a rule set, a seeded i.i.d.
 * fill, a step count and circle modifications (brushes of a random material,
 * explosions of EMPTY) at given steps.  Admission is FIFO by HBM budget (one
 * arena, no allocation per request); admitted worlds advance together in rounds
 * (continuous batching) and each result equals the world stepped alone.
 * With slo_ms > 0 admission is also by deadline: the loop predicts each
 * request's completion from the rounds it needs and the round time (a fixed
 * part plus the round's work over a capacity, both fitted online to the
 * rounds' GPU time); a queued request is admitted only when it and every
 * admitted request are predicted to finish within arrival + 0.85 slo_ms (the
 * rest is headroom for host delays and model error), waits while it could
 * still make that, and is shed (state 4) once it cannot. */
typedef struct {
    double rate;              /* requests per second */
    double duration_s;        /* arrival window */
    uint64_t seed;
    uint32_t side_min, side_max;   /* side log-uniform, rounded to even */
    uint32_t steps_min, steps_max; /* steps log-uniform */
    uint32_t mods_max;        /* modifications per request uniform in [0, mods_max] */
    uint32_t quantum;         /* steps per round for multi-tile worlds (capped at the tile halo); 0 = 8 */
    uint32_t quantum_small;   /* steps per round for worlds inside one LDS tile; 0 = 64 */
    int kernel;               /* 1 batched LDS kernel, 0 plain kernel looped per world */
    int max_inflight;         /* rounds queued on the device at once; 0 = 2 */
    double drain_timeout_s;   /* stop this long after the last arrival (0 = run to completion) */
    size_t hbm_budget;        /* bytes of world state admitted at once */
    uint32_t verify_every;    /* keep a hash of every n-th request's final grid (0 = none) */
    double slo_ms;            /* deadline admission (see above); 0 = FIFO by HBM budget only */
    double capacity_gcups;    /* starting value of the online capacity estimate (G per second of the
                                 round model's work unit: cell-updates of whole LDS tiles for the batched
                                 kernel, cell-updates for the plain one); 0 = 1500 */
} se_serve_config;

typedef struct {
    uint64_t requests, completed, rejected, unfinished, rounds;
    double p50_ms, p99_ms, mean_ms, max_ms; /* arrival -> completion, completed requests */
    double wall_s, cell_updates, gcups;     /* completed work over the wall time */
    double offered_gcups;                   /* arriving work per second of the arrival window */
    size_t peak_bytes;                      /* arena high-water mark */
    double processed_gcups;                 /* all steps run (also of unfinished requests) over wall */
    double p50_ms_censored, p99_ms_censored;/* over every arrived request, unfinished ones at the end time */
    uint64_t shed, verified;                /* requests shed by deadline admission; final grids hashed */
    double shed_cell_frac;                  /* shed requests' share of the offered cell-updates */
    double slo_met_frac;                    /* completed requests done within slo_ms (1 when slo_ms is 0) */
    double capacity_gcups_est;              /* the capacity estimate at the end of the run (same unit) */
    double round_fixed_ms;                  /* the fixed part of the round-time model */
    double pred_ratio_p50, pred_ratio_p99;  /* completed requests' latency over the latency predicted at
                                               admission (deadline admission; 0 otherwise) */
} se_serve_stats;

typedef struct {
    double arrival_s, done_s;
    uint32_t side, steps, ruleset, nmods;
    uint64_t fill_seed;
    int state;    /* 0 waiting, 1 admitted, 2 done, 3 rejected (larger than the budget), 4 shed (deadline) */
    int verified; /* hash holds FNV-1a 64 of the final grid (row-major u32) */
    uint64_t hash;
} se_serve_request;

int se_serve_create(se_ctx* ctx, const se_ruleset* const* rulesets, int n_rulesets, const se_serve_config* cfg,
                    se_serve** out);
void se_serve_destroy(se_serve* s);
/* The i.i.d. fill of requests that use rule set `rs` (see se_world_fill_seeded). */
int se_serve_set_fill(se_serve* s, int rs, const uint32_t* ids, const uint32_t* thresholds, int n);
/* Runs the arrival window (and the drain); once per se_serve (SE_ERR_STATE after). */
int se_serve_run(se_serve* s, se_serve_stats* stats);
uint64_t se_serve_request_count(const se_serve* s);
int se_serve_request_info(const se_serve* s, uint64_t i, se_serve_request* out);
int se_serve_request_mod(const se_serve* s, uint64_t i, uint32_t j, uint32_t* step, se_modification* m);
/* FNV-1a 64 of a byte range (the hash se_serve_request.hash uses). */
uint64_t se_fnv1a64(const void* data, size_t bytes);

/* ---------------------------------------------------------------- batched step without allocation
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, serving loops that carve worlds from
 * their own arena, harnesses that forbid hipMalloc: the batched LDS step of
 * se_batch_run (same kernels, same results) over grids the caller owns.  The
 * per-call device data (launch records, tile prefix sums, modifications) is
 * staged in a fixed array of the library, SE_NOALLOC_SCRATCH_BYTES per device,
 * located when a context is created on the device, through a pinned host
 * buffer of the context; the call allocates nothing.  Limits: at most
 * SE_NOALLOC_MAX_GRIDS grids and SE_NOALLOC_MAX_MODS modifications per call
 * (SE_ERR_ARG beyond), one rule set per call.  Every call on a device shares
 * its array and overwrites it, so calls that overlap on one device (other
 * threads, or other streams before the previous call's work has run) race;
 * calls on one stream are ordered. */
typedef struct {
    uint32_t* grid[2];  /* device buffers of pitch*h u32 cells, row-major; the caller owns them */
    uint32_t w, h, pitch;
    int32_t cur;        /* in: the buffer holding the current grid; out: the one holding the result */
    int32_t frame;      /* in: Params.frame before the steps; out: after them */
    uint32_t nsteps;    /* steps (Simulation::run calls) this grid advances */
    uint32_t mod_first; /* modifications for the grid's first step: mods[mod_first .. mod_first + mod_count) */
    uint32_t mod_count; /* at most SE_MAX_MODIFICATIONS; cut at the first of size 0, as se_world_step does */
} se_grid_step;
#define SE_NOALLOC_MAX_GRIDS 2048
#define SE_NOALLOC_MAX_MODS 2048
#define SE_NOALLOC_SCRATCH_BYTES (224 * 1024)
/* Enqueues the steps of n grids on `stream` (a hipStream_t; NULL: the
 * context's stream) and updates grids[i].cur and .frame; returns without
 * waiting for the device.  A grid whose next frame is the clearing frame 1
 * takes that step first (its modifications are dropped, as there). */
int se_grids_step_noalloc(se_ctx* ctx, const se_ruleset* rs, se_grid_step* grids, uint32_t n,
                          const se_modification* mods, uint32_t nmods, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* SANDENGINE_HIP_H */
