"""Benchmark of the sandengine grid-update step on one MI355X.

Default workload: one 16384x16384 world (u32 cells, 1 GiB per buffer), the
four-material rule set, seed 1, bottom 30% of rows stone and the rest i.i.d.
60% EMPTY / 25% sand / 15% water; --steps timed steps after --warmup untimed
ones.  Both step paths run from the same initial grid:
  plain   : one HBM pass per step (se_step_plain)
  blocked : LDS temporal blocking, several sub-steps per launch (se_step_lds_*)
and their final grids must be identical (checked after the timed region).

Prints ONE JSON line: value = whole-job Gcell-updates/s of the faster path,
roofline of that path's kernel (algorithmic bytes: 8 B per cell per launch --
one u32 read and one u32 write), per-step latency p50/p99 for both paths, and
the CPU baseline (the reference's GLSL program under Mesa llvmpipe on a bounded
sample, or the oracle's C restatement where the llvmpipe build is absent).
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (side, ruleset, seed, ids, percents, stone fraction (num, den), stone id)
    "single16k": (16384, "four_materials", 1, [0, 3, 4], [60, 25, 15], (3, 10), 5),
    "single256": (256, "four_materials", 0, [0, 3, 4, 5], [50, 25, 15, 10], (0, 1), 0),
    "rich8k": (8192, "twelve_materials", 2, [0] + list(range(3, 15)), [40] + [5] * 12, (0, 1), 0),
}
BATCH = {
    # name: (n worlds, side range, step range, seed)
    "batch4096": (4096, (64, 1024), (100, 1000), 3),
    "batch512": (512, (64, 1024), (100, 1000), 3),
}
BATCH_FILLS = {"four_materials": ([0, 3, 4, 5], [50, 25, 15, 10]),
               "twelve_materials": ([0] + list(range(3, 15)), [40] + [5] * 12)}
DESCR = {
    "single16k": "single world 16384x16384 u32, 4 materials, seed 1, bottom 30% stone, 60/25/15 empty/sand/water",
    "single256": "single world 256x256 u32, 4 materials, seed 0, 50/25/15/10 empty/sand/water/stone",
    "rich8k": "single world 8192x8192 u32, 12 materials with SET/RAND, seed 2, 40% empty + 12x5%",
    "batch4096": "4096 independent worlds, sides log-uniform in [64,1024] rounded to even, 4- or 12-material "
                 "rules (coin flip), i.i.d. fills, steps uniform in [100,1000], seed 3",
    "batch512": "512 independent worlds drawn like batch4096 (first 512 of the same draw)",
}


def batch_specs(name):
    """World parameters of a batched workload, from one seeded generator."""
    n, (lo, hi), (slo, shi), seed = BATCH[name]
    rng = np.random.RandomState(seed)
    specs = []
    for i in range(BATCH["batch4096"][0]):
        side = float(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        side = int(round(side / 2.0)) * 2
        rs = "four_materials" if rng.rand() < 0.5 else "twelve_materials"
        steps = int(rng.randint(slo, shi + 1))
        specs.append((side, rs, steps, seed * 100000 + i))
    return specs[:n]


def run_serve(args):
    """Serving loop (Poisson arrivals, HBM-budget admission, continuous batching):
    p50/p99 request latency and aggregate throughput, batched LDS kernel vs the
    plain kernel looped per world."""
    import sandengine_amd as S
    ctx = S.Context(0)
    rsets = [S.RuleSet(ctx, S.bundled_ruleset(k)) for k in BATCH_FILLS]
    fills = [BATCH_FILLS[k] for k in BATCH_FILLS]
    out_runs = {}
    for kernel in [int(k) for k in args.serve_kernels.split(",")]:
        rate = args.rate if kernel == 1 else args.plain_rate
        dur = args.duration if kernel == 1 else args.plain_duration
        srv = S.Serve(ctx, rsets, fills, rate=rate, duration_s=dur, seed=4, side_min=128, side_max=4096,
                      steps_min=10, steps_max=5000, mods_max=20, kernel=kernel, max_inflight=2,
                      drain_timeout_s=args.drain, hbm_budget=int(args.budget_gb * (1 << 30)), verify_every=0)
        t0 = time.perf_counter()
        st = srv.run()
        st["rate"] = rate
        st["duration_s"] = dur
        st["offered_over_processed"] = st["offered_gcups"] / max(st["processed_gcups"], 1e-9)
        name = "batched_lds" if kernel == 1 else "plain_looped"
        out_runs[name] = st
        log(f"{name}: rate {rate}/s for {dur}s: {st['completed']}/{st['requests']} done "
            f"({st['unfinished']} unfinished), p50 {st['p50_ms']:.0f} ms, p99 {st['p99_ms']:.0f} ms, "
            f"{st['processed_gcups']:.0f} Gcell-upd/s processed ({st['gcups']:.0f} in completed requests) vs "
            f"{st['offered_gcups']:.0f} offered, censored p50/p99 {st['p50_ms_censored']:.0f}/"
            f"{st['p99_ms_censored']:.0f} ms, rounds {st['rounds']}, "
            f"peak {st['peak_bytes'] / 2**30:.1f} GiB, {time.perf_counter() - t0:.1f}s")
        srv.close()
    best = out_runs.get("batched_lds") or next(iter(out_runs.values()))
    out = {
        "metric": "Gcell-updates/s & HBM-roofline %; p50/p99 step latency, batched vs plain",
        "value": best["processed_gcups"], "unit": "Gcell-updates/s", "n_gpus": 1, "steps": None, "warmup": 0,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic (seeded Poisson arrivals and SplitMix64 fills)",
        "config": {"workload": "serving: Poisson arrivals, side log-uniform [128,4096], 4- or 12-material rules, "
                               "steps log-uniform [10,5000], 0-20 circle modifications (brush r 4-64 / "
                               "explosion r 8-128), seed 4, HBM admission budget %.0f GiB" % args.budget_gb},
        "runs": out_runs,
    }
    print(json.dumps(out))


def run_batch(args):
    """Batched multi-world workload: autotuned batched LDS kernel (one hipGraph)
    against looping the plain kernel world by world, from identical grids."""
    import sandengine_amd as S
    ctx = S.Context(0)
    rsets = {k: S.RuleSet(ctx, S.bundled_ruleset(k)) for k in BATCH_FILLS}
    specs = batch_specs(args.workload)
    t0 = time.perf_counter()
    fast, slow = [], []
    for side, rs, steps, fseed in specs:
        ids, pct = BATCH_FILLS[rs]
        for lst in (fast, slow):
            w = S.World(ctx, rsets[rs], side, side)
            w.frame = 1
            w.fill_seeded(fseed, ids, pct)
            lst.append(w)
    cells = float(sum(side * side * steps for side, _, steps, _ in specs))
    log(f"{len(specs)} worlds created and filled in {time.perf_counter() - t0:.1f}s; "
        f"{cells:.3e} cell-updates")
    bf, bs = S.Batch(ctx), S.Batch(ctx)
    for (side, rs, steps, _), wf, ws in zip(specs, fast, slow):
        bf.add(wf, steps)
        bs.add(ws, steps)
    t0 = time.perf_counter()
    bf.autotune(rounds=args.tune_rounds)
    tune_s = time.perf_counter() - t0
    tuned_by_rules = {}
    for name, rs in rsets.items():
        vi, k, persistent = bf.tuned(rs)
        w0 = next(w for (_, r, _, _), w in zip(specs, fast) if r == name)
        info = w0.variant_info(1, vi)
        tuned_by_rules[name] = {"variant": vi, "tile": [info[0], info[1]], "halo": info[2] // 1000,
                                "threads": info[2] % 1000, "k": k,
                                "launch": "persistent" if persistent else "one workgroup per work item"}
        log(f"autotune {name}: variant {vi} {info} k={k} persistent={persistent}")
    log(f"autotune: {tune_s:.1f}s")
    t0 = time.perf_counter()
    ms_batch = bf.run(mode=1, graph=True)
    wall_batch = time.perf_counter() - t0
    t0 = time.perf_counter()
    ms_loop = bs.run(mode=0)
    wall_loop = time.perf_counter() - t0
    log(f"batched: {ms_batch:.1f} ms GPU ({wall_batch:.2f}s wall incl. plan+graph); "
        f"looped plain: {ms_loop:.1f} ms GPU ({wall_loop:.2f}s wall)")
    identical = True
    if not args.no_check:
        for wf, ws in zip(fast, slow):
            if wf.frame != ws.frame or not np.array_equal(wf.download(), ws.download()):
                identical = False
                break
        log(f"batched == looped plain for every world: {identical}")
    out = {
        "metric": "Gcell-updates/s & HBM-roofline %; p50/p99 step latency, batched vs plain",
        "value": cells / (wall_batch) / 1e9,
        "unit": "Gcell-updates/s",
        "n_gpus": 1, "steps": int(sum(s[2] for s in specs)), "warmup": 0,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32", "data": "synthetic (seeded SplitMix64 fills generated on the device)",
        "config": {"workload": DESCR[args.workload], "path": "batched LDS kernel, autotuned, one hipGraph",
                   "tuned_variant": tuned_by_rules},
        "batched": {"gpu_ms": ms_batch, "wall_s": wall_batch, "gcups_gpu": cells / (ms_batch / 1e3) / 1e9,
                    "gcups_wall": cells / wall_batch / 1e9, "autotune_s": tune_s},
        "looped_plain": {"gpu_ms": ms_loop, "wall_s": wall_loop, "gcups_gpu": cells / (ms_loop / 1e3) / 1e9,
                         "gcups_wall": cells / wall_loop / 1e9},
        "speedup_batched_vs_looped_plain": wall_loop / wall_batch,
        "cell_updates": cells,
        "parity": {"identical": identical},
    }
    print(json.dumps(out))
    if not identical:
        sys.exit(1)


# se_render per cell: id after and before the step (4 + 4), light in (16), colour
# noise (4, precomputed per world), light and colour out (16 + 16)
RENDER_BYTES_PER_CELL = 60


def run_render(args):
    """Grid step plus the render side-outputs of every step (colour noise and
    light propagation) on the 16384^2 four-material world: the plain step kernel
    then se_render, per step.  se_render dominates; its roofline uses 60
    algorithmic bytes per cell and its launch time is measured alone with HIP
    events.  The grid must equal a world stepped without render outputs."""
    import sandengine_amd as S
    ctx = S.Context(0)
    side, rsname = WORKLOADS["single16k"][:2]
    rs = S.RuleSet(ctx, S.bundled_ruleset(rsname))
    cells = side * side
    wc = make_world(S, ctx, rs, "single16k")  # HBM reference: a streaming copy of the grid
    wc.copy_timed(4)
    copy_ms = wc.copy_timed(args.copy_reps) / args.copy_reps
    wc.close()
    copy_gbs = 8.0 * cells / (copy_ms / 1e3) / 1e9
    w = make_world(S, ctx, rs, "single16k")
    w.set_render(True)
    wall, gpu_ms = time_path(S, w, "plain", args.steps, args.warmup, 0)
    reps = max(4, args.copy_reps)
    w.render_timed(2)
    render_ms = w.render_timed(reps) / reps
    step_ms = gpu_ms / args.steps
    p50, p99 = latencies(w, "plain", 0, args.latency_samples)
    log(f"step+render: {cells * args.steps / wall / 1e9:.1f} Gcell-upd/s wall, {step_ms:.3f} ms/step by events; "
        f"se_render alone {render_ms:.3f} ms/launch -> {RENDER_BYTES_PER_CELL * cells / (render_ms / 1e3) / 1e9:.0f} "
        f"GB/s algorithmic")
    identical = None
    if not args.no_check:
        wp = make_world(S, ctx, rs, "single16k")
        wp.step(w.frame - wp.frame)
        identical = bool(np.array_equal(wp.download(), w.download()))
        wp.close()
        log(f"grid with render outputs == grid without: {identical}")
    w.close()
    kernel = "se_render"
    prof = load_profile(kernel, args.workload)
    achieved = RENDER_BYTES_PER_CELL * cells / (render_ms / 1e3) / 1e9
    cpu = None if args.no_cpu_baseline else cpu_baseline("single16k")
    out = {
        "metric": "Gcell-updates/s & HBM-roofline %; p50/p99 step latency, batched vs plain",
        "value": cells * args.steps / wall / 1e9, "unit": "Gcell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u32 grid, f32 colour and light",
        "data": "synthetic (seeded SplitMix64 fill generated on the device)",
        "config": {"workload": DESCR["single16k"] + "; colour and light side-outputs of every step",
                   "path": "plain step + se_render per step", "rules": rsname, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": prof.get("bytes_per_launch"),
                     "algorithmic_bytes_per_launch": RENDER_BYTES_PER_CELL * cells, "avg_launch_ms": render_ms},
        "cpu_baseline": cpu,
        "copy_reference": {"gbs": copy_gbs, "ms": copy_ms, "frac_of_peak": copy_gbs / HBM_PEAK_GBS},
        "paths": {"step_and_render": {"gpu_ms_per_step": step_ms, "render_ms_per_launch": render_ms,
                                      "step_latency_ms": {"p50": p50, "p99": p99}}},
        "parity": {"grid_identical_without_render": identical},
    }
    print(json.dumps(out))
    if identical is False:
        sys.exit(1)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_world(S, ctx, rs, wl):
    side, _, seed, ids, pct, (num, den), stone = WORKLOADS[wl]
    w = S.World(ctx, rs, side, side)
    w.frame = 1  # the clearing frame has run; the seeded grid starts at frame 2
    w.fill_seeded(seed, ids, pct, stone_rows=(side * num) // den, stone_id=stone)
    return w


def time_path(S, w, mode, steps, warmup, sub):
    adv = (lambda n: w.step(n)) if mode == "plain" else (lambda n: w.step_blocked(n, sub))
    adv(warmup)
    w.sync()
    t0 = time.perf_counter()
    gpu_ms = w.step_timed(steps, 0 if mode == "plain" else 1, sub)
    wall = time.perf_counter() - t0
    return wall, gpu_ms


def latencies(w, mode, sub, n):
    """Per-step latency (ms): each sample is one synchronous step (or one blocked
    launch divided by its sub-steps)."""
    out = []
    for _ in range(n):
        if mode == "plain":
            out.append(w.step_timed(1, 0, 0))
        else:
            out.append(w.step_timed(sub, 1, sub) / sub)
    return np.percentile(out, 50), np.percentile(out, 99)


def cpu_baseline(wl, side=1024, steps=12, warm=2):
    """Bounded sample of the same workload on the host cores, in a child process."""
    code = f"""
import json, os, sys, time
sys.path.insert(0, {ROOT!r})
import numpy as np
from oracle import lang_ref as L, oracle as O, llvmpipe as LP
side, steps, warm = {side}, {steps}, {warm}
_, rsname, seed, ids, pct, (num, den), stone = {WORKLOADS[wl]!r}
text = open(os.path.join({ROOT!r}, 'sandengine_amd', 'rulesets', rsname + '.yaml')).read()
res = L.parse_string(text)
g = O.fill(seed, side, side, ids, pct, (side * num) // den, stone)
cores = min(16, os.cpu_count() or 1)
if LP.available():
    LP.lib(cores)
    w = LP.World(LP.program_for(res), side, side)
    w.run(1); w.upload(g)
    w.run(warm)
    dt = w.time_run(steps)
    kind, extra = 'reference', LP.info()
else:
    blob = O.compile_ruleset(res)
    os.environ['OMP_NUM_THREADS'] = str(cores)
    g = O.run(blob, g, 2, warm)
    t = time.perf_counter(); O.run(blob, g, 2 + warm, steps); dt = time.perf_counter() - t
    kind, extra = 'port', 'oracle C restatement (OpenMP)'
print(json.dumps(dict(value=side * side * steps / dt / 1e9, unit='Gcell-updates/s', cores=cores, kind=kind,
                      sample=f'{{side}}x{{side}} world of the same fill and rules, {{steps}} timed steps after '
                             f'{{warm}} warm-up; {{extra}}')))
"""
    env = dict(os.environ)
    env["LP_NUM_THREADS"] = str(min(16, os.cpu_count() or 1))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=600)
    if r.returncode != 0:
        log("cpu baseline failed:", r.stderr[-2000:])
        return None
    return json.loads(r.stdout.strip().splitlines()[-1])


def load_profile(kernel_prefix, workload):
    """Per-launch figures rocprofv3 measured for a kernel on a workload
    (profiles/traffic.json): HBM bytes (FETCH_SIZE + WRITE_SIZE, scaled by the
    copy-kernel calibration) and VALU / LDS wavefront instructions."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(p):
        return {}
    d = json.load(open(p))
    for k, v in d.get("workloads", {}).get(workload, {}).get("kernels", {}).items():
        if k.startswith(kernel_prefix):
            return v
    return {}


# One wave64 VALU instruction per cycle per CU (4 SIMD16 units), 256 CUs, 2.4 GHz: the
# issue rate measured on the MI355X for the forms the block loop is made of (VOP3, SDWA,
# 32-bit multiplies, VOP2 with a scalar or constant operand).  VOP2 ops with two VGPR
# sources can issue faster, so this is the ceiling for this loop, not for every kernel.
VALU_PEAK_WAVE_INSTR_PER_S = 256 * 2.4e9


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=24)
    ap.add_argument("--workload", default="single16k",
                    choices=sorted(WORKLOADS) + sorted(BATCH) + ["serve", "render16k"])
    ap.add_argument("--tune-rounds", type=int, default=2)
    ap.add_argument("--rate", type=float, default=2000.0, help="serve: requests/s (batched kernel)")
    ap.add_argument("--duration", type=float, default=60.0, help="serve: arrival window (s)")
    ap.add_argument("--plain-rate", type=float, default=200.0)
    ap.add_argument("--plain-duration", type=float, default=10.0)
    ap.add_argument("--drain", type=float, default=30.0, help="serve: stop this long after the last arrival")
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--serve-kernels", default="1,0")
    ap.add_argument("--sub", type=int, default=8, help="sub-steps per LDS launch")
    ap.add_argument("--paths", default="plain,blocked")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--latency-samples", type=int, default=50)
    ap.add_argument("--copy-reps", type=int, default=20)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (every box has one GPU). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    if args.workload in BATCH:
        return run_batch(args)
    if args.workload == "serve":
        return run_serve(args)
    if args.workload == "render16k":
        return run_render(args)
    import sandengine_amd as S
    ctx = S.Context(0)
    side, rsname = WORKLOADS[args.workload][:2]
    t0 = time.perf_counter()
    rs = S.RuleSet(ctx, S.bundled_ruleset(rsname))
    log(f"ruleset {rsname}: JIT {time.perf_counter() - t0:.2f}s")
    cells = side * side
    results, worlds = {}, {}
    # HBM reference: a streaming copy of the same grid (reads and writes 8 B per cell)
    wc = make_world(S, ctx, rs, args.workload)
    wc.copy_timed(4)
    copy_ms = wc.copy_timed(args.copy_reps) / args.copy_reps
    wc.close()
    copy_gbs = 8.0 * cells / (copy_ms / 1e3) / 1e9
    log(f"copy reference: {copy_ms:.3f} ms per 2x{4 * cells / 2**30:.2f} GiB -> {copy_gbs:.0f} GB/s")
    lds_kernel = None
    for mode in args.paths.split(","):
        w = make_world(S, ctx, rs, args.workload)
        if mode == "blocked":
            tw, th, c = w.variant_info(1, w.variant_info(2, args.sub)[0])
            lds_kernel = f"se_step_lds_{tw}x{th}_h{c // 1000}_t{c % 1000}"
        wall, gpu_ms = time_path(S, w, mode, args.steps, args.warmup, args.sub)
        launches = args.steps if mode == "plain" else -(-args.steps // args.sub)
        avg_launch_ms = gpu_ms / launches
        r = dict(gcups=cells * args.steps / wall / 1e9, gcups_gpu=cells * args.steps / (gpu_ms / 1e3) / 1e9,
                 ms_per_step=wall * 1e3 / args.steps, avg_launch_ms=avg_launch_ms, launches=launches,
                 steps_per_launch=1 if mode == "plain" else args.sub,
                 kernel="se_step_plain_r1_nt1" if mode == "plain" else lds_kernel)
        r["achieved_gbs"] = 8.0 * cells / (avg_launch_ms / 1e3) / 1e9
        r["roofline_frac"] = r["achieved_gbs"] / HBM_PEAK_GBS
        r["frac_of_copy"] = r["achieved_gbs"] / copy_gbs
        p50, p99 = latencies(w, mode, args.sub, args.latency_samples)
        r["step_latency_ms"] = {"p50": p50, "p99": p99}
        log(f"{mode}: {r['gcups']:.1f} Gcell-upd/s wall, {r['gcups_gpu']:.1f} by events, "
            f"{r['avg_launch_ms']:.3f} ms/launch, {r['achieved_gbs']:.0f} GB/s algorithmic")
        results[mode] = r
        worlds[mode] = w
    check = None
    if not args.no_check and "plain" in worlds and "blocked" in worlds:
        # both advanced warmup + steps + latency steps from the same grid; align frames
        a, b = worlds["plain"], worlds["blocked"]
        if a.frame < b.frame:
            a.step(b.frame - a.frame)
        elif b.frame < a.frame:
            b.step(a.frame - b.frame)
        ga = a.download()
        gb = b.download()
        check = dict(identical=bool(np.array_equal(ga, gb)), frame=a.frame)
        del ga, gb
        log(f"plain vs blocked after frame {a.frame}: identical={check['identical']}")
    for w in worlds.values():
        w.close()
    best = max(results, key=lambda m: results[m]["gcups"])
    R = results[best]
    kernel = "se_step_plain_r1_nt1" if best == "plain" else lds_kernel
    prof = load_profile(kernel, args.workload)
    traffic = prof.get("bytes_per_launch")
    valu = None
    if prof.get("valu_wave_instr_per_launch"):
        rate = prof["valu_wave_instr_per_launch"] / (R["avg_launch_ms"] / 1e3)
        valu = {"bound": "valu", "kernel": kernel, "achieved": rate, "peak": VALU_PEAK_WAVE_INSTR_PER_S,
                "unit": "wave64 VALU instructions/s", "frac": rate / VALU_PEAK_WAVE_INSTR_PER_S,
                "instr_per_launch": prof["valu_wave_instr_per_launch"],
                "source": "SQ_INSTS_VALU per launch from profiles/traffic.json over this run's launch time"}
    cpu = None if args.no_cpu_baseline else cpu_baseline(args.workload)
    out = {
        "metric": "Gcell-updates/s & HBM-roofline %; p50/p99 step latency, batched vs plain",
        "value": R["gcups"],
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": R["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded SplitMix64 fill generated on the device)",
        "config": {"workload": DESCR[args.workload], "path": best, "rules": rsname,
                   "steps_per_launch": R["steps_per_launch"], "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": R["achieved_gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": R["achieved_gbs"] / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": 8 * cells, "avg_launch_ms": R["avg_launch_ms"],
                     "binding_limit": "hbm" if best == "plain" else "valu issue and lds latency (see roofline_valu)"},
        "roofline_valu": valu,
        "cpu_baseline": cpu,
        "copy_reference": {"gbs": copy_gbs, "ms": copy_ms, "frac_of_peak": copy_gbs / HBM_PEAK_GBS},
        "speedup_blocked_vs_plain": (results["blocked"]["gcups"] / results["plain"]["gcups"]
                                     if "blocked" in results and "plain" in results else None),
        "paths": results,
        "parity": check,
    }
    print(json.dumps(out))
    if check is not None and not check["identical"]:
        sys.exit(1)


if __name__ == "__main__":
    main()
