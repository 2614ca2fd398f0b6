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


SERVE_MIX = dict(side_min=128, side_max=4096, steps_min=10, steps_max=5000, mods_max=20)


def replay_alone(S, ctx, rsets, fills, r):
    """FNV-1a 64 of a served request's final grid when the same request runs
    alone on the plain kernel (same fill, same modifications at the same steps)."""
    ids, pct = fills[r["ruleset"]]
    w = S.World(ctx, rsets[r["ruleset"]], r["side"], r["side"])
    w.frame = 1
    w.fill_seeded(r["fill_seed"], ids, pct)
    by_step = {}
    for step, m in r["mods"]:
        by_step.setdefault(step, []).append(m)
    done = 0
    for st in sorted(by_step) + [r["steps"]]:
        if st > done:
            w.step(st - done)
            done = st
        if st < r["steps"]:
            for m in by_step[st]:
                w.push_modification(*m)
    h = S.fnv1a64(w.download())
    w.close()
    return h


def serve_once(S, ctx, rsets, fills, args, kernel, rate, duration, drain, slo_ms, verify_every):
    srv = S.Serve(ctx, rsets, fills, rate=rate, duration_s=duration, seed=4, kernel=kernel, max_inflight=2,
                  drain_timeout_s=drain, hbm_budget=int(args.budget_gb * (1 << 30)), verify_every=verify_every,
                  slo_ms=slo_ms, **SERVE_MIX)
    t0 = time.perf_counter()
    st = srv.run()
    st.update(rate=rate, duration_s=duration, drain_s=drain, slo_ms=slo_ms or None,
              offered_over_processed=st["offered_gcups"] / max(st["processed_gcups"], 1e-9),
              censored=st["unfinished"] > 0)
    log(f"kernel {kernel} rate {rate}/s {duration}s slo {slo_ms or '-'}: {st['completed']}/{st['requests']} done, "
        f"{st['shed']} shed ({st['shed_cell_frac']:.2f} of the work), {st['unfinished']} unfinished, "
        f"p50 {st['p50_ms']:.0f} ms, p99 {st['p99_ms']:.0f} ms, {st['processed_gcups']:.0f} Gcell/s processed "
        f"({st['gcups']:.0f} completed) vs {st['offered_gcups']:.0f} offered, censored p99 "
        f"{st['p99_ms_censored']:.0f} ms, capacity est {st['capacity_gcups_est']:.0f} Gcell/s, round fixed "
        f"{st['round_fixed_ms']:.3f} ms, actual/predicted latency p50 {st['pred_ratio_p50']:.2f} p99 "
        f"{st['pred_ratio_p99']:.2f}, peak {st['peak_bytes'] / 2**30:.1f} GiB, {time.perf_counter() - t0:.1f}s")
    if verify_every:
        t1 = time.perf_counter()
        n = eq = 0
        for i in range(len(srv)):
            r = srv.request(i)
            if r["verified"]:
                n += 1
                eq += int(replay_alone(S, ctx, rsets, fills, r) == r["hash"])
        st["verified"] = {"every": verify_every, "requests": n, "equal_alone": eq,
                          "check": "final grid (FNV-1a 64) equals the request run alone on the plain kernel"}
        log(f"verified {eq}/{n} requests equal to their run alone ({time.perf_counter() - t1:.1f}s)")
    reqs = [srv.request(i) for i in range(min(len(srv), 64))]
    srv.close()
    return st, reqs


def run_serve(args):
    """Serving loop (Poisson arrivals, HBM-budget admission, continuous batching):
    p50/p99 request latency and aggregate throughput.  Runs, all on the served mix:
    the batched LDS kernel with FIFO admission at --rate; the same with deadline
    admission (--slo-ms); an arrival-rate sweep of both; the plain kernel looped
    per world at --plain-rate (censored when it cannot keep up)."""
    import sandengine_amd as S
    ctx = S.Context(0)
    names = list(BATCH_FILLS)
    rsets = [S.RuleSet(ctx, S.bundled_ruleset(k)) for k in names]
    fills = [BATCH_FILLS[k] for k in names]
    runs, sweep = {}, []
    ver = args.verify
    runs["batched_fifo"], reqs = serve_once(S, ctx, rsets, fills, args, 1, args.rate, args.duration, args.drain, 0, ver)
    if args.slo_ms > 0:
        runs["batched_deadline"], _ = serve_once(S, ctx, rsets, fills, args, 1, args.rate, args.duration, args.drain,
                                                 args.slo_ms, ver)
    for rate in [float(x) for x in args.sweep.split(",") if x]:
        for slo in ([0.0, args.slo_ms] if args.slo_ms > 0 else [0.0]):
            st, _ = serve_once(S, ctx, rsets, fills, args, 1, rate, args.sweep_duration, args.sweep_drain, slo, 0)
            sweep.append({k: st[k] for k in ("rate", "slo_ms", "requests", "completed", "shed", "unfinished",
                                              "p50_ms", "p99_ms", "p99_ms_censored", "slo_met_frac",
                                              "shed_cell_frac", "processed_gcups", "gcups", "offered_gcups",
                                              "offered_over_processed", "censored")})
    if args.plain_duration > 0:
        runs["plain_looped"], _ = serve_once(S, ctx, rsets, fills, args, 0, args.plain_rate, args.plain_duration,
                                             args.plain_drain, 0, 0)
    nb = min(16, len(reqs))
    cpu = None if args.no_cpu_baseline else cpu_baseline(
        [(names[r["ruleset"]], r["side"], r["fill_seed"], fills[r["ruleset"]][0], fills[r["ruleset"]][1], 0, 0)
         for r in reqs[:nb]], steps=8, warm=1,
        what=f"offline: the first {nb} requests of the same arrival sequence (sides, rule sets, fills), one after "
             f"another")
    best = runs["batched_fifo"]
    out = {
        "metric": "Gcell-updates/s & HBM-roofline %; p50/p99 step latency, batched vs plain",
        "value": best["processed_gcups"], "unit": "Gcell-updates/s", "n_gpus": 1, "steps": None, "warmup": 0,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic (seeded Poisson arrivals and SplitMix64 fills)",
        "config": {"workload": "serving: Poisson arrivals at %g/s for %g s, side log-uniform [128,4096], 4- or "
                               "12-material rules, steps log-uniform [10,5000], 0-20 circle modifications (brush "
                               "r 4-64 / explosion r 8-128), seed 4, HBM admission budget %.0f GiB"
                               % (args.rate, args.duration, args.budget_gb),
                   "value_is": "batched_fifo processed Gcell-updates/s (all steps run over the wall time)",
                   "slo_ms": args.slo_ms or None},
        "runs": runs,
        "rate_sweep": sweep,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    bad = [k for k, v in runs.items() if isinstance(v.get("verified"), dict)
           and v["verified"]["equal_alone"] != v["verified"]["requests"]]
    if bad:
        sys.exit(1)


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
        for lst in ((fast,) if args.no_looped else (fast, slow)):
            w = S.World(ctx, rsets[rs], side, side)
            w.frame = 1
            w.fill_seeded(fseed, ids, pct)
            lst.append(w)
    cells = float(sum(side * side * steps for side, _, steps, _ in specs))
    log(f"{len(specs)} worlds created and filled in {time.perf_counter() - t0:.1f}s; "
        f"{cells:.3e} cell-updates")
    bf, bs = S.Batch(ctx), S.Batch(ctx)
    for (side, rs, steps, _), wf in zip(specs, fast):
        bf.add(wf, steps)
    for (side, rs, steps, _), ws in zip(specs, slow):
        bs.add(ws, steps)
    t0 = time.perf_counter()
    tuned_by_rules = {}
    if args.tune_rounds > 0:
        bf.autotune(rounds=args.tune_rounds)
    tune_s = time.perf_counter() - t0
    for name, rs in rsets.items():
        w0 = next(w for (_, r, _, _), w in zip(specs, fast) if r == name)
        if args.tune_rounds > 0:
            vi, k, persistent = bf.tuned(rs)
        else:  # the defaults se_batch_run uses untuned
            vi, k, persistent = w0.variant_info(2, 0)[0], 0, False
        info = w0.variant_info(1, vi)
        tuned_by_rules[name] = {"variant": vi, "tile": [info[0], info[1]], "halo": info[2] // 1000,
                                "threads": info[2] % 1000, "k": k or info[2] // 1000,
                                "launch": "persistent" if persistent else "one workgroup per work item",
                                "tuned": args.tune_rounds > 0}
        log(f"{'autotune' if args.tune_rounds > 0 else 'default'} {name}: variant {vi} {info} k={k} "
            f"persistent={persistent}")
    log(f"autotune: {tune_s:.1f}s")
    t0 = time.perf_counter()
    ms_batch = bf.run(mode=1, graph=True)
    wall_batch = time.perf_counter() - t0
    ms_loop = wall_loop = None
    if not args.no_looped:
        t0 = time.perf_counter()
        ms_loop = bs.run(mode=0)
        wall_loop = time.perf_counter() - t0
        log(f"batched: {ms_batch:.1f} ms GPU ({wall_batch:.2f}s wall incl. plan+graph); "
            f"looped plain: {ms_loop:.1f} ms GPU ({wall_loop:.2f}s wall)")
    identical = True
    if not args.no_check and not args.no_looped:
        for wf, ws in zip(fast, slow):
            if wf.frame != ws.frame or not np.array_equal(wf.download(), ws.download()):
                identical = False
                break
        log(f"batched == looped plain for every world: {identical}")
    nb = min(64, len(specs))
    cpu = None if args.no_cpu_baseline else cpu_baseline(
        [(rs, side, fseed, BATCH_FILLS[rs][0], BATCH_FILLS[rs][1], 0, 0) for side, rs, _, fseed in specs[:nb]],
        steps=24, what=f"the first {nb} worlds of the same draw, one after another (one Simulation per world, as the "
             f"reference runs them)")
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
        "looped_plain": None if args.no_looped else {
            "gpu_ms": ms_loop, "wall_s": wall_loop, "gcups_gpu": cells / (ms_loop / 1e3) / 1e9,
            "gcups_wall": cells / wall_loop / 1e9},
        "speedup_batched_vs_looped_plain": None if args.no_looped else wall_loop / wall_batch,
        "cpu_baseline": cpu,
        "cell_updates": cells,
        "parity": {"identical": None if (args.no_check or args.no_looped) else identical,
                   "check": "every world's grid after the batched run equals the looped plain kernel's"},
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
    cpu = None if args.no_cpu_baseline else cpu_baseline(
        single_sample("single16k"), steps=40,
        what="2048x2048 world of the same fill and rules as single16k (the reference program computes colour and "
             "light in every step too)")
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


def host_info():
    """The host the CPU baseline runs on: logical CPUs, the CPUs this process may
    use, and the CPU model."""
    model = ""
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    return {"nproc": os.cpu_count(), "cpus_allowed": allowed, "cpu_model": model}


# Host threads for the CPU baseline: the CPUs this process may use, at most 16
# (the host share of one GPU on the measurement machines).
BASELINE_THREADS = max(1, min(16, host_info()["cpus_allowed"]))


def cpu_baseline(samples, steps=12, warm=2, what=""):
    """The reference's GLSL program under Mesa llvmpipe (or, where that build is
    absent, the oracle's C restatement) on a bounded sample of the workload, in a
    child process: `samples` are worlds (rule set, side, fill seed, ids,
    percents, stone rows, stone id), each run for `warm` untimed and `steps`
    timed steps; the value is their cell-updates over the summed timed runs."""
    code = f"""
import json, os, sys, time
sys.path.insert(0, {ROOT!r})
import numpy as np
from oracle import lang_ref as L, oracle as O, llvmpipe as LP
samples, steps, warm, threads = {samples!r}, {steps}, {warm}, {BASELINE_THREADS}
progs, blobs, cells, dt = {{}}, {{}}, 0.0, 0.0
use_lp = LP.available()
if use_lp:
    LP.lib(threads)
else:
    os.environ['OMP_NUM_THREADS'] = str(threads)
for rsname, side, seed, ids, pct, stone_rows, stone_id in samples:
    if rsname not in progs:
        res = L.parse_string(open(os.path.join({ROOT!r}, 'sandengine_amd', 'rulesets', rsname + '.yaml')).read())
        progs[rsname] = LP.program_for(res) if use_lp else None
        blobs[rsname] = O.compile_ruleset(res)
    g = O.fill(seed, side, side, ids, pct, stone_rows, stone_id)
    if use_lp:
        w = LP.World(progs[rsname], side, side)
        w.run(1); w.upload(g)
        w.run(warm)
        dt += w.time_run(steps)
    else:
        g = O.run(blobs[rsname], g, 2, warm)
        t = time.perf_counter(); O.run(blobs[rsname], g, 2 + warm, steps); dt += time.perf_counter() - t
    cells += float(side) * side * steps
kind, extra = ('reference', LP.info()) if use_lp else ('port', 'oracle C restatement (OpenMP)')
print(json.dumps(dict(value=cells / dt / 1e9, unit='Gcell-updates/s', cores=threads, kind=kind, seconds=dt,
                      engine=extra)))
"""
    env = dict(os.environ)
    env["LP_NUM_THREADS"] = str(BASELINE_THREADS)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=900)
    if r.returncode != 0:
        log("cpu baseline failed:", r.stderr[-2000:])
        return None
    out = json.loads(r.stdout.strip().splitlines()[-1])
    out["sample"] = (f"{what}; {warm} untimed + {steps} timed steps per world; reference dispatch geometry "
                     f"((W+7)/8, (H+7)/8) x 16x16; {out.pop('engine')}")
    out["host"] = dict(host_info(), threads_used=BASELINE_THREADS,
                       note="LP_NUM_THREADS = threads_used: the CPUs allowed to this process, at most 16 (the "
                            "host share of one GPU)")
    log(f"cpu baseline: {out['value']:.4f} Gcell-upd/s on {BASELINE_THREADS} threads ({out['seconds']:.1f} s timed)")
    return out


def single_sample(wl, side=2048):
    """A side x side world of the same rules and fill as workload `wl`."""
    _, rsname, seed, ids, pct, (num, den), stone = WORKLOADS[wl]
    return [(rsname, side, seed, ids, pct, (side * num) // den, stone)]


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


def binding_limit(path, traffic, launch_ms, valu):
    """What holds the path's kernel back, from the measured figures: HBM traffic
    per launch over the launch time against the VALU issue rate."""
    if path == "plain":
        return "hbm"
    hbm = traffic / (launch_ms / 1e3) / (HBM_PEAK_GBS * 1e9) if traffic else None
    v = valu["frac"] if valu else None
    if hbm is None or v is None:
        return "unmeasured: no counter profile for this kernel in profiles/traffic.json"
    if v >= hbm:
        return (f"valu: the block loop issues VALU at {v:.2f} of the 614 G/s peak, while the measured HBM traffic "
                f"(`traffic`) moves at {hbm:.2f} of 8 TB/s")
    return (f"hbm: the tile loads and stores move the measured traffic (`traffic`) at {hbm:.2f} of 8 TB/s, the "
            f"block loop issues VALU at {v:.2f} of its peak")


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
    ap.add_argument("--duration", type=float, default=20.0, help="serve: arrival window (s)")
    ap.add_argument("--drain", type=float, default=30.0, help="serve: stop this long after the last arrival")
    ap.add_argument("--slo-ms", type=float, default=1000.0, help="serve: latency objective of deadline admission")
    ap.add_argument("--sweep", default="800,1200,1600,2000", help="serve: arrival rates of the sweep")
    ap.add_argument("--sweep-duration", type=float, default=10.0)
    ap.add_argument("--sweep-drain", type=float, default=20.0)
    ap.add_argument("--plain-rate", type=float, default=2000.0)
    ap.add_argument("--plain-duration", type=float, default=5.0)
    ap.add_argument("--plain-drain", type=float, default=5.0)
    ap.add_argument("--verify", type=int, default=0, help="serve: replay every n-th request alone (0: none)")
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--sub", type=int, default=8, help="sub-steps per LDS launch")
    ap.add_argument("--paths", default="plain,blocked,blocked_dense,blocked_tuned")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-looped", action="store_true", help="batch: skip the looped plain-kernel comparison")
    ap.add_argument("--latency-samples", type=int, default=50)
    ap.add_argument("--roofline-launches", type=int, default=10)
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
    for mode in args.paths.split(","):
        w = make_world(S, ctx, rs, args.workload)
        if mode == "blocked_dense":  # the LDS kernel with tile sleeping off: every tile loaded and stored
            w.set_sleep(False)
        sub, tune, lds_kernel = args.sub, None, None
        if mode == "blocked_tuned":  # the variant se_world_autotune picks for this run length (untimed)
            t0 = time.perf_counter()
            vi, sub, trial_ms = w.autotune(args.steps, rounds=2)
            tune = {"variant": vi, "tile": list(w.variant_info(1, vi)[:2]), "halo": sub,
                    "trial_ms_per_step": trial_ms, "seconds": time.perf_counter() - t0}
            log(f"autotune for {args.steps} steps: variant {vi} {w.variant_info(1, vi)}, {sub} sub-steps per launch")
        if mode.startswith("blocked"):
            tw, th, c = w.variant_info(1, w.variant_info(2, sub)[0])
            lds_kernel = f"se_step_lds_{tw}x{th}_h{c // 1000}_t{c % 1000}"
        wall, gpu_ms = time_path(S, w, mode, args.steps, args.warmup, sub)
        launches = args.steps if mode == "plain" else -(-args.steps // sub)
        avg_launch_ms = gpu_ms / launches
        r = dict(gcups=cells * args.steps / wall / 1e9, gcups_gpu=cells * args.steps / (gpu_ms / 1e3) / 1e9,
                 ms_per_step=wall * 1e3 / args.steps, avg_launch_ms=avg_launch_ms, launches=launches,
                 steps_per_launch=1 if mode == "plain" else sub,
                 kernel="se_step_plain_r1_nt1" if mode == "plain" else lds_kernel)
        if tune is not None:
            r["autotune"] = tune
        # roofline timing: launches that each run the full `sub` sub-steps (the timed
        # steps may end on a shorter launch), between HIP events on the world's stream
        if mode.startswith("blocked"):
            w.step_timed(2 * sub, 1, sub)
            full_ms = w.step_timed(sub * args.roofline_launches, 1, sub) / args.roofline_launches
        else:
            full_ms = avg_launch_ms
        r["full_launch_ms"] = full_ms
        r["achieved_gbs"] = 8.0 * cells / (full_ms / 1e3) / 1e9
        r["roofline_frac"] = r["achieved_gbs"] / HBM_PEAK_GBS
        r["frac_of_copy"] = r["achieved_gbs"] / copy_gbs
        p50, p99 = latencies(w, mode, sub, args.latency_samples)
        r["step_latency_ms"] = {"p50": p50, "p99": p99}
        log(f"{mode}: {r['gcups']:.1f} Gcell-upd/s wall, {r['gcups_gpu']:.1f} by events, "
            f"{r['avg_launch_ms']:.3f} ms/launch in the timed steps, {full_ms:.3f} ms per full launch, "
            f"{r['achieved_gbs']:.0f} GB/s algorithmic")
        results[mode] = r
        worlds[mode] = w
    check = None
    if not args.no_check and "plain" in worlds and len(worlds) > 1:
        # every path advanced warmup + steps + latency steps from the same grid;
        # align the frames and compare each blocked path's grid with the plain one's
        a = worlds["plain"]
        top = max(w.frame for w in worlds.values())
        for w in worlds.values():
            if w.frame < top:
                w.step(top - w.frame)
        ga = a.download()
        check = {"frame": top, "identical": True}
        for mode, w in worlds.items():
            if mode != "plain":
                same = bool(np.array_equal(ga, w.download()))
                check[f"{mode}_equals_plain"] = same
                check["identical"] = check["identical"] and same
        del ga
        log(f"blocked paths vs plain after frame {top}: {check}")
    for w in worlds.values():
        w.close()
    # headline: the fastest of the plain kernel, the LDS kernel at the default
    # variant and sub-steps, and the LDS kernel as se_world_autotune picks it for
    # this run length (its trials run before the warmup, outside the timed steps)
    best = max((m for m in results if m in ("plain", "blocked", "blocked_tuned")), key=lambda m: results[m]["gcups"])
    R = results[best]
    kernel = R["kernel"]
    prof = load_profile(kernel, args.workload)
    traffic = prof.get("bytes_per_launch")
    valu = None
    if prof.get("valu_wave_instr_per_launch"):
        rate = prof["valu_wave_instr_per_launch"] / (R["full_launch_ms"] / 1e3)
        valu = {"bound": "valu", "kernel": kernel, "achieved": rate, "peak": VALU_PEAK_WAVE_INSTR_PER_S,
                "unit": "wave64 VALU instructions/s", "frac": rate / VALU_PEAK_WAVE_INSTR_PER_S,
                "instr_per_launch": prof["valu_wave_instr_per_launch"],
                "source": "SQ_INSTS_VALU per launch from profiles/traffic.json over this run's launch time"}
    cpu = None if args.no_cpu_baseline else cpu_baseline(
        single_sample(args.workload), steps=40,
        what=f"2048x2048 world of the same fill and rules as {args.workload} (llvmpipe holds at most 8192^2 RGBA32F)")
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
                     "algorithmic_bytes_per_launch": 8 * cells, "avg_launch_ms": R["full_launch_ms"],
                     "launch_timing": ("the timed steps' launches" if best == "plain" else
                                       f"{args.roofline_launches} launches of {R['steps_per_launch']} sub-steps "
                                       f"after the timed steps"),
                     "binding_limit": binding_limit(best, traffic, R["full_launch_ms"], valu)},
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
