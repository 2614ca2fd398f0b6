"""Serving loop: continuous batching must not change any result.

Each request's final grid (FNV-1a 64 recorded by the loop) must equal the same
request stepped alone through the plain single-world path -- same fill, same
modifications at the same steps -- whatever else shared its rounds.  Covered
for the batched LDS kernel and the plain kernel, under a roomy budget and a
tight one that forces FIFO queueing, with a request larger than the budget
rejected; at the served workload's own request mix (sides 128-4096, steps
10-5000, 0-20 brushes of radius 4-64 and explosions of radius 8-128, arrivals
at 2000 per second) with every completed request replayed alone and the
smallest ones also against the oracle's C restatement; and under deadline
admission, where completed requests stay within the latency objective and the
others are shed.
"""
import numpy as np
import pytest

import sandengine_amd as S
from oracle import lang_ref as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FILLS = [([0, 3, 4, 5], [50, 25, 15, 10]), ([0] + list(range(3, 15)), [40] + [5] * 12)]
_ctx = None
_rs = None


def setup():
    global _ctx, _rs
    if _ctx is None:
        _ctx = S.Context(0)
        _rs = [S.RuleSet(_ctx, S.bundled_ruleset("four_materials")),
               S.RuleSet(_ctx, S.bundled_ruleset("twelve_materials"))]
    return _ctx, _rs


def replay_alone(ctx, rs, r):
    ids, pct = FILLS[r["ruleset"]]
    w = S.World(ctx, rs[r["ruleset"]], r["side"], r["side"])
    w.frame = 1
    w.fill_seeded(r["fill_seed"], ids, pct)
    by_step = {}
    for step, m in r["mods"]:
        by_step.setdefault(step, []).append(m)
    done = 0
    for s in sorted(by_step) + [r["steps"]]:
        if s > done:
            w.step(s - done)
            done = s
        if s < r["steps"]:
            for m in by_step[s]:
                w.push_modification(*m)
    h = S.fnv1a64(w.download())
    w.close()
    return h


@pytest.mark.parametrize("kernel,budget_mb", [(1, 256), (0, 256), (1, 6)])
def test_serving_results_equal_alone(kernel, budget_mb):
    print(f"\n[gpu] serve kernel={kernel} budget={budget_mb}MB", flush=True)
    ctx, rs = setup()
    srv = S.Serve(ctx, rs, FILLS, rate=300.0, duration_s=0.6, seed=9 + kernel + budget_mb, side_min=8,
                  side_max=420, steps_min=2, steps_max=150, mods_max=6, kernel=kernel, max_inflight=2,
                  hbm_budget=budget_mb << 20, verify_every=1)
    st = srv.run()
    n = len(srv)
    assert st["requests"] == n and n > 50
    assert st["completed"] + st["rejected"] == n and st["unfinished"] == 0
    assert st["p50_ms"] > 0 and st["p99_ms"] >= st["p50_ms"]
    checked = 0
    for i in range(n):
        r = srv.request(i)
        if r["state"] == 3:
            continue
        assert r["state"] == 2 and r["verified"] == 1
        assert replay_alone(ctx, rs, r) == r["hash"], r
        checked += 1
    assert checked == st["completed"]
    srv.close()


def test_oversized_request_is_rejected():
    ctx, rs = setup()
    srv = S.Serve(ctx, rs, FILLS, rate=50.0, duration_s=0.2, seed=3, side_min=600, side_max=700, steps_min=1,
                  steps_max=3, mods_max=0, kernel=1, hbm_budget=2 << 20, verify_every=0)
    st = srv.run()
    assert st["rejected"] == st["requests"] > 0 and st["completed"] == 0
    srv.close()


# the served workload's request mix
CONFIG4 = dict(side_min=128, side_max=4096, steps_min=10, steps_max=5000, mods_max=20)
RULE_NAMES = ["four_materials", "twelve_materials"]


def oracle_replay(r, blobs):
    """The request's final grid from the oracle: seeded fill, then its steps with
    each step's modifications applied to that step (frames 2, 3, ...)."""
    ids, pct = FILLS[r["ruleset"]]
    g = O.fill(r["fill_seed"], r["side"], r["side"], ids, pct)
    by_step = {}
    for step, m in r["mods"]:
        by_step.setdefault(step, []).append(m)
    done, frame, pending = 0, 2, None
    for s in sorted(by_step) + [r["steps"]]:
        if s > done:
            g = O.run(blobs[r["ruleset"]], g, frame, s - done, pending)
            frame += s - done
            done, pending = s, None
        if s < r["steps"]:
            pending = by_step[s]
    return S.fnv1a64(g)


def budget(ctx, want=200 << 30):
    free = ctx.mem_info()[1]
    return int(min(want, free - (16 << 30)))


def test_served_mix_results_equal_alone_and_oracle():
    print("\n[gpu] serve: served mix, 2000 req/s for 0.25 s, verify every request", flush=True)
    ctx, rs = setup()
    srv = S.Serve(ctx, rs, FILLS, rate=2000.0, duration_s=0.25, seed=4, kernel=1, max_inflight=2,
                  hbm_budget=budget(ctx), verify_every=1, **CONFIG4)
    st = srv.run()
    n = len(srv)
    print(f"[gpu] serve: {n} requests, {st['completed']} done, p50 {st['p50_ms']:.0f} ms, "
          f"p99 {st['p99_ms']:.0f} ms, {st['processed_gcups']:.0f} Gcell/s", flush=True)
    assert n > 300 and st["completed"] == n and st["verified"] == n
    reqs = [srv.request(i) for i in range(n)]
    assert max(r["side"] for r in reqs) > 2048 and max(r["steps"] for r in reqs) > 2500
    assert sum(len(r["mods"]) for r in reqs) > 10 * n - 200
    assert max(m[3] for r in reqs for _, m in r["mods"]) > 100  # explosions of radius > 100
    for r in reqs:
        assert r["state"] == 2 and r["verified"] == 1
        assert replay_alone(ctx, rs, r) == r["hash"], {k: r[k] for k in ("side", "steps", "ruleset", "fill_seed")}
    # the smallest requests with modifications, against the oracle's C restatement
    blobs = [O.compile_ruleset(L.parse_string(S.bundled_ruleset(k))) for k in RULE_NAMES]
    small = sorted((r for r in reqs if r["mods"]), key=lambda r: r["side"] ** 2 * r["steps"])
    picked = [r for r in small if r["ruleset"] == 0][:3] + [r for r in small if r["ruleset"] == 1][:3]
    assert len(picked) == 6
    for r in picked:
        assert oracle_replay(r, blobs) == r["hash"], {k: r[k] for k in ("side", "steps", "ruleset", "fill_seed")}
    srv.close()


def test_deadline_admission_meets_objective_or_sheds():
    print("\n[gpu] serve: deadline admission at 4000 req/s, objective 300 ms", flush=True)
    ctx, rs = setup()
    srv = S.Serve(ctx, rs, FILLS, rate=4000.0, duration_s=1.0, seed=5, kernel=1, max_inflight=2,
                  hbm_budget=budget(ctx, 64 << 30), verify_every=7, slo_ms=300.0, **CONFIG4)
    st = srv.run()
    n = len(srv)
    print(f"[gpu] serve: {n} requests, {st['completed']} done, {st['shed']} shed "
          f"({st['shed_cell_frac']:.2f} of the work), p50 {st['p50_ms']:.0f} ms, p99 {st['p99_ms']:.0f} ms, "
          f"within objective {st['slo_met_frac']:.3f}, capacity {st['capacity_gcups_est']:.0f} Gcell/s, "
          f"round fixed {st['round_fixed_ms']:.3f} ms, actual/predicted p50 {st['pred_ratio_p50']:.2f} "
          f"p99 {st['pred_ratio_p99']:.2f}", flush=True)
    assert st["completed"] + st["shed"] + st["rejected"] == n and st["unfinished"] == 0
    assert st["shed"] > 0 and st["completed"] > n // 4  # overload: some shed, most served
    assert st["slo_met_frac"] >= 0.95
    checked = 0
    for i in range(0, n, 7):
        r = srv.request(i)
        assert r["state"] in (2, 4)
        if r["state"] == 2:
            assert r["verified"] == 1 and replay_alone(ctx, rs, r) == r["hash"]
            checked += 1
    assert checked > 0
    srv.close()
