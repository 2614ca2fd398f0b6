"""Serving loop: continuous batching must not change any result.

Each request's final grid (FNV-1a 64 recorded by the loop) must equal the same
request stepped alone through the plain single-world path -- same fill, same
modifications at the same steps -- whatever else shared its rounds.  Covered
for the batched LDS kernel and the plain kernel, under a roomy budget and a
tight one that forces FIFO queueing, with a request larger than the budget
rejected.
"""
import numpy as np
import pytest

import sandengine_amd as S

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
