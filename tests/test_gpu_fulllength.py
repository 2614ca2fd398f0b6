"""GPU paths against the oracle over the benchmark workloads' full lengths.

tests/golden/fulllength.json holds SHA-256 digests written by the oracle's C
restatement of the step (tests/golden/make_fulllength.py), which is pinned to
the reference's GLSL program under Mesa llvmpipe (tests/golden/cases.json) and
reproduces every llvmpipe digest of tests/golden/fullsize.json at these shapes:
  single16k  16384^2, 4 materials, 2000 steps: after 4 steps and every 250;
  rich8k     8192^2, 12 materials, 5000 steps: after 16 steps and every 500;
  batch4096  all 4096 worlds of the batch draw, each over all its steps.
The plain kernel and the LDS kernel must reach every checkpoint bit for bit,
and the batched kernel (one hipGraph) every world's final grid.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import sandengine_amd as S

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fulllength.json")
CASES = json.load(open(FIXTURE))["cases"] if os.path.exists(FIXTURE) else {}
_ctx = None
_rs = {}


def rs(name):
    global _ctx
    if _ctx is None:
        _ctx = S.Context(0)
    if name not in _rs:
        _rs[name] = S.RuleSet(_ctx, S.bundled_ruleset(name))
    return _rs[name]


def sha(g):
    return hashlib.sha256(memoryview(np.ascontiguousarray(g, dtype=np.uint32)).cast("B")).hexdigest()


SINGLE = [k for k in ("single16k", "rich8k") if k in CASES]


@pytest.mark.parametrize("mode", ["plain", "blocked"])
@pytest.mark.parametrize("name", SINGLE)
def test_single_world_full_length(name, mode):
    c = CASES[name]
    print(f"\n[gpu] full length {name} {mode}: {c['W']}x{c['H']}, {c['checkpoints'][-1]['steps']} steps", flush=True)
    r = rs(c["ruleset"])  # creates the context on first use
    w = S.World(_ctx, r, c["W"], c["H"])
    w.frame = c["first_frame"] - 1  # the clearing frame has run
    w.fill_seeded(c["seed"], c["ids"], c["percents"], stone_rows=c["stone_rows"], stone_id=c["stone_id"])
    assert sha(w.download()) == c["input_sha256"]
    done = 0
    for cp in c["checkpoints"]:
        n = cp["steps"] - done
        w.step(n) if mode == "plain" else w.step_blocked(n, 0)
        done = cp["steps"]
        assert w.frame == c["first_frame"] - 1 + done
        got = sha(w.download())
        print(f"  step {done}: {got[:16]} {'ok' if got == cp['sha256'] else 'DIFFERS'}", flush=True)
        assert got == cp["sha256"], f"{name} {mode}: grid differs from the oracle after {done} steps"
    w.close()


@pytest.mark.skipif("batch4096" not in CASES, reason="no batch fixture")
def test_batch4096_full_length():
    c = CASES["batch4096"]
    print(f"\n[gpu] full length batch4096: {len(c['worlds'])} worlds, batched kernel in one hipGraph", flush=True)
    b = S.Batch(rs("four_materials").ctx)  # creates the context on first use
    worlds = []
    for side, rsname, steps, fseed, _ in c["worlds"]:
        ids, pct = c["fills"][rsname]
        w = S.World(_ctx, rs(rsname), side, side)
        w.frame = c["first_frame"] - 1
        w.fill_seeded(fseed, ids, pct)
        b.add(w, steps)
        worlds.append(w)
    ms = b.run(mode=1, graph=True)
    print(f"  batched run {ms:.1f} ms", flush=True)
    bad = []
    for i, ((side, rsname, steps, fseed, want), w) in enumerate(zip(c["worlds"], worlds)):
        assert w.frame == c["first_frame"] - 1 + steps
        if sha(w.download())[:32] != want:
            bad.append(i)
        w.close()
    b.close()
    assert not bad, f"{len(bad)} worlds differ from the oracle, first {bad[:8]}"
