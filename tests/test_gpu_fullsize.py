"""GPU path against the llvmpipe oracle at the benchmark workloads' real shapes.

tests/golden/fullsize.json holds SHA-256 digests of grids the reference's own
GLSL program produced under Mesa llvmpipe (tests/golden/make_fullsize.py): the
first steps of the 16384^2 four-material world and of the 8192^2 twelve-material
world, and whole runs of small worlds of the 4096-world batch.  The device fill
must reproduce the recorded input, and the plain kernel, the LDS kernel and the
batched kernel must all reproduce the recorded output, bit for bit.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import sandengine_amd as S

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fullsize.json")
CASES = json.load(open(FIXTURE))["cases"] if os.path.exists(FIXTURE) else []
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
    return hashlib.sha256(np.ascontiguousarray(g, dtype=np.uint32).tobytes()).hexdigest()


def world(c):
    r = rs(c["ruleset"])  # creates the context on first use
    w = S.World(_ctx, r, c["W"], c["H"])
    w.frame = c["first_frame"] - 1  # the clearing frame has run
    w.fill_seeded(c["seed"], c["ids"], c["percents"], stone_rows=c["stone_rows"], stone_id=c["stone_id"])
    return w


SINGLE = [c for c in CASES if not c["name"].startswith("batch")]
BATCHED = [c for c in CASES if c["name"].startswith("batch")]


@pytest.mark.parametrize("case", SINGLE, ids=[c["name"] for c in SINGLE])
@pytest.mark.parametrize("mode", ["plain", "blocked"])
def test_single_world_matches_llvmpipe(case, mode):
    print(f"\n[gpu] fullsize {case['name']} {mode}", flush=True)
    w = world(case)
    assert sha(w.download()) == case["input_sha256"]
    if mode == "plain":
        w.step(case["steps"])
    else:
        w.step_blocked(case["steps"], 0)
    assert w.frame == case["first_frame"] - 1 + case["steps"]
    assert sha(w.download()) == case["output_sha256"]
    w.close()


@pytest.mark.skipif(not BATCHED, reason="no batched fixtures")
def test_batched_worlds_match_llvmpipe():
    print("\n[gpu] fullsize batch", flush=True)
    rs(BATCHED[0]["ruleset"])  # creates the context
    b = S.Batch(_ctx)
    worlds = []
    for c in BATCHED:
        w = world(c)
        assert sha(w.download()) == c["input_sha256"], c["name"]
        b.add(w, c["steps"])
        worlds.append(w)
    b.run(mode=1, graph=True)
    for c, w in zip(BATCHED, worlds):
        assert sha(w.download()) == c["output_sha256"], c["name"]
        w.close()
    b.close()
