"""Case definitions shared by make_golden.py (which runs them under llvmpipe) and
the tests (which regenerate each input from these parameters and compare).

A case: rule set, size, seeded categorical fill (oracle.fill), optional special
ids sprinkled in (NULL, WALL and unknown ids that must decode to NULL), the
first frame, the step count and modifications {step: [[x, y, shape, size, matID], ...]}.
"""
from __future__ import annotations

import numpy as np

from oracle import oracle as O

RULESETS = {
    "four": "sandengine_amd/rulesets/four_materials.yaml",
    "twelve": "sandengine_amd/rulesets/twelve_materials.yaml",
    "features": "sandengine_amd/rulesets/features.yaml",
    # the reference's own data/materials.yaml, run with its checked-in program
    "reference_default": None,
}

FILLS = {
    "four": ([0, 3, 4, 5], [50, 25, 15, 10]),
    "twelve": ([0] + list(range(3, 15)), [40] + [5] * 12),
    "features": ([0, 3, 4, 5, 6, 7], [40, 15, 10, 15, 5, 15]),
    "reference_default": ([0, 3, 4, 5, 6, 7, 8, 9, 10], [40, 10, 10, 10, 5, 10, 5, 5, 5]),
}
NMAT = {"four": 6, "twelve": 15, "features": 8, "reference_default": 11}
SPECIAL_IDS = [1, 2, 200, 65535, 16777215]


def make_input(case) -> np.ndarray:
    ids, pct = FILLS[case["ruleset"]]
    if "fill" in case:
        ids, pct = case["fill"]
    W, H = case["W"], case["H"]
    g = O.fill(case["seed"], W, H, ids, pct, case.get("stone_rows", 0), case.get("stone_id", 0))
    sp = case.get("specials", 0)
    if sp:
        r = O.cell_rand(case["seed"] + 1000, np.arange(W * H, dtype=np.uint64)).reshape(H, W)
        pick = r < np.uint32(int(sp * 2 ** 32))
        which = (r % np.uint32(len(SPECIAL_IDS))).astype(np.int64)
        g[pick] = np.asarray(SPECIAL_IDS, np.uint32)[which[pick]]
    return g


def _rand_mods(rng, W, H, nmat, n, allow_weird=True):
    mods = []
    for _ in range(n):
        x = int(rng.randint(-4, W + 4))
        y = int(rng.randint(-4, H + 4))
        shape = int(rng.randint(0, 2)) if not allow_weird else int(rng.choice([0, 0, 1, 1, 2]))
        size = int(rng.randint(1, max(2, min(W, H) // 2 + 2)))
        mat = int(rng.randint(-1, nmat + 2)) if allow_weird else int(rng.randint(0, nmat))
        mods.append([x, y, shape, size, mat])
    return mods


def _build():
    cases = []
    cases.append(dict(name="config0_256_four_seed0", ruleset="four", W=256, H=256, seed=0,
                      steps=1000, checkpoints=[1, 10, 100, 500]))
    rng = np.random.RandomState(12345)
    sizes = [(1, 1), (1, 6), (6, 1), (2, 2), (3, 3), (5, 7), (8, 8), (17, 13), (32, 32), (64, 48),
             (97, 61)]
    for rs in ["four", "twelve", "features", "reference_default"]:
        for i, (W, H) in enumerate(sizes):
            steps = int(rng.randint(20, 120))
            c = dict(name=f"small_{rs}_{W}x{H}", ruleset=rs, W=W, H=H, seed=100 + 17 * i,
                     steps=steps, specials=0.03 if i % 2 else 0.0)
            if i % 3 != 1:
                mods_at = {}
                for k in sorted(set(int(v) for v in rng.randint(0, steps, size=3))):
                    mods_at[str(k)] = _rand_mods(rng, W, H, NMAT[rs], int(rng.randint(1, 6)))
                c["mods_at"] = mods_at
            cases.append(c)
    # circle / square boundaries: Pythagorean radii, overlapping mods, last-wins,
    # NULL material hiding an earlier match, size 0 ending the list
    circ = [[100, 100, 0, 5, 5], [30, 30, 0, 13, 4], [60, 140, 0, 25, 3], [120, 120, 0, 65, 5],
            [170, 30, 0, 1, 3], [180, 60, 0, 2, 4], [150, 180, 0, 3, 5], [20, 180, 1, 6, 3],
            [60, 150, 1, 4, 4], [125, 120, 0, 10, 1], [110, 110, 0, 8, 200], [10, 10, 1, 3, 2],
            [-3, 100, 0, 7, 5], [199, 199, 0, 9, 4]]
    cases.append(dict(name="mods_circle_boundaries", ruleset="four", W=200, H=200, seed=7,
                      steps=3, mods_at={"0": circ}))
    cases.append(dict(name="mods_size0_terminates", ruleset="four", W=64, H=64, seed=8, steps=2,
                      mods_at={"0": [[10, 10, 0, 5, 5], [40, 40, 1, 0, 4], [50, 50, 0, 9, 3]]}))
    cases.append(dict(name="mods_at_clearing_frame", ruleset="four", W=32, H=32, seed=9, steps=4,
                      first_frame=1, mods_at={"0": [[16, 16, 0, 8, 5]], "2": [[5, 5, 1, 3, 3]]}))
    cases.append(dict(name="large_frame_four", ruleset="four", W=64, H=64, seed=10, steps=50,
                      first_frame=1000003))
    cases.append(dict(name="large_frame_twelve", ruleset="twelve", W=64, H=40, seed=11, steps=50,
                      first_frame=2147483000))
    cases.append(dict(name="config2_like_256_twelve_seed2", ruleset="twelve", W=256, H=256, seed=2,
                      steps=300, checkpoints=[50, 150]))
    cases.append(dict(name="config1_like_128_four_seed1", ruleset="four", W=128, H=128, seed=1,
                      steps=300, fill=([0, 3, 4], [60, 25, 15]), stone_rows=38, stone_id=5))
    cases.append(dict(name="features_128x96", ruleset="features", W=128, H=96, seed=12,
                      steps=400, specials=0.01,
                      mods_at={"5": [[64, 48, 0, 20, 7], [10, 90, 1, 5, 2]],
                               "200": [[30, 30, 0, 12, 0], [100, 10, 1, 8, 1]]}))
    cases.append(dict(name="reference_default_128x128", ruleset="reference_default", W=128,
                      H=128, seed=13, steps=400, specials=0.01,
                      mods_at={"3": [[64, 64, 0, 16, 7], [20, 100, 1, 6, 9]]}))
    return cases


CASES = _build()
