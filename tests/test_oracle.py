"""Pins the oracle (oracle/) before anything is checked against it.

* against the llvmpipe golden fixtures (the reference's own GLSL program run
  under Mesa llvmpipe, tests/golden/),
* against the reference's checked-in generated GLSL (SHA-256 fixtures),
* against closed forms: the lowbias32 hash, the u32->f32 rand thresholds, the
  integer circle test, the 4-phase Margolus offset,
* against invariants: per-material counts under swap-only rules, WALL/NULL
  immobility, one-block-per-cell == one-block-per-block evaluation,
* and, where oracle/_ref is built (this container), live against llvmpipe on
  fresh random cases.
"""
import hashlib
import os

import numpy as np
import pytest

from oracle import lang_ref as L
from oracle import oracle as O
from tests.golden import cases as C
from tests.golden import load

ALL_CASES = load.cases()


def _blob(ruleset):
    txt = load.ruleset_text(ruleset)
    if txt is None:
        pytest.skip("reference tree absent")
    return O.compile_ruleset(L.parse_string(txt))


@pytest.mark.parametrize("case", ALL_CASES, ids=[c["name"] for c in ALL_CASES])
def test_oracle_matches_llvmpipe_golden(case):
    blob = _blob(case["ruleset"])
    g = C.make_input(case)
    assert hashlib.sha256(g.tobytes()).hexdigest() == case["input_sha256"]
    ma = load.mods_at(case)
    f = case.get("first_frame", 2)
    cps = set(case.get("checkpoints", []))
    for s in range(case["steps"]):
        g = O.step(blob, g, f + s, ma.get(s))
        if s + 1 in cps:
            np.testing.assert_array_equal(g, load.grid(f"{case['name']}@{s + 1}"))
    np.testing.assert_array_equal(g, load.grid(case["name"]))


def test_run_helper_equals_stepping():
    case = [c for c in ALL_CASES if c["name"] == "config1_like_128_four_seed1"][0]
    blob = _blob("four")
    np.testing.assert_array_equal(O.run(blob, C.make_input(case), 2, case["steps"]),
                                  load.grid(case["name"]))


def test_glsl_emitter_matches_reference_gen():
    txt = load.ruleset_text("reference_default")
    if txt is None:
        pytest.skip("reference tree absent")
    m, r = L.create_glsl(L.parse_string(txt))
    want = load.glsl_sha256()
    assert hashlib.sha256(m.encode()).hexdigest() == want["materials.glsl"]
    assert hashlib.sha256(r.encode()).hexdigest() == want["rules.glsl"]


def test_include_resolver_matches_reference_program():
    from oracle import llvmpipe as LP
    if not LP.available():
        pytest.skip("oracle/_ref not built")
    asm = LP.assemble(None, None)
    assert hashlib.sha256(asm.encode()).hexdigest() == load.glsl_sha256()["falling_sand.glsl"]


def _lowbias(h):
    h = np.uint64(h)
    m = np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(16); h = (h * np.uint64(0x7FEB352D)) & m
    h ^= h >> np.uint64(15); h = (h * np.uint64(0x846CA68B)) & m
    h ^= h >> np.uint64(16)
    return int(h)


def test_hash43_closed_form():
    rng = np.random.RandomState(0)
    pts = [(0, 0, 0), (0xFFFFFFFF, 5, 2), (7, 0xFFFFFFFF, 1000003), (1, 1, 2147483647)]
    pts += [tuple(int(v) for v in rng.randint(0, 2 ** 32, size=3, dtype=np.uint64))
            for _ in range(200)]
    for x, y, f in pts:
        u, r = O.hash43(x, y, f)
        base = (x * 461 + y * 2131 + f * 2131 * 2131) & 0xFFFFFFFF
        for k, c in enumerate([213, 2131, 21313, 213132]):
            h = _lowbias((base * c) & 0xFFFFFFFF)
            assert int(u[k]) == h
            assert r[k] == np.float32(h) * np.float32(2.0 ** -32)


def _threshold(p, le):
    """max u with RNE_f32(u)*2^-32 (<= or <) p, by bisection over u32."""
    pf = np.float32(p)
    lo, hi = 0, 2 ** 32 - 1
    ok = (lambda u: np.float32(u) * np.float32(2.0 ** -32) <= pf) if le else \
        (lambda u: np.float32(u) * np.float32(2.0 ** -32) < pf)
    while lo < hi:
        mid = (lo + hi + 1) // 2
        if ok(mid):
            lo = mid
        else:
            hi = mid - 1
    return lo


def test_rand_thresholds_closed_form():
    assert _threshold(0.5, le=False) == 2 ** 31 - 65
    assert _threshold(0.5, le=True) == 2147483776
    assert _threshold(0.1, le=True) == 429496751
    assert _threshold(0.3, le=True) == 1288490304
    assert _threshold(0.004, le=True) == 17179870
    assert _threshold(0.001, le=True) == 4294967
    assert _threshold(0.02, le=True) == 85899348


def test_margolus_offset_cycle():
    # a 2x2 world with one sand grain at (0,0).  Offsets for frame % 4 = 0,1,2,3
    # are (0,0),(1,1),(0,1),(1,0).  Only offset (0,0) makes the grain the SELF cell
    # of a block holding the cell below it; then it falls unless the block is
    # mirrored (rand.x < 0.5 puts it in RIGHT, where the SELF-keyed rule is off).
    # With x offset 1 it is RIGHT of an out-of-bounds WALL, and the guarded
    # mirror swap refuses to move it.
    blob = _blob("four")
    seen = set()
    for f in range(2, 60):
        g = np.array([[3, 0], [0, 0]], np.uint32)
        out = O.step(blob, g, f)
        falls = out[1, 0] == 3 or out[1, 1] == 3
        _, r = O.hash43(0, 0, f)
        assert falls == (f % 4 == 0 and not r[0] < 0.5), f
        seen.add(bool(falls))
    assert seen == {True, False}


def test_circle_integer_closed_form():
    blob = _blob("four")
    W = H = 140
    g = np.zeros((H, W), np.uint32)
    for r in [1, 2, 3, 5, 13, 25, 65]:
        out = O.step(blob, g, 2, [[70, 70, 0, r, 5]])
        yy, xx = np.mgrid[0:H, 0:W]
        want = ((xx - 70) ** 2 + (yy - 70) ** 2 <= r * r)
        np.testing.assert_array_equal(out == 5, want)
        sq = O.step(blob, g, 2, [[70, 70, 1, r, 4]])
        np.testing.assert_array_equal(sq == 4, (abs(xx - 70) <= r) & (abs(yy - 70) <= r))


def test_last_match_wins_and_null_hides():
    blob = _blob("four")
    g = np.zeros((16, 16), np.uint32)
    out = O.step(blob, g, 2, [[8, 8, 1, 3, 3], [8, 8, 1, 1, 4]])
    assert out[8, 8] == 4 and out[6, 6] == 3
    out = O.step(blob, g, 2, [[8, 8, 1, 3, 3], [8, 8, 1, 1, 77]])  # unknown -> NULL
    assert out[8, 8] == 0 and out[6, 6] == 3
    out = O.step(blob, g, 2, [[8, 8, 1, 3, 3], [0, 0, 1, 0, 4], [8, 8, 1, 1, 4]])
    assert out[8, 8] == 3  # size 0 ends the list


def test_percell_order_equals_block_order():
    rng = np.random.RandomState(1)
    for rs in ["four", "twelve", "features"]:
        blob = _blob(rs)
        for W, H in [(1, 1), (3, 5), (16, 9), (33, 20)]:
            g = rng.randint(0, 20, size=(H, W)).astype(np.uint32)
            for f in range(2, 8):
                mods = [[int(rng.randint(0, W)), int(rng.randint(0, H)), int(rng.randint(0, 2)),
                         int(rng.randint(1, 4)), int(rng.randint(0, 8))]] if f % 2 else None
                np.testing.assert_array_equal(O.step(blob, g, f, mods),
                                              O.step(blob, g, f, mods, percell=True))


def test_swap_only_rules_conserve_counts_and_walls_stay():
    blob = _blob("four")
    g = O.fill(5, 96, 80, [0, 3, 4, 5, 2, 1], [40, 20, 20, 10, 5, 5])
    counts = np.bincount(g.ravel(), minlength=6)
    fixed = (g == 1) | (g == 2)
    for f in range(2, 200):
        g = O.step(blob, g, f)
    np.testing.assert_array_equal(np.bincount(g.ravel(), minlength=6), counts)
    np.testing.assert_array_equal((g == 1) | (g == 2), fixed)


def test_unknown_ids_become_null_after_one_step():
    blob = _blob("four")
    g = np.full((6, 6), 999, np.uint32)
    g[0, 0] = 0
    out = O.step(blob, g, 2)
    assert set(np.unique(out)) <= {0, 1}
    assert (out == 1).sum() == 35


def test_clearing_frame_ignores_mods():
    blob = _blob("four")
    g = O.fill(3, 10, 10, [0, 3, 4, 5], [25, 25, 25, 25])
    out = O.step(blob, g, 1, [[5, 5, 0, 4, 5]])
    assert not out.any()


@pytest.mark.parametrize("rs", ["four", "twelve", "features", "reference_default"])
def test_oracle_live_against_llvmpipe(rs):
    from oracle import llvmpipe as LP
    if not LP.available():
        pytest.skip("oracle/_ref not built")
    txt = load.ruleset_text(rs)
    if txt is None:
        pytest.skip("reference tree absent")
    res = L.parse_string(txt)
    blob = O.compile_ruleset(res)
    prog = LP.program_for(None if rs == "reference_default" else res)
    rng = np.random.RandomState(hash(rs) % 2 ** 31)
    for trial in range(3):
        W, H = int(rng.randint(1, 50)), int(rng.randint(1, 50))
        case = dict(ruleset=rs, W=W, H=H, seed=int(rng.randint(0, 10 ** 6)), specials=0.02)
        g = C.make_input(case)
        mods = {int(rng.randint(0, 30)): C._rand_mods(rng, W, H, C.NMAT[rs], 3)}
        want = LP.seeded_run(prog, g, 30, mods)
        got = g
        for s in range(30):
            got = O.step(blob, got, 2 + s, mods.get(s))
        np.testing.assert_array_equal(got, want)


def test_lang_error_categories():
    ok_types = "types:\n  t:\n"
    ok_mats = "materials:\n  m:\n    type: t\n    color: [1, 2, 3]\n    density: 1.0\n"
    ok_rules = "rules:\n  r:\n    if: isType_t(DOWN)\n    do: SWAP SELF DOWN\n"
    cases = [
        (ok_types + ok_mats, "MissingField"),
        (ok_rules + ok_mats, "MissingField"),
        (ok_rules + ok_types, "MissingField"),
        ("rules: 3\n" + ok_types + ok_mats, "InvalidType"),
        ("rules:\n  r:\n    do: SWAP SELF DOWN\n" + ok_types + ok_mats, "MissingField"),
        ("rules:\n  r:\n    if: isType_t(DOWN)\n" + ok_types + ok_mats, "MissingField"),
        ("rules:\n  r:\n    if: isType_zz(DOWN)\n    do: SWAP SELF DOWN\n" + ok_types + ok_mats,
         "NotFound"),
        ("rules:\n  r:\n    if: DOWN.mat == qq\n    do: SWAP SELF DOWN\n" + ok_types + ok_mats,
         "NotFound"),
        ("rules:\n  r:\n    if: isType_t(DOWN)\n    do: swap(SELF, DOWN)\n" + ok_types + ok_mats,
         "NotRecognized"),
        ("rules:\n  r:\n    if: isType_t(DOWN)\n    do: SWAP SELF UP\n" + ok_types + ok_mats,
         "NotFound"),
        (ok_rules + ok_types + "materials:\n  m:\n    type: nope\n    color: [1, 2, 3]\n"
                              "    density: 1.0\n", "NotFound"),
        (ok_rules + ok_types + "materials:\n  m:\n    type: t\n    density: 1.0\n", "MissingField"),
        (ok_rules + ok_types + "materials:\n  m:\n    type: t\n    color: [1, 2]\n"
                              "    density: 1.0\n", "InvalidType"),
        (ok_rules + ok_types + "materials:\n  m:\n    type: t\n    color: [1, 2, 3]\n"
                              "    density: heavy\n", "InvalidType"),
        (ok_rules + "types:\n  t:\n    base_rules: [nope]\n" + ok_mats, "NotFound"),
        (ok_rules + "types:\n  t:\n    base_rules: r\n" + ok_mats, "InvalidType"),
        ("rules:\n  r:\n    if: isType_t(DOWN)\n    do: SWAP SELF DOWN\n    mirrored: 3\n"
         + ok_types + ok_mats, "InvalidType"),
        ("rules:\n  r:\n    if: isType_t(DOWN)\n    do: SWAP SELF DOWN\n    probability: x\n"
         + ok_types + ok_mats, "InvalidType"),
    ]
    for text, kind in cases:
        with pytest.raises(L.ParsingErr) as e:
            L.parse_string(text)
        assert e.value.kind == kind, (text, e.value.message)
    L.parse_string(ok_rules + ok_types + ok_mats)
