"""GPU parity of the batched multi-world path (se_batch_*).

Every world of a batch must end bit-identical to a twin world advanced alone by
the plain kernel (itself pinned to the oracle in test_gpu_parity.py), whatever
the mix: sizes that fit one LDS tile and sizes that need rounds, two rule sets,
uneven step counts, a world at its clearing frame, a world with queued
modifications, graph and stream submission, tuned and default variants.
"""
import numpy as np
import pytest

import sandengine_amd as S
from oracle import lang_ref as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FILLS = {"four_materials": ([0, 3, 4, 5], [50, 25, 15, 10]),
         "twelve_materials": ([0] + list(range(3, 15)), [40] + [5] * 12)}
SHAPES = [(2, 2), (1, 7), (64, 64), (100, 36), (192, 192), (194, 190), (200, 200), (256, 256), (300, 140),
          (513, 257), (1024, 64), (70, 700), (386, 386), (130, 1030), (112, 112), (114, 110), (224, 113)]
_ctx = None
_rs = {}


def ctx():
    global _ctx
    if _ctx is None:
        _ctx = S.Context(0)
    return _ctx


def rs(name):
    if name not in _rs:
        _rs[name] = S.RuleSet(ctx(), S.bundled_ruleset(name))
    return _rs[name]


def make_worlds(seed, n_extra=10):
    rng = np.random.RandomState(seed)
    specs = []
    for i, (W, H) in enumerate(SHAPES + [(int(rng.randint(2, 600)), int(rng.randint(2, 600)))
                                         for _ in range(n_extra)]):
        name = "four_materials" if rng.rand() < 0.5 else "twelve_materials"
        steps = int(rng.randint(0, 90))
        specs.append((name, W, H, steps, int(rng.randint(0, 10 ** 6))))
    worlds, twins = [], []
    for i, (name, W, H, steps, fs) in enumerate(specs):
        ids, pct = FILLS[name]
        g0 = O.fill(fs, W, H, ids, pct)
        for lst in (worlds, twins):
            w = S.World(ctx(), rs(name), W, H)
            if i == 3:
                w.frame = 0  # next step is the clearing frame
            else:
                w.frame = 1 + (fs % 50)
            w.upload(g0)
            if i == 5:
                w.push_modification(W // 2, H // 2, 0, 9, 4)
            lst.append(w)
    return specs, worlds, twins


@pytest.mark.parametrize("graph", [False, True])
def test_batch_equals_per_world_plain(graph):
    print(f"\n[gpu] batch graph={graph}", flush=True)
    specs, worlds, twins = make_worlds(11 if graph else 12)
    b = S.Batch(ctx())
    for w, sp in zip(worlds, specs):
        b.add(w, sp[3])
    for rep in range(2):
        b.run(mode=1, graph=graph)
        for w, t, sp in zip(worlds, twins, specs):
            t.step(sp[3])
            assert w.frame == t.frame
            np.testing.assert_array_equal(w.download(), t.download(), err_msg=str(sp))
    b.close()


def test_batch_modes_and_autotune_agree():
    specs, worlds, twins = make_worlds(13, n_extra=4)
    b = S.Batch(ctx())
    for w, sp in zip(worlds, specs):
        b.add(w, sp[3])
    v, k, persistent, score = b.autotune(rounds=2)
    assert k >= 2 and score > 0
    for name in sorted({sp[0] for sp in specs}):  # both rule sets have worlds that take part
        tv, tk, tp = b.tuned(rs(name))
        assert tv >= 0 and tk >= 2
    b.run(mode=1)
    b.run(mode=0)
    b.run(mode=2, graph=True)
    b.run(mode=3)
    for w, t, sp in zip(worlds, twins, specs):
        t.step(4 * sp[3])
        np.testing.assert_array_equal(w.download(), t.download(), err_msg=str(sp))
    b.close()


def test_batch_small_worlds_against_oracle():
    """Single-tile worlds advance hundreds of steps inside one work item."""
    text = S.bundled_ruleset("twelve_materials")
    blob = O.compile_ruleset(L.parse_string(text))
    b = S.Batch(ctx())
    cases = []
    for i, (W, H, steps) in enumerate([(64, 64, 300), (150, 90, 257), (192, 192, 120), (31, 17, 999)]):
        g0 = O.fill(40 + i, W, H, *FILLS["twelve_materials"])
        w = S.World(ctx(), rs("twelve_materials"), W, H)
        w.frame = 1
        w.upload(g0)
        b.add(w, steps)
        cases.append((w, g0, steps))
    b.run(mode=1, graph=True)
    for w, g0, steps in cases:
        np.testing.assert_array_equal(w.download(), O.run(blob, g0, 2, steps))
    b.close()


def test_batch_wall_changing_table_against_oracle():
    """A full-table rule set whose entries change WALL cells (sand sets the WALL
    below it to water): the batched kernel's edge tiles restore the world's WALL
    border, in single-tile worlds and in worlds that need rounds."""
    from tests.test_gpu_parity import _erode_rules
    text = _erode_rules()
    rset = S.RuleSet(ctx(), text)
    assert rset.kernel_form == "table"
    blob = O.compile_ruleset(L.parse_string(text))
    b = S.Batch(ctx())
    cases = []
    for i, (W, H, steps) in enumerate([(64, 64, 120), (150, 90, 77), (300, 260, 41), (31, 17, 200)]):
        g0 = O.fill(70 + i, W, H, [0, 3, 4, 5], [50, 25, 15, 10])
        w = S.World(ctx(), rset, W, H)
        w.frame = 1
        w.upload(g0)
        b.add(w, steps)
        cases.append((w, g0, steps))
    b.run(mode=1, graph=True)
    for w, g0, steps in cases:
        want = O.run(blob, g0, 2, steps)
        assert (want == 3).sum() < (g0 == 3).sum()
        np.testing.assert_array_equal(w.download(), want)
    b.close()


def test_batch_three_class_table_against_oracle():
    """A full table with six hash classes per tuple (tile codes left unscaled)
    in the batched kernel: single-tile worlds and worlds that need rounds, with
    unknown ids (NULL once simulated) in some of them, against the oracle."""
    from tests.test_gpu_parity import _three_class_rules
    text = _three_class_rules()
    rset = S.RuleSet(ctx(), text)
    assert rset.kernel_form == "table"
    blob = O.compile_ruleset(L.parse_string(text))
    b = S.Batch(ctx())
    cases = []
    for i, (W, H, steps) in enumerate([(64, 64, 90), (150, 90, 57), (300, 260, 33), (31, 17, 120), (226, 114, 40)]):
        g0 = O.fill(90 + i, W, H, [0, 3, 4, 5], [45, 25, 20, 10])
        if i % 2 == 0:
            g0[H // 3:H // 3 + 4, W // 4:W // 4 + 6] = 77
        w = S.World(ctx(), rset, W, H)
        w.frame = 1
        w.upload(g0)
        b.add(w, steps)
        cases.append((w, g0, steps))
    b.run(mode=1, graph=True)
    for w, g0, steps in cases:
        np.testing.assert_array_equal(w.download(), O.run(blob, g0, 2, steps))
    b.close()
