"""Tile sleeping of the LDS kernel leaves every result unchanged.

Worlds with large settled regions (a stone floor, an empty sky) are advanced by
step_blocked with sleeping on and compared cell for cell with the plain kernel
(and with step_blocked with sleeping off) while everything that must wake a
sleeping tile happens: brushes and explosions inside the settled region,
plain steps and uploads between blocked launches, a change of LDS variant and
of sub-steps per launch, and the clearing frame.
"""
import numpy as np
import pytest

import sandengine_amd as S

pytestmark = pytest.mark.gpu

_ctx = None
_rs = {}


def rs(name):
    global _ctx
    if _ctx is None:
        _ctx = S.Context(0)
    if name not in _rs:
        _rs[name] = S.RuleSet(_ctx, S.bundled_ruleset(name))
    return _rs[name]


def floor_world(r, side, seed, sleep=True):
    w = S.World(_ctx, r, side, side)
    w.frame = 1
    if r.nmat == 6:
        w.fill_seeded(seed, [0, 3, 4], [85, 10, 5], stone_rows=side // 2, stone_id=5)
    else:
        w.fill_seeded(seed, [0, 3, 5], [90, 5, 5], stone_rows=side // 2, stone_id=5)
    w.set_sleep(sleep)
    return w


SCRIPT = [  # (kind, argument)
    ("blocked", 16), ("blocked", 8),
    ("mods", [(300, 900, 0, 40, 0), (700, 1300, 1, 12, 3)]),  # explosion and brush in the stone floor
    ("blocked", 24),
    ("plain", 3), ("blocked", 16),
    ("mods", [(1000, 1500, 0, 64, 4)]),
    ("blocked", 5), ("blocked", 11),
    ("variant", 1), ("blocked", 16), ("variant", -1),
    ("reupload", None), ("blocked", 16),
]


@pytest.mark.parametrize("name", ["four_materials", "twelve_materials"])
def test_sleeping_equals_plain_and_dense(name):
    print(f"\n[gpu] tile sleeping {name}: 1536^2 world, brushes, plain steps, uploads, variant change", flush=True)
    r = rs(name)
    worlds = {"plain": floor_world(r, 1536, 5), "sleep": floor_world(r, 1536, 5, True),
              "dense": floor_world(r, 1536, 5, False)}
    for kind, arg in SCRIPT:
        for label, w in worlds.items():
            if kind == "blocked":
                w.step(arg) if label == "plain" else w.step_blocked(arg, 8)
            elif kind == "plain":
                w.step(arg)
            elif kind == "mods":
                for m in arg:
                    w.push_modification(*m)
            elif kind == "variant" and label != "plain":
                w.set_variant(3, arg)
            elif kind == "reupload":
                w.upload(w.download())
        want = worlds["plain"].download()
        for label in ("sleep", "dense"):
            got = worlds[label].download()
            bad = int((got != want).sum())
            assert bad == 0, f"{label} differs from plain in {bad} cells after {kind} {arg}"
    assert worlds["sleep"].frame == worlds["plain"].frame
    for w in worlds.values():
        w.close()


def test_sleeping_through_the_clearing_frame():
    r = rs("four_materials")
    a, b = floor_world(r, 700, 2), floor_world(r, 700, 2)
    a.step_blocked(16, 8)
    b.step(16)
    for w in (a, b):
        w.frame = 0  # the next step is the clearing frame
    a.step_blocked(20, 8)
    b.step(20)
    assert np.array_equal(a.download(), b.download())
    a.close()
    b.close()
