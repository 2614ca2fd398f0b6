"""se_world_autotune: the world comes back as it was (grid, frame, queued
modifications), the chosen variant is one of the module's, and the steps it
then runs equal the plain kernel's."""
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


FILL = {"four_materials": ([0, 3, 4], [60, 25, 15]), "twelve_materials": ([0] + list(range(3, 15)), [40] + [5] * 12)}
MODS = [(400, 300, 0, 30, 0), (900, 1100, 1, 12, 3)]


def world(r, name, side):
    w = S.World(_ctx, r, side, side)
    w.frame = 1
    ids, pct = FILL[name]
    w.fill_seeded(7, ids, pct, stone_rows=side // 4, stone_id=5)
    for m in MODS:
        w.push_modification(*m)
    return w


@pytest.mark.parametrize("nsteps", [20, 100])
@pytest.mark.parametrize("name", ["four_materials", "twelve_materials"])
def test_autotune_restores_world_and_keeps_results(name, nsteps):
    print(f"\n[gpu] autotune {name} for {nsteps} steps: 1280^2 world with queued brushes", flush=True)
    r = rs(name)
    a, b = world(r, name, 1280), world(r, name, 1280)
    g0, f0 = a.download(), a.frame
    vi, k, ms = a.autotune(nsteps, rounds=1)
    n_lds = 0
    while True:
        try:
            a.variant_info(1, n_lds)
        except S.SandengineError:
            break
        n_lds += 1
    assert 0 <= vi < n_lds
    assert k == a.variant_info(1, vi)[2] // 1000 and ms > 0
    assert a.frame == f0 and np.array_equal(a.download(), g0)
    a.step_blocked(nsteps, 0)  # the queued brushes are still there and land at the first step
    b.step(nsteps)
    assert a.frame == b.frame
    np.testing.assert_array_equal(a.download(), b.download())
    a.close()
    b.close()


def test_autotune_refused_with_render_on():
    r = rs("four_materials")
    w = world(r, "four_materials", 64)
    w.set_render(True)
    with pytest.raises(S.SandengineError):
        w.autotune(20)
    w.close()
