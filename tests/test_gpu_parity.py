"""GPU parity: the HIP kernels against the llvmpipe golden fixtures and the oracle.

Every comparison is bit-exact (integer cell ids).  Sizes the oracle finishes in
seconds are compared cell for cell; at the 16384^2 benchmark size the checks are
size-independent properties (plain == LDS-blocked, per-material counts
conserved by swap-only rules, WALL/NULL cells fixed).
"""
import numpy as np
import pytest

import sandengine_amd as S
from oracle import lang_ref as L
from oracle import oracle as O
from tests.golden import cases as C
from tests.golden import load

pytestmark = pytest.mark.gpu

ALL_CASES = load.cases()
_ctx = None
_rulesets = {}


def ctx():
    global _ctx
    if _ctx is None:
        _ctx = S.Context(0)
    return _ctx


def ruleset(name):
    text = load.ruleset_text(name)
    if text is None:
        pytest.skip("reference rule file not present on this machine")
    if name not in _rulesets:
        _rulesets[name] = S.RuleSet(ctx(), text)
    return _rulesets[name]


def run_protocol(rs, g0, first_frame, steps, mods_at, mode, sub=0, checkpoints=()):
    """create (frame 0) -> one clearing step -> upload -> frames first_frame.."""
    H, W = g0.shape
    w = S.World(ctx(), rs, W, H)
    w.step(1)
    w.upload(g0)
    w.frame = first_frame - 1
    stops = sorted(set([k for k in mods_at if k < steps] + [c for c in checkpoints if c < steps] + [steps]))
    snaps = {}
    s = 0
    adv = w.step if mode == "plain" else (lambda n: w.step_blocked(n, sub))
    for stop in stops:
        if s in mods_at:
            for m in mods_at[s]:
                w.push_modification(*m)
        if stop > s:
            adv(stop - s)
            s = stop
        if s in checkpoints:
            snaps[s] = w.download()
    out = w.download()
    frame = w.frame
    w.close()
    return out, snaps, frame


@pytest.mark.parametrize("mode,sub", [("plain", 0), ("blocked", 0), ("blocked", 3)])
@pytest.mark.parametrize("case", ALL_CASES, ids=[c["name"] for c in ALL_CASES])
def test_golden(case, mode, sub):
    print(f"\n[gpu] golden {case['name']} {mode} sub={sub}", flush=True)
    rs = ruleset(case["ruleset"])
    g0 = C.make_input(case)
    cps = tuple(case.get("checkpoints", ()))
    out, snaps, frame = run_protocol(rs, g0, case.get("first_frame", 2), case["steps"], load.mods_at(case),
                                     mode, sub, cps)
    assert frame == case.get("first_frame", 2) - 1 + case["steps"]
    for k, v in snaps.items():
        np.testing.assert_array_equal(v, load.grid(f"{case['name']}@{k}"))
    np.testing.assert_array_equal(out, load.grid(case["name"]))


RANDOM_SHAPES = [(1, 1), (2, 1), (3, 7), (255, 3), (256, 5), (257, 9), (300, 301), (513, 64), (700, 33),
                 (1030, 17), (64, 1000), (112, 112), (226, 114), (97, 250)]


@pytest.mark.parametrize("rsname", ["four", "twelve", "features"])
@pytest.mark.parametrize("shape", RANDOM_SHAPES, ids=[f"{w}x{h}" for w, h in RANDOM_SHAPES])
def test_random_vs_oracle(rsname, shape):
    W, H = shape
    print(f"\n[gpu] random {rsname} {W}x{H}", flush=True)
    rs = ruleset(rsname)
    blob = O.compile_ruleset(L.parse_string(load.ruleset_text(rsname)))
    rng = np.random.RandomState(W * 7919 + H)
    case = dict(ruleset=rsname, W=W, H=H, seed=int(rng.randint(0, 10 ** 6)), specials=0.02)
    g0 = C.make_input(case)
    steps = 37
    mods_at = {0: C._rand_mods(rng, W, H, C.NMAT[rsname], 3), 11: C._rand_mods(rng, W, H, C.NMAT[rsname], 2),
               20: C._rand_mods(rng, W, H, C.NMAT[rsname], 4)}
    first = int(rng.randint(2, 50))
    want = g0
    for s in range(steps):
        want = O.step(blob, want, first + s, mods_at.get(s))
    for mode, sub in [("plain", 0), ("blocked", 0), ("blocked", 5)]:
        got, _, _ = run_protocol(rs, g0, first, steps, mods_at, mode, sub)
        np.testing.assert_array_equal(got, want, err_msg=f"{mode} sub={sub}")


@pytest.mark.parametrize("rsname", ["four", "twelve", "features"])
@pytest.mark.parametrize("wall_brush", [False, True])
def test_interior_tiles_vs_oracle(rsname, wall_brush):
    """A world with tiles wholly inside it and no NULL/WALL cells, so the LDS
    kernel's interior tiles run without the NULL/WALL test; with wall_brush a
    brush places WALL inside the world at a launch's first sub-step, after
    which those tiles must take the test again."""
    W, H = 520, 300
    print(f"\n[gpu] interior {rsname} wall_brush={wall_brush}", flush=True)
    rs = ruleset(rsname)
    blob = O.compile_ruleset(L.parse_string(load.ruleset_text(rsname)))
    rng = np.random.RandomState(4242 + wall_brush)
    g0 = C.make_input(dict(ruleset=rsname, W=W, H=H, seed=int(rng.randint(0, 10 ** 6))))
    steps = 29
    mods_at = {}
    if wall_brush:
        mods_at = {8: [[260, 150, 0, 30, 2], [100, 200, 1, 12, 1], [400, 60, 0, 9, 2]], 16: [[300, 120, 1, 20, 0]]}
    first = 6
    want = g0
    for s in range(steps):
        want = O.step(blob, want, first + s, mods_at.get(s))
    for mode, sub in [("blocked", 0), ("blocked", 8)]:
        got, _, _ = run_protocol(rs, g0, first, steps, mods_at, mode, sub)
        np.testing.assert_array_equal(got, want, err_msg=f"{mode} sub={sub}")


def _petrify_rules():
    """The twelve-material rule set plus a rule whose result holds WALL (acid
    turns sand below it into WALL): too large for the full table and refused by
    the coded table, so the LDS kernels run the generated chain."""
    t = S.bundled_ruleset("twelve_materials")
    t2 = t.replace("    extra_rules: [corrode]\n", "    extra_rules: [corrode, petrify]\n")
    t2 = t2.replace("types:\n", "  petrify:\n    if: DOWN.mat == sand\n    probability: 0.05\n"
                                 "    do: SET DOWN WALL\ntypes:\n", 1)
    assert t2 != t
    return t2


def test_kernel_forms():
    print("\n[gpu] kernel forms", flush=True)
    assert ruleset("four").kernel_form == "table"
    assert ruleset("twelve").kernel_form == "coded"
    assert ruleset("features").kernel_form == "chain"  # Right rules: no mirror factoring
    assert S.RuleSet(ctx(), _petrify_rules()).kernel_form == "chain"


@pytest.mark.parametrize("shape", [(300, 260), (131, 77)], ids=["300x260", "131x77"])
def test_coded_table_fallback_vs_oracle(shape):
    """A rule set that the coded table refuses at load time still steps exactly."""
    W, H = shape
    print(f"\n[gpu] petrify {W}x{H}", flush=True)
    text = _petrify_rules()
    rs = S.RuleSet(ctx(), text)
    blob = O.compile_ruleset(L.parse_string(text))
    rng = np.random.RandomState(W + 17 * H)
    g0 = O.fill(int(rng.randint(0, 10 ** 6)), W, H, [0] + list(range(3, 15)), [40] + [5] * 12)
    steps, first = 23, 9
    mods_at = {5: C._rand_mods(rng, W, H, 15, 3)}
    want = g0
    for s in range(steps):
        want = O.step(blob, want, first + s, mods_at.get(s))
    assert (want == 2).sum() > 0  # WALL was created
    for mode, sub in [("plain", 0), ("blocked", 0), ("blocked", 6)]:
        got, _, _ = run_protocol(rs, g0, first, steps, mods_at, mode, sub)
        np.testing.assert_array_equal(got, want, err_msg=f"{mode} sub={sub}")


def _erode_rules():
    """The four-material rule set plus a rule that replaces WALL (sand sets the
    WALL below it to water) and one that then sinks sand into water.  At the
    world's bottom border the first writes water into the WALL cell outside the
    world (dropped there, as in the reference) and the second swaps the sand
    into it, so the sand turns into water.  Entries of the full table change WALL
    cells, so edge tiles must restore the WALL outside the world after every
    block."""
    t = S.bundled_ruleset("four_materials")
    t2 = t.replace("    extra_rules: [sand_fall]\n", "    extra_rules: [sand_sink, sand_erode, sand_fall]\n", 1)
    t2 = t2.replace("types:\n", "  sand_erode:\n    if: DOWN.mat.type == TYPE_WALL\n    probability: 0.5\n"
                                 "    do: SET DOWN water\n"
                                 "  sand_sink:\n    if: DOWN.mat == water\n    do: SWAP SELF DOWN\ntypes:\n", 1)
    assert t2 != t
    return t2


@pytest.mark.parametrize("shape", [(300, 260), (131, 77)], ids=["300x260", "131x77"])
def test_wall_changing_table_vs_oracle(shape):
    """A full-table rule set whose entries change WALL cells steps exactly in the
    LDS kernel's edge tiles (which then restore the WALL border); the bundled
    four-material set, whose table keeps WALL, skips that restoring."""
    W, H = shape
    print(f"\n[gpu] erode {W}x{H}", flush=True)
    text = _erode_rules()
    rs = S.RuleSet(ctx(), text)
    assert rs.kernel_form == "table"
    blob = O.compile_ruleset(L.parse_string(text))
    rng = np.random.RandomState(W + 31 * H)
    g0 = O.fill(int(rng.randint(0, 10 ** 6)), W, H, [0, 3, 4, 5], [50, 25, 15, 10])
    steps, first = 23, 9
    mods_at = {5: C._rand_mods(rng, W, H, 6, 3)}
    want = g0
    for s in range(steps):
        want = O.step(blob, want, first + s, mods_at.get(s))
    assert (want == 3).sum() < (g0 == 3).sum()  # sand turned into water at the border
    for mode, sub in [("plain", 0), ("blocked", 0), ("blocked", 6)]:
        got, _, _ = run_protocol(rs, g0, first, steps, mods_at, mode, sub)
        np.testing.assert_array_equal(got, want, err_msg=f"{mode} sub={sub}")


def _three_class_rules():
    """The four-material rule set plus two rules with different probabilities,
    so hash lane 1 has two thresholds (three classes) beside the mirror lane's
    two: R = 6 table entries per tuple.  4R = 24 is not a power of two, so the
    LDS tile keeps plain ids and the table offset is scaled per block (the
    bundled set, R = 2, keeps ids times 8)."""
    t = S.bundled_ruleset("four_materials")
    t2 = t.replace("    extra_rules: [sand_fall]\n", "    extra_rules: [sand_sink, sand_fall]\n", 1)
    t2 = t2.replace("    extra_rules: [water_flow]\n", "    extra_rules: [water_jitter, water_flow]\n", 1)
    t2 = t2.replace("types:\n", "  sand_sink:\n    if: DOWN.mat == water\n    probability: 0.3\n"
                                 "    do: SWAP SELF DOWN\n"
                                 "  water_jitter:\n    if: isType_EMPTY(DOWNRIGHT)\n    probability: 0.6\n"
                                 "    do: SWAP SELF DOWNRIGHT\ntypes:\n", 1)
    assert t2.count("probability") == 2
    return t2


@pytest.mark.parametrize("shape", [(300, 260), (131, 77)], ids=["300x260", "131x77"])
def test_three_class_table_vs_oracle(shape):
    """A full table with six hash classes per tuple steps exactly in every kernel."""
    W, H = shape
    print(f"\n[gpu] three-class table {W}x{H}", flush=True)
    text = _three_class_rules()
    rs = S.RuleSet(ctx(), text)
    assert rs.kernel_form == "table"
    blob = O.compile_ruleset(L.parse_string(text))
    rng = np.random.RandomState(W + 37 * H)
    g0 = O.fill(int(rng.randint(0, 10 ** 6)), W, H, [0, 3, 4, 5], [45, 25, 20, 10])
    steps, first = 23, 9
    mods_at = {5: C._rand_mods(rng, W, H, 6, 3)}
    want = g0
    for s in range(steps):
        want = O.step(blob, want, first + s, mods_at.get(s))
    for mode, sub in [("plain", 0), ("blocked", 0), ("blocked", 6)]:
        got, _, _ = run_protocol(rs, g0, first, steps, mods_at, mode, sub)
        np.testing.assert_array_equal(got, want, err_msg=f"{mode} sub={sub}")


def test_clear_frame_and_api_errors():
    rs = ruleset("four")
    w = S.World(ctx(), rs, 40, 30)
    w.upload(np.full((30, 40), 3, np.uint32))
    w.push_modification(20, 15, 0, 5, 5)
    w.step(1)  # frame 1 clears, modifications included
    assert w.frame == 1
    assert not w.download().any()
    with pytest.raises(S.SandengineError):
        w.frame = -3
    for i in range(S.MAX_MODIFICATIONS):
        w.push_modification(0, 0, 0, 1, 3)
    with pytest.raises(S.SandengineError):
        w.push_modification(0, 0, 0, 1, 3)
    with pytest.raises((S.SandengineError, ValueError)):
        w.upload(np.zeros((3, 3), np.uint32))
    w.close()


def test_simulation_api_matches_oracle():
    text = S.bundled_ruleset("twelve_materials")
    blob = O.compile_ruleset(L.parse_string(text))
    sim = S.Simulation((96, 64), text, ctx=ctx())
    sim.run()  # frame 1 clears
    g0 = O.fill(4, 96, 64, [0] + list(range(3, 15)), [40] + [5] * 12)
    sim.upload(g0)
    want = g0
    for f in range(2, 40):
        mods = [[48, 10, 0, 8, 7]] if f == 10 else None
        if mods:
            sim.modifications.append(S.SimModification((48, 10), S.MODSHAPE_CIRCLE, 8, 7))
        sim.run()
        want = O.step(blob, want, f, mods)
    assert sim.params.frame == 39
    np.testing.assert_array_equal(sim.grid, want)


def test_input_loop_matches_oracle():
    """Scripted input through the headless event loop: the brush follows the
    cursor, number keys switch material, the wheel resizes it."""
    text = S.bundled_ruleset("four_materials")
    res = S.parse_string(text)
    blob = O.compile_ruleset(L.parse_string(text))
    sim = S.Simulation((80, 60), text, ctx=ctx())
    il = S.InputLoop(sim, res)
    il.frame()  # frame 1 clears
    g0 = O.fill(9, 80, 60, [0, 3, 4, 5], [50, 25, 15, 10])
    sim.upload(g0)
    want = g0
    script = {3: ("key", 1), 4: ("move", (200, 100)), 5: ("press", True), 9: ("wheel", 2.0), 12: ("key", 2),
              13: ("move", (500, 300)), 16: ("press", False), 20: ("key", 0), 21: ("press", True)}
    for f in range(2, 30):
        ev = script.get(f)
        if ev:
            kind, arg = ev
            if kind == "key":
                il.key(arg)
            elif kind == "move":
                il.cursor_moved(arg[0], arg[1], (640, 480))
            elif kind == "press":
                il.mouse_left(arg)
            else:
                il.wheel(arg)
        queued_before = len(sim.modifications)
        il.frame()
        p = sim.params
        mods = None
        if p.mousePressed:
            mx = int(np.float32(np.float32(p.mousePos[0]) * np.float32(80)))
            my = int(np.float32(np.float32(p.mousePos[1]) * np.float32(60)))
            mods = [[mx, my, S.MODSHAPE_CIRCLE, p.brushSize, p.brushMaterial.id if p.brushMaterial else 0]]
        assert queued_before == 0
        want = O.step(blob, want, f, mods)
    assert sim.params.frame == 29
    np.testing.assert_array_equal(sim.grid, want)


def test_save_load_continues_exactly(tmp_path):
    """A world saved mid-run (with modifications queued) and loaded into a new
    world continues exactly as the original; other rule sets and bad files are
    refused."""
    rs4 = ruleset("four")
    rs12 = ruleset("twelve")
    W, H = 301, 517
    g0 = O.fill(21, W, H, [0] + list(range(3, 15)), [40] + [5] * 12)
    a = S.World(ctx(), rs12, W, H)
    a.step(1)
    a.upload(g0)
    a.frame = 40
    a.step(13)
    a.push_modification(150, 200, 0, 30, 9)
    a.push_modification(10, 10, 1, 5, 0)
    path = str(tmp_path / "w.sew")
    a.save(path)
    b = S.World.load(ctx(), rs12, path)
    assert (b.W, b.H, b.frame) == (W, H, a.frame)
    np.testing.assert_array_equal(b.download(), a.download())
    a.step(20)
    b.step_blocked(20, 0)
    np.testing.assert_array_equal(b.download(), a.download())
    with pytest.raises(S.SandengineError):
        S.World.load(ctx(), rs4, path)
    with open(path, "r+b") as f:
        f.truncate(100)
    with pytest.raises(S.SandengineError):
        S.World.load(ctx(), rs12, path)
    with open(path, "wb") as f:
        f.write(b"not a world at all, not a world at all")
    with pytest.raises(S.SandengineError):
        S.World.load(ctx(), rs12, path)
    a.close()
    b.close()


def test_save_load_large_world_in_chunks(tmp_path):
    """4096^2 (64 MiB per grid: several 16 MiB chunks each way)."""
    rs4 = ruleset("four")
    a = S.World(ctx(), rs4, 4096, 4096)
    a.frame = 1
    _config1(a)
    a.step(7)
    path = str(tmp_path / "big.sew")
    a.save(path)
    b = S.World.load(ctx(), rs4, path)
    a.step(9)
    b.step(9)
    assert b.frame == a.frame
    np.testing.assert_array_equal(b.download(), a.download())
    a.close()
    b.close()


def _config1(w):
    w.fill_seeded(1, [0, 3, 4], [60, 25, 15], stone_rows=(w.H * 3) // 10, stone_id=5)


def test_device_fill_matches_generator():
    rs = ruleset("four")
    w = S.World(ctx(), rs, 333, 211)
    _config1(w)
    want = O.fill(1, 333, 211, [0, 3, 4], [60, 25, 15], (211 * 3) // 10, 5)
    np.testing.assert_array_equal(w.download(), want)
    w.close()


@pytest.mark.parametrize("side,steps", [(4096, 64), (16384, 24)])
def test_large_plain_equals_blocked_and_conserves(side, steps):
    print(f"\n[gpu] large {side}^2 {steps} steps plain vs blocked", flush=True)
    rs = ruleset("four")
    a = S.World(ctx(), rs, side, side)
    b = S.World(ctx(), rs, side, side)
    for w in (a, b):
        w.frame = 1
        _config1(w)
    g0 = a.download()
    counts = np.bincount(g0.ravel(), minlength=6)
    del g0
    a.step(steps)
    b.step_blocked(steps)
    ga, gb = a.download(), b.download()
    assert a.frame == b.frame == 1 + steps
    np.testing.assert_array_equal(ga, gb)
    np.testing.assert_array_equal(np.bincount(ga.ravel(), minlength=6), counts)
    a.close()
    b.close()


@pytest.mark.parametrize("rsname", ["four", "twelve", "features"])
def test_every_kernel_variant_matches_oracle(rsname):
    """All plain and LDS instantiations (rows per wave, store kind, tile, halo,
    threads) give the oracle's grid, with modifications in the run."""
    print(f"\n[gpu] variants {rsname}", flush=True)
    rs = ruleset(rsname)
    blob = O.compile_ruleset(L.parse_string(load.ruleset_text(rsname)))
    W, H, steps, first = 530, 301, 35, 6
    case = dict(ruleset=rsname, W=W, H=H, seed=77, specials=0.01)
    g0 = C.make_input(case)
    rng = np.random.RandomState(5)
    mods_at = {0: C._rand_mods(rng, W, H, C.NMAT[rsname], 4), 17: C._rand_mods(rng, W, H, C.NMAT[rsname], 3)}
    want = g0
    for s in range(steps):
        want = O.step(blob, want, first + s, mods_at.get(s))
    probe = S.World(ctx(), rs, 8, 8)
    n_lds = 0
    while True:
        try:
            probe.variant_info(1, n_lds)
        except S.SandengineError:
            break
        n_lds += 1
    probe.close()
    assert n_lds >= 3
    for pv in range(6):
        for lv, sub in [(-1, 0)] + [(i, s) for i in range(n_lds) for s in (0, 3)] if pv == 3 else [(-1, 0)]:
            w = S.World(ctx(), rs, W, H)
            w.set_variant(pv, lv)
            w.step(1)
            w.upload(g0)
            w.frame = first - 1
            stops = [0, 17, steps]
            for a, b in zip(stops[:-1], stops[1:]):
                for m in mods_at.get(a, []):
                    w.push_modification(*m)
                if lv < 0:
                    w.step(b - a)
                else:
                    w.step_blocked(b - a, sub)
            np.testing.assert_array_equal(w.download(), want, err_msg=f"plain {pv} lds {lv} sub {sub}")
            w.close()


@pytest.mark.parametrize("sub", [0, 5])
def test_full_table_null_tiles_against_oracle(sub):
    """NULL inside the world under the full table: a four-material world with
    unknown ids (NULL once simulated) in a block of interior cells and NULL
    down part of one edge, brushes that place WALL in interior and edge tiles,
    and a later brush that paints over the NULL block, against the oracle."""
    print(f"\n[gpu] full table NULL tiles sub={sub}", flush=True)
    rs = ruleset("four")
    blob = O.compile_ruleset(L.parse_string(load.ruleset_text("four")))
    W, H, steps, first = 530, 301, 30, 6
    g0 = O.fill(77, W, H, [0, 3, 4, 5], [50, 25, 15, 10])
    g0[200:210, 240:260] = 200     # unknown ids: NULL when their block is simulated
    g0[0:40, W - 3] = 1            # NULL on the right edge
    mods_at = {9: [[300, 150, 0, 6, 2], [8, 290, 1, 5, 2], [265, 4, 0, 3, 2]],
               18: [[300, 150, 0, 9, 0], [250, 205, 1, 30, 3]]}
    want = g0
    for s in range(steps):
        want = O.step(blob, want, first + s, mods_at.get(s))
        if s == 17:
            assert (want[200:210, 240:260] == 1).sum() > 0  # the unknown ids became NULL
    assert (want == 1).sum() == 40 and (want[200:210, 240:260] == 3).all()
    got, _, frame = run_protocol(rs, g0, first, steps, mods_at, "blocked", sub)
    assert frame == first - 1 + steps
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("sub", [0, 5])
def test_coded_table_edge_tiles_against_oracle(sub):
    """Edge tiles of the coded table run without the NULL/WALL test while no
    cell inside the world holds either: a 12-material world with no invalid
    ids, then brushes that place WALL and NULL inside the world near its edges
    (those tiles must take the tested path from the brush's sub-step on, and
    from the next launch's load), against the oracle."""
    print(f"\n[gpu] coded table edge tiles sub={sub}", flush=True)
    rs = ruleset("twelve")
    blob = O.compile_ruleset(L.parse_string(load.ruleset_text("twelve")))
    W, H, steps, first = 530, 301, 30, 6
    g0 = C.make_input(dict(ruleset="twelve", W=W, H=H, seed=91))
    mods_at = {9: [[20, 20, 0, 6, 2], [510, 280, 1, 5, 1], [265, 4, 0, 3, 2]],
               18: [[25, 22, 0, 9, 0], [300, 150, 0, 40, 14]]}
    want = g0
    for s in range(steps):
        want = O.step(blob, want, first + s, mods_at.get(s))
    got, _, frame = run_protocol(rs, g0, first, steps, mods_at, "blocked", sub)
    assert frame == first - 1 + steps
    np.testing.assert_array_equal(got, want)
