"""Headless input loop (sandengine_amd.InputLoop) against the event handling of
sandengine-core/src/lib.rs:19-120: which modification each frame queues.

The simulation is a stand-in that records what run() receives, so these tests
need no GPU; tests/test_gpu_parity.py runs the loop on a real world.
"""
import numpy as np

import sandengine_amd as S


class FakeSim:
    def __init__(self, size):
        self.size = size
        self.params = S.Params()
        self.modifications = []
        self.runs = []

    def run(self):
        self.runs.append(list(self.modifications))
        self.modifications.clear()
        self.params.frame += 1


def loop(size=(64, 64), ruleset="four_materials"):
    sim = FakeSim(size)
    return sim, S.InputLoop(sim, S.parse_string(S.bundled_ruleset(ruleset)))


def test_selectable_materials_follow_the_rule_file():
    sim, il = loop()
    # EMPTY is selectable, NULL and WALL are not; rule-file materials default to selectable
    assert [m.name for m in il.selectable] == ["EMPTY", "sand", "water", "stone"]


def test_no_modification_without_the_button_held():
    sim, il = loop()
    il.cursor_moved(100, 100, (640, 480))
    il.frame()
    assert sim.runs == [[]] and sim.params.frame == 1


def test_brush_at_the_cursor_with_defaults():
    sim, il = loop()
    il.cursor_moved(320, 240, (640, 480))
    il.mouse_left(True)
    il.frame()
    m = sim.runs[0][0]
    assert m.position == (32, 32) and m.mod_shape == S.MODSHAPE_CIRCLE
    assert m.mod_size == 5 and m.mod_matID == 0  # SandMaterial::default()


def test_number_keys_pick_selectable_materials():
    sim, il = loop()
    il.key(1)
    assert sim.params.brushMaterial.name == "sand"
    il.key("3")
    assert sim.params.brushMaterial.name == "stone"
    il.key(7)  # past the list: unchanged
    assert sim.params.brushMaterial.name == "stone"
    il.key("F1")  # any other key selects index 0
    assert sim.params.brushMaterial.name == "EMPTY"


def test_wheel_changes_brush_size_by_its_sign():
    sim, il = loop()
    il.wheel(3.0)
    assert sim.params.brushSize == 6
    il.wheel(-0.5)
    assert sim.params.brushSize == 5
    for _ in range(10):
        il.wheel(-1.0)
    assert sim.params.brushSize == 1  # never below 1
    il.wheel(0.0)  # f32::signum(+0.0) is +1
    assert sim.params.brushSize == 2
    il.wheel(float("nan"))  # NaN as i32 is 0
    assert sim.params.brushSize == 2


def test_position_is_truncated_f32_product():
    sim, il = loop(size=(97, 61))
    il.mouse_left(True)
    for (x, y, dims) in [(639, 479, (640, 480)), (1, 1, (640, 480)), (333.7, 12.2, (1000, 700))]:
        il.cursor_moved(x, y, dims)
        il.frame()
        fx = np.float32(np.float32(x) / np.float32(dims[0])) * np.float32(97)
        fy = np.float32(np.float32(y) / np.float32(dims[1])) * np.float32(61)
        assert sim.runs[-1][0].position == (int(fx), int(fy))
    il.cursor_moved(-10, 5, (640, 480))  # outside the window: negative, truncated toward zero
    il.frame()
    assert sim.runs[-1][0].position[0] == int(np.float32(np.float32(-10) / np.float32(640)) * np.float32(97))
