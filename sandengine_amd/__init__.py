"""sandengine_amd: MI355X-native sandengine grid-update step (Python host side).

Mirrors the reference's interfaces over the C ABI in include/sandengine_hip.h:

* ``parse_string(yaml)``    sandengine_lang::parse_string (sandengine-lang/src/parser.rs:93)
  returning a ``ParsingResult`` with ``rules`` / ``types`` / ``materials``;
  errors raise ``ParsingErr`` whose ``kind`` is the reference's category
  (MissingField, InvalidType, NotFound, NotRecognized; parser.rs:44-73).
* ``create_glsl_from_parser`` sandengine-lang/src/lib.rs:17 (returns the two
  generated GLSL texts instead of writing gen/*.glsl).
* ``Simulation``            sandengine_core::simulation::Simulation
  (sandengine-core/src/simulation.rs:97-253): ``run()`` is one step
  (frame += 1, then the step; frame 1 clears), ``modifications`` is the list the
  caller pushes SimModification records onto, applied to the next run() and
  then cleared, ``params.frame`` is the frame counter.

The compute path is the HIP library libsandengine_hip.so built in this
directory; there is no CPU fallback: without the library every entry point
raises ``SandengineError``.
"""
from __future__ import annotations

import ctypes
import math
import json
import os
import struct
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsandengine_hip.so")
RULESETS_DIR = os.path.join(HERE, "rulesets")

MODSHAPE_CIRCLE = 0   # simulation.rs:41
MODSHAPE_SQUARE = 1   # simulation.rs:42
MAX_MODIFICATIONS = 256  # simulation.rs:43

SE_OK, SE_ERR_ARG, SE_ERR_PARSE, SE_ERR_CODEGEN, SE_ERR_JIT, SE_ERR_HIP, SE_ERR_OOM, SE_ERR_FULL, \
    SE_ERR_STATE = 0, -1, -2, -3, -4, -5, -6, -7, -8


class SandengineError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code
        self.message = message


class ParsingErr(SandengineError):
    """Rule-file error; ``kind`` is the reference's ParsingErr category."""

    def __init__(self, kind: str, message: str):
        super().__init__(SE_ERR_PARSE, message)
        self.kind = kind


class SeModification(ctypes.Structure):
    _fields_ = [("position", ctypes.c_int32 * 2), ("mod_shape", ctypes.c_int32),
                ("mod_size", ctypes.c_int32), ("mod_matID", ctypes.c_int32),
                ("_pad4", ctypes.c_int32 * 3)]


assert ctypes.sizeof(SeModification) == 32


class SeGridStep(ctypes.Structure):
    """se_grid_step: one caller-owned grid of se_grids_step_noalloc."""
    _fields_ = [("grid", ctypes.c_void_p * 2), ("w", ctypes.c_uint32), ("h", ctypes.c_uint32),
                ("pitch", ctypes.c_uint32), ("cur", ctypes.c_int32), ("frame", ctypes.c_int32),
                ("nsteps", ctypes.c_uint32), ("mod_first", ctypes.c_uint32), ("mod_count", ctypes.c_uint32)]

_lib = None


def lib():
    """The HIP library; raises if it has not been built (no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SandengineError(SE_ERR_STATE, f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        c_p, c_sz, c_u32, c_i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_int32
        PP = ctypes.POINTER(ctypes.c_void_p)
        L.se_last_error.restype = ctypes.c_char_p
        L.se_version.restype = ctypes.c_char_p
        L.se_free.argtypes = [c_p]
        L.se_parse_ir_json.argtypes = [ctypes.c_char_p, c_sz, PP, PP]
        L.se_emit_glsl.argtypes = [ctypes.c_char_p, c_sz, PP, PP]
        L.se_emit_hip.argtypes = [ctypes.c_char_p, c_sz, PP]
        L.se_jit_check.argtypes = [ctypes.c_char_p, c_sz, ctypes.POINTER(c_sz)]
        L.se_ctx_create.argtypes = [ctypes.c_int, PP]
        L.se_ctx_destroy.argtypes = [c_p]
        L.se_ctx_mem_info.argtypes = [c_p, ctypes.POINTER(c_sz), ctypes.POINTER(c_sz)]
        L.se_ruleset_from_yaml.argtypes = [c_p, ctypes.c_char_p, c_sz, PP]
        L.se_ruleset_destroy.argtypes = [c_p]
        L.se_ruleset_material_id.argtypes = [c_p, ctypes.c_char_p, ctypes.POINTER(c_i32)]
        L.se_ruleset_material_count.argtypes = [c_p, ctypes.POINTER(c_i32)]
        L.se_ruleset_kernel_form.argtypes = [c_p, ctypes.POINTER(c_i32)]
        L.se_world_create.argtypes = [c_p, c_p, c_u32, c_u32, PP]
        L.se_world_destroy.argtypes = [c_p]
        L.se_world_upload.argtypes = [c_p, ctypes.POINTER(c_u32), c_sz]
        L.se_world_download.argtypes = [c_p, ctypes.POINTER(c_u32), c_sz]
        L.se_world_push_modification.argtypes = [c_p, ctypes.POINTER(SeModification)]
        L.se_world_step.argtypes = [c_p, c_u32]
        L.se_world_step_blocked.argtypes = [c_p, c_u32, c_u32]
        L.se_world_get_frame.argtypes = [c_p, ctypes.POINTER(c_i32)]
        L.se_world_save.argtypes = [c_p, ctypes.c_char_p]
        L.se_world_set_render.argtypes = [c_p, ctypes.c_int]
        L.se_world_render_download.argtypes = [c_p, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
        L.se_world_render_ptrs.argtypes = [c_p, PP, PP, ctypes.POINTER(c_u32)]
        L.se_world_render_timed.argtypes = [c_p, c_u32, ctypes.POINTER(ctypes.c_float)]
        L.se_world_load.argtypes = [c_p, c_p, ctypes.c_char_p, PP]
        L.se_world_set_frame.argtypes = [c_p, c_i32]
        L.se_world_sync.argtypes = [c_p]
        L.se_world_device_ptr.argtypes = [c_p, PP, ctypes.POINTER(c_u32)]
        L.se_world_fill_seeded.argtypes = [c_p, ctypes.c_uint64, ctypes.POINTER(c_u32),
                                           ctypes.POINTER(c_u32), ctypes.c_int, c_u32, c_u32]
        L.se_world_step_timed.argtypes = [c_p, c_u32, ctypes.c_int, c_u32, ctypes.POINTER(ctypes.c_float)]
        L.se_world_set_variant.argtypes = [c_p, c_i32, c_i32]
        L.se_world_autotune.argtypes = [c_p, c_u32, c_u32, ctypes.POINTER(c_i32), ctypes.POINTER(c_u32),
                                        ctypes.POINTER(ctypes.c_float)]
        L.se_world_set_sleep.argtypes = [c_p, ctypes.c_int]
        L.se_world_copy_timed.argtypes = [c_p, c_u32, ctypes.POINTER(ctypes.c_float)]
        L.se_serve_create.argtypes = [c_p, ctypes.POINTER(c_p), ctypes.c_int, c_p, PP]
        L.se_serve_destroy.argtypes = [c_p]
        L.se_serve_set_fill.argtypes = [c_p, ctypes.c_int, ctypes.POINTER(c_u32), ctypes.POINTER(c_u32), ctypes.c_int]
        L.se_serve_run.argtypes = [c_p, c_p]
        L.se_serve_request_count.argtypes = [c_p]
        L.se_serve_request_count.restype = ctypes.c_uint64
        L.se_serve_request_info.argtypes = [c_p, ctypes.c_uint64, c_p]
        L.se_serve_request_mod.argtypes = [c_p, ctypes.c_uint64, c_u32, ctypes.POINTER(c_u32), c_p]
        L.se_fnv1a64.argtypes = [c_p, c_sz]
        L.se_fnv1a64.restype = ctypes.c_uint64
        L.se_batch_create.argtypes = [c_p, PP]
        L.se_batch_destroy.argtypes = [c_p]
        L.se_batch_tuned.argtypes = [c_p, c_p, ctypes.POINTER(c_i32), ctypes.POINTER(c_u32), ctypes.POINTER(c_i32)]
        L.se_batch_add.argtypes = [c_p, c_p, c_u32]
        L.se_batch_clear.argtypes = [c_p]
        L.se_batch_run.argtypes = [c_p, ctypes.c_int, c_i32, c_u32, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
        L.se_batch_autotune.argtypes = [c_p, c_u32, ctypes.POINTER(c_i32), ctypes.POINTER(c_u32),
                                        ctypes.POINTER(c_i32), ctypes.POINTER(ctypes.c_float)]
        L.se_variant_info.argtypes = [c_p, c_i32, c_i32] + [ctypes.POINTER(c_i32)] * 3
        L.se_grids_step_noalloc.argtypes = [c_p, c_p, ctypes.POINTER(SeGridStep), c_u32,
                                            ctypes.POINTER(SeModification), c_u32, c_p]
        _lib = L
    return _lib


def _check(rc: int):
    if rc != SE_OK:
        raise SandengineError(rc, lib().se_last_error().decode(errors="replace"))


def _take_str(p: ctypes.c_void_p) -> Optional[str]:
    if not p.value:
        return None
    s = ctypes.string_at(p.value).decode()
    lib().se_free(p)
    return s


# ---------------------------------------------------------------- language layer
@dataclass
class SandRule:
    name: str
    ruletype: str
    if_conds: List[str]
    do_actions: List[str]
    probabilities: List[float]
    mirror: bool
    precondition: Optional[str]
    used: bool


@dataclass
class SandType:
    id: int
    name: str
    inherits: str
    children: List[str]
    base_rules: List[str]


@dataclass
class SandMaterial:
    id: int
    name: str
    mattype: str
    color: List[float]
    emission: List[float]
    selectable: bool
    density: float
    extra_rules: List[str]


@dataclass
class ParsingResult:
    rules: List[SandRule]
    types: List[SandType]
    materials: List[SandMaterial]
    source: str = ""
    ir: dict = field(default_factory=dict)

    def material_id(self, name: str) -> int:
        for m in self.materials:
            if m.name == name:
                return m.id
        raise KeyError(name)


def parse_string(yaml_text: str) -> ParsingResult:
    b = yaml_text.encode()
    out, kind = ctypes.c_void_p(), ctypes.c_void_p()
    rc = lib().se_parse_ir_json(b, len(b), ctypes.byref(out), ctypes.byref(kind))
    if rc == SE_ERR_PARSE:
        k = _take_str(kind) or "Yaml"
        raise ParsingErr(k, lib().se_last_error().decode(errors="replace"))
    _check(rc)
    ir = json.loads(_take_str(out))
    return ParsingResult([SandRule(**r) for r in ir["rules"]], [SandType(**t) for t in ir["types"]],
                         [SandMaterial(**m) for m in ir["materials"]], yaml_text, ir)


def parse_path(path: str) -> ParsingResult:
    """sandengine-lang/src/lib.rs:10-13."""
    with open(path) as f:
        return parse_string(f.read())


def create_glsl_from_parser(result: ParsingResult) -> Tuple[str, str]:
    """(gen/materials.glsl, gen/rules.glsl) text for a parsed rule file."""
    b = result.source.encode()
    m, r = ctypes.c_void_p(), ctypes.c_void_p()
    _check(lib().se_emit_glsl(b, len(b), ctypes.byref(m), ctypes.byref(r)))
    return _take_str(m), _take_str(r)


def emit_hip(yaml_text: str) -> str:
    b = yaml_text.encode()
    s = ctypes.c_void_p()
    rc = lib().se_emit_hip(b, len(b), ctypes.byref(s))
    _check(rc)
    return _take_str(s)


def jit_check(yaml_text: str) -> int:
    """hiprtc-compiles the rule set for gfx950 (no device needed); returns code bytes."""
    b = yaml_text.encode()
    n = ctypes.c_size_t()
    _check(lib().se_jit_check(b, len(b), ctypes.byref(n)))
    return n.value


def bundled_ruleset(name: str) -> str:
    with open(os.path.join(RULESETS_DIR, name + ".yaml")) as f:
        return f.read()


BUNDLED = ["four_materials", "twelve_materials", "features"]


def jit_check_bundled(verbose: bool = False):
    for n in BUNDLED:
        size = jit_check(bundled_ruleset(n))
        if verbose:
            print(f"hiprtc gfx950 {n}: {size} bytes")


# ---------------------------------------------------------------- device runtime
class Context:
    """One device: JIT cache + stream (se_ctx)."""

    def __init__(self, device: int = 0):
        p = ctypes.c_void_p()
        _check(lib().se_ctx_create(device, ctypes.byref(p)))
        self._p = p
        self.device = device

    def mem_info(self) -> Tuple[int, int]:
        t, f = ctypes.c_size_t(), ctypes.c_size_t()
        _check(lib().se_ctx_mem_info(self._p, ctypes.byref(t), ctypes.byref(f)))
        return t.value, f.value

    def close(self):
        if getattr(self, "_p", None) and self._p.value:
            lib().se_ctx_destroy(self._p)
            self._p = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RuleSet:
    """Parsed + JIT-compiled rule set (se_ruleset)."""

    def __init__(self, ctx: Context, yaml_text: str):
        b = yaml_text.encode()
        p = ctypes.c_void_p()
        rc = lib().se_ruleset_from_yaml(ctx._p, b, len(b), ctypes.byref(p))
        if rc == SE_ERR_PARSE:
            parse_string(yaml_text)  # raises ParsingErr with the category
        _check(rc)
        self._p, self.ctx, self.source = p, ctx, yaml_text
        n = ctypes.c_int32()
        _check(lib().se_ruleset_material_count(self._p, ctypes.byref(n)))
        self.nmat = n.value

    @property
    def kernel_form(self) -> str:
        """How the LDS kernels evaluate a block: "chain" (the generated rule
        chain), "table" (full transition table) or "coded" (coded table)."""
        f = ctypes.c_int32()
        _check(lib().se_ruleset_kernel_form(self._p, ctypes.byref(f)))
        return ("chain", "table", "coded")[f.value]

    def material_id(self, name: str) -> int:
        v = ctypes.c_int32()
        _check(lib().se_ruleset_material_id(self._p, name.encode(), ctypes.byref(v)))
        return v.value

    def close(self):
        if getattr(self, "_p", None) and self._p.value:
            lib().se_ruleset_destroy(self._p)
            self._p = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class World:
    """One simulated world resident in HBM (se_world)."""

    def __init__(self, ctx: Context, rules: RuleSet, w: int, h: int):
        p = ctypes.c_void_p()
        _check(lib().se_world_create(ctx._p, rules._p, w, h, ctypes.byref(p)))
        self._p, self.ctx, self.rules, self.W, self.H = p, ctx, rules, w, h

    def upload(self, grid: np.ndarray):
        g = np.ascontiguousarray(grid, dtype=np.uint32)
        if g.shape != (self.H, self.W):
            raise ValueError(f"grid shape {g.shape} != {(self.H, self.W)}")
        _check(lib().se_world_upload(self._p, g.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), g.size))

    def download(self) -> np.ndarray:
        g = np.empty((self.H, self.W), np.uint32)
        _check(lib().se_world_download(self._p, g.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), g.size))
        return g

    def push_modification(self, x: int, y: int, shape: int, size: int, mat: int):
        m = SeModification()
        m.position[0], m.position[1] = x, y
        m.mod_shape, m.mod_size, m.mod_matID = shape, size, mat
        _check(lib().se_world_push_modification(self._p, ctypes.byref(m)))

    def step(self, n: int = 1):
        _check(lib().se_world_step(self._p, n))

    def step_blocked(self, n: int = 1, sub: int = 0):
        _check(lib().se_world_step_blocked(self._p, n, sub))

    def step_timed(self, n: int, mode: int = 0, sub: int = 0) -> float:
        ms = ctypes.c_float()
        _check(lib().se_world_step_timed(self._p, n, mode, sub, ctypes.byref(ms)))
        return ms.value

    def copy_timed(self, reps: int) -> float:
        """GPU ms of `reps` streaming copies of the grid (HBM reference)."""
        ms = ctypes.c_float()
        _check(lib().se_world_copy_timed(self._p, reps, ctypes.byref(ms)))
        return ms.value

    def set_variant(self, plain: int = 3, lds: int = -1):
        _check(lib().se_world_set_variant(self._p, plain, lds))

    def autotune(self, nsteps: int, rounds: int = 2) -> Tuple[int, int, float]:
        """Times the next min(nsteps, 48) blocked steps with every LDS variant
        (grid, frame and queued modifications restored after each trial) and keeps
        the fastest for step_blocked(n, 0).  Returns (variant, sub-steps per
        launch, ms per step)."""
        v, k, ms = ctypes.c_int32(), ctypes.c_uint32(), ctypes.c_float()
        _check(lib().se_world_autotune(self._p, nsteps, rounds, ctypes.byref(v), ctypes.byref(k), ctypes.byref(ms)))
        return v.value, k.value, ms.value

    def set_sleep(self, enable: bool = True):
        """Tile sleeping of step_blocked (default on; results are the same)."""
        _check(lib().se_world_set_sleep(self._p, 1 if enable else 0))

    def variant_info(self, kind: int, index: int) -> Tuple[int, int, int]:
        a, b, c = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _check(lib().se_variant_info(self._p, kind, index, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def set_render(self, enable: bool = True) -> None:
        """Colour and light side-outputs of every step (se_world_set_render)."""
        _check(lib().se_world_set_render(self._p, 1 if enable else 0))

    def render_download(self) -> Tuple[np.ndarray, np.ndarray]:
        """(colour, light) of the last step, each (H, W, 4) float32 RGBA."""
        col = np.empty((self.H, self.W, 4), np.float32)
        light = np.empty((self.H, self.W, 4), np.float32)
        fp = ctypes.POINTER(ctypes.c_float)
        _check(lib().se_world_render_download(self._p, col.ctypes.data_as(fp), light.ctypes.data_as(fp)))
        return col, light

    def render_timed(self, n: int) -> float:
        """GPU ms of n more render launches on the last step's grids (timing
        only: the light advances n propagation steps, the grid is unchanged)."""
        ms = ctypes.c_float()
        _check(lib().se_world_render_timed(self._p, n, ctypes.byref(ms)))
        return ms.value

    def save(self, path: str) -> None:
        """Writes the grid, frame and queued modifications (se_world_save)."""
        _check(lib().se_world_save(self._p, os.fsencode(path)))

    @classmethod
    def load(cls, ctx: "Context", ruleset: "RuleSet", path: str) -> "World":
        """A world saved by save(); the rule set must be the one it was saved with."""
        p = ctypes.c_void_p()
        _check(lib().se_world_load(ctx._p, ruleset._p, os.fsencode(path), ctypes.byref(p)))
        with open(path, "rb") as f:
            W, H = struct.unpack("<8sIII", f.read(20))[2:4]
        w = cls.__new__(cls)
        w._p, w.ctx, w.rules, w.W, w.H = p, ctx, ruleset, W, H
        return w

    @property
    def frame(self) -> int:
        v = ctypes.c_int32()
        _check(lib().se_world_get_frame(self._p, ctypes.byref(v)))
        return v.value

    @frame.setter
    def frame(self, f: int):
        _check(lib().se_world_set_frame(self._p, f))

    def sync(self):
        _check(lib().se_world_sync(self._p))

    def fill_seeded(self, seed: int, ids: Sequence[int], percents: Sequence[int], stone_rows: int = 0,
                    stone_id: int = 0):
        """Device-side seeded fill; percents are integers summing to 100."""
        if sum(percents) != 100 or len(ids) != len(percents):
            raise ValueError("percents must sum to 100, one per id")
        thr = percent_thresholds(percents)
        a = (ctypes.c_uint32 * len(ids))(*ids)
        t = (ctypes.c_uint32 * len(ids))(*thr)
        _check(lib().se_world_fill_seeded(self._p, seed, a, t, len(ids), stone_rows, stone_id))

    def close(self):
        if getattr(self, "_p", None) and self._p.value:
            lib().se_world_destroy(self._p)
            self._p = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ServeConfig(ctypes.Structure):
    _fields_ = [("rate", ctypes.c_double), ("duration_s", ctypes.c_double), ("seed", ctypes.c_uint64),
                ("side_min", ctypes.c_uint32), ("side_max", ctypes.c_uint32), ("steps_min", ctypes.c_uint32),
                ("steps_max", ctypes.c_uint32), ("mods_max", ctypes.c_uint32), ("quantum", ctypes.c_uint32),
                ("quantum_small", ctypes.c_uint32), ("kernel", ctypes.c_int), ("max_inflight", ctypes.c_int),
                ("drain_timeout_s", ctypes.c_double), ("hbm_budget", ctypes.c_size_t),
                ("verify_every", ctypes.c_uint32), ("slo_ms", ctypes.c_double), ("capacity_gcups", ctypes.c_double)]


class ServeStats(ctypes.Structure):
    _fields_ = [("requests", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("rejected", ctypes.c_uint64),
                ("unfinished", ctypes.c_uint64), ("rounds", ctypes.c_uint64), ("p50_ms", ctypes.c_double),
                ("p99_ms", ctypes.c_double), ("mean_ms", ctypes.c_double), ("max_ms", ctypes.c_double),
                ("wall_s", ctypes.c_double), ("cell_updates", ctypes.c_double), ("gcups", ctypes.c_double),
                ("offered_gcups", ctypes.c_double), ("peak_bytes", ctypes.c_size_t),
                ("processed_gcups", ctypes.c_double), ("p50_ms_censored", ctypes.c_double),
                ("p99_ms_censored", ctypes.c_double), ("shed", ctypes.c_uint64), ("verified", ctypes.c_uint64),
                ("shed_cell_frac", ctypes.c_double), ("slo_met_frac", ctypes.c_double),
                ("capacity_gcups_est", ctypes.c_double), ("round_fixed_ms", ctypes.c_double),
                ("pred_ratio_p50", ctypes.c_double), ("pred_ratio_p99", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServeRequest(ctypes.Structure):
    _fields_ = [("arrival_s", ctypes.c_double), ("done_s", ctypes.c_double), ("side", ctypes.c_uint32),
                ("steps", ctypes.c_uint32), ("ruleset", ctypes.c_uint32), ("nmods", ctypes.c_uint32),
                ("fill_seed", ctypes.c_uint64), ("state", ctypes.c_int), ("verified", ctypes.c_int),
                ("hash", ctypes.c_uint64)]


def fnv1a64(a: np.ndarray) -> int:
    b = np.ascontiguousarray(a)
    return lib().se_fnv1a64(b.ctypes.data_as(ctypes.c_void_p), b.nbytes)


def percent_thresholds(percents: Sequence[int]) -> List[int]:
    """Cumulative integer percents -> u32 thresholds floor(cum * 2^32 / 100)."""
    thr, cum = [], 0
    for p in percents:
        cum += p
        thr.append(min((cum << 32) // 100, 0xFFFFFFFF))
    return thr


class Serve:
    """Serving loop with continuous batching under an HBM budget (se_serve)."""

    def __init__(self, ctx: Context, rulesets: Sequence[RuleSet], fills, **cfg):
        c = ServeConfig()
        for k, v in cfg.items():
            setattr(c, k, v)
        arr = (ctypes.c_void_p * len(rulesets))(*[r._p.value for r in rulesets])
        p = ctypes.c_void_p()
        _check(lib().se_serve_create(ctx._p, arr, len(rulesets), ctypes.byref(c), ctypes.byref(p)))
        self._p, self.ctx, self.rulesets, self.cfg = p, ctx, list(rulesets), c
        for i, (ids, pct) in enumerate(fills):
            a = (ctypes.c_uint32 * len(ids))(*ids)
            t = (ctypes.c_uint32 * len(ids))(*percent_thresholds(pct))
            _check(lib().se_serve_set_fill(self._p, i, a, t, len(ids)))

    def run(self) -> dict:
        st = ServeStats()
        _check(lib().se_serve_run(self._p, ctypes.byref(st)))
        return st.as_dict()

    def __len__(self):
        return lib().se_serve_request_count(self._p)

    def request(self, i: int) -> dict:
        r = ServeRequest()
        _check(lib().se_serve_request_info(self._p, i, ctypes.byref(r)))
        d = {k: getattr(r, k) for k, _ in r._fields_}
        mods = []
        for j in range(r.nmods):
            step = ctypes.c_uint32()
            m = SeModification()
            _check(lib().se_serve_request_mod(self._p, i, j, ctypes.byref(step), ctypes.byref(m)))
            mods.append((step.value, [m.position[0], m.position[1], m.mod_shape, m.mod_size, m.mod_matID]))
        d["mods"] = mods
        return d

    def close(self):
        if getattr(self, "_p", None) and self._p.value:
            lib().se_serve_destroy(self._p)
            self._p = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def grids_step_noalloc(ctx: "Context", rules: "RuleSet", grids: Sequence[dict], mods=(), stream: int = 0):
    """se_grids_step_noalloc: the batched step over caller-owned device grids,
    allocating nothing.  Each grid is a dict with keys grid0, grid1 (device
    addresses of pitch*h u32 cells), w, h, pitch, cur, frame, nsteps and
    optionally mod_first, mod_count (into `mods`: (x, y, shape, size, matID)
    tuples).  `stream` is a hipStream_t handle (0: the context's stream).  The
    work is enqueued, not waited for; returns the (cur, frame) of each grid after
    it."""
    arr = (SeGridStep * max(1, len(grids)))()
    for i, g in enumerate(grids):
        a = arr[i]
        a.grid[0], a.grid[1] = g["grid0"], g["grid1"]
        a.w, a.h, a.pitch, a.cur, a.frame, a.nsteps = g["w"], g["h"], g["pitch"], g["cur"], g["frame"], g["nsteps"]
        a.mod_first, a.mod_count = g.get("mod_first", 0), g.get("mod_count", 0)
    m = (SeModification * max(1, len(mods)))()
    for i, md in enumerate(mods):
        m[i].position[0], m[i].position[1] = md[0], md[1]
        m[i].mod_shape, m[i].mod_size, m[i].mod_matID = md[2], md[3], md[4]
    _check(lib().se_grids_step_noalloc(ctx._p, rules._p, arr, len(grids), m, len(mods), stream or None))
    return [(arr[i].cur, arr[i].frame) for i in range(len(grids))]


class Batch:
    """Many worlds advanced together (se_batch): looped plain kernel (mode 0) or
    the batched LDS kernel, launched one workgroup per work item (mode 2),
    persistently (mode 3) or as tuned (mode 1), optionally through one hipGraph."""

    def __init__(self, ctx: Context):
        p = ctypes.c_void_p()
        _check(lib().se_batch_create(ctx._p, ctypes.byref(p)))
        self._p, self.ctx, self.worlds = p, ctx, []

    def add(self, world: World, nsteps: int):
        _check(lib().se_batch_add(self._p, world._p, nsteps))
        self.worlds.append((world, nsteps))

    def clear(self):
        _check(lib().se_batch_clear(self._p))
        self.worlds = []

    def run(self, mode: int = 1, variant: int = -1, k: int = 0, graph: bool = False) -> float:
        ms = ctypes.c_float()
        _check(lib().se_batch_run(self._p, mode, variant, k, 1 if graph else 0, ctypes.byref(ms)))
        return ms.value

    def autotune(self, rounds: int = 3) -> Tuple[int, int, bool, float]:
        """(LDS variant, k, persistent launch, ns per 1000 cell-updates) of the last group tuned."""
        v, k, p, s = ctypes.c_int32(), ctypes.c_uint32(), ctypes.c_int32(), ctypes.c_float()
        _check(lib().se_batch_autotune(self._p, rounds, ctypes.byref(v), ctypes.byref(k), ctypes.byref(p),
                                       ctypes.byref(s)))
        return v.value, k.value, bool(p.value), s.value

    def tuned(self, rules: "RuleSet") -> Tuple[int, int, bool]:
        """(LDS variant, k, persistent launch) tuned for one rule set's group."""
        v, k, p = ctypes.c_int32(), ctypes.c_uint32(), ctypes.c_int32()
        _check(lib().se_batch_tuned(self._p, rules._p, ctypes.byref(v), ctypes.byref(k), ctypes.byref(p)))
        return v.value, k.value, bool(p.value)

    def close(self):
        if getattr(self, "_p", None) and self._p.value:
            lib().se_batch_destroy(self._p)
            self._p = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- reference-shaped API
@dataclass
class SimModification:
    """sandengine-core/src/simulation.rs:46-56."""
    position: Tuple[int, int] = (0, 0)
    mod_shape: int = MODSHAPE_CIRCLE
    mod_size: int = 0
    mod_matID: int = 0


@dataclass
class Params:
    """simulation.rs:69-91 Params: the frame the step reads, and the brush state
    the input loop keeps (sandengine-core/src/lib.rs:57-120)."""
    frame: int = 0
    mousePos: Tuple[float, float] = (0.0, 0.0)
    mousePressed: bool = False
    brushSize: int = 5
    brushMaterial: Optional[SandMaterial] = None  # None: SandMaterial::default(), id 0
    time: float = 0.0


class Simulation:
    """Drop-in for sandengine_core::simulation::Simulation over the HIP path."""

    def __init__(self, size: Tuple[int, int], rules, device: int = 0, ctx: Optional[Context] = None,
                 blocked: bool = False):
        self.ctx = ctx or Context(device)
        text = rules.source if isinstance(rules, ParsingResult) else rules
        self.ruleset = RuleSet(self.ctx, text)
        self.size = size
        self.world = World(self.ctx, self.ruleset, size[0], size[1])
        self.modifications: List[SimModification] = []
        self.params = Params()
        self.blocked = blocked

    def run(self):
        """simulation.rs:195-252: one step; queued modifications apply to it."""
        w = self.world
        w.frame = self.params.frame
        for m in self.modifications[:MAX_MODIFICATIONS]:
            w.push_modification(m.position[0], m.position[1], m.mod_shape, m.mod_size, m.mod_matID)
        (w.step_blocked if self.blocked else w.step)(1)
        self.params.frame = w.frame
        self.modifications.clear()

    @property
    def grid(self) -> np.ndarray:
        return self.world.download()

    def upload(self, grid: np.ndarray):
        self.world.upload(grid)


def _f32_to_i32(v) -> int:
    """Rust `f32 as i32`: truncation toward zero, saturating, NaN -> 0."""
    v = float(v)
    if v != v:
        return 0
    if v >= 2147483647.0:
        return 2147483647
    if v <= -2147483648.0:
        return -2147483648
    return int(v)


class InputLoop:
    """Headless form of sandengine_core::run's event loop
    (sandengine-core/src/lib.rs:19-120): scripted window events drive the same
    brush state, and frame() is one MainEventsCleared -- a circle brush at the
    cursor while the left button is held, then Simulation.run().  Rendering,
    the UI and frame timing are not part of it.

    key(k)                 number keys 0-9 pick selectable_materials[k]; any
                           other key picks index 0 (lib.rs:87-101)
    cursor_moved(x, y, d)  mousePos = position / framebuffer size, in f32 (lib.rs:103-106)
    mouse_left(pressed)    left button state (lib.rs:107-113)
    wheel(lines)           brushSize = max(1, brushSize + signum(lines)) (lib.rs:114-118)
    """

    def __init__(self, sim: "Simulation", parsing_result: ParsingResult):
        self.sim = sim
        self.selectable = [m for m in parsing_result.materials if m.selectable]  # lib.rs:21-27

    def key(self, key) -> None:
        digits = {str(d): d for d in range(10)}
        idx = key if isinstance(key, int) and 0 <= key <= 9 else digits.get(str(key), 0)
        if idx < len(self.selectable):
            self.sim.params.brushMaterial = self.selectable[idx]

    def cursor_moved(self, x: float, y: float, framebuffer_dims: Tuple[int, int]) -> None:
        f = np.float32
        self.sim.params.mousePos = (float(f(x) / f(framebuffer_dims[0])), float(f(y) / f(framebuffer_dims[1])))

    def mouse_left(self, pressed: bool) -> None:
        self.sim.params.mousePressed = bool(pressed)

    def wheel(self, lines_y: float) -> None:
        y = float(lines_y)
        sign = 0 if y != y else (1 if math.copysign(1.0, y) > 0 else -1)  # f32::signum: +0.0 -> 1
        self.sim.params.brushSize = max(1, int(self.sim.params.brushSize) + sign)

    def frame(self) -> None:
        p = self.sim.params
        if p.mousePressed:
            f = np.float32
            self.sim.modifications.append(SimModification(
                position=(_f32_to_i32(f(p.mousePos[0]) * f(self.sim.size[0])),
                          _f32_to_i32(f(p.mousePos[1]) * f(self.sim.size[1]))),
                mod_shape=MODSHAPE_CIRCLE, mod_size=int(p.brushSize),
                mod_matID=p.brushMaterial.id if p.brushMaterial is not None else 0))
        self.sim.run()
