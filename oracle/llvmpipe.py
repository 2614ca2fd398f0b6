"""ctypes front end of oracle/_ref/libse_llvmpipe.so (the reference GLSL under llvmpipe).

TEST INFRASTRUCTURE ONLY: used by tests/golden/make_golden.py to write fixtures,
by tests that pin the oracle where the library is present, and by bench.py's
cpu_baseline leg.  The library is built by ``make -C oracle _ref/libse_llvmpipe.so``
from the reference tree; where it is absent these helpers raise.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

from . import lang_ref as L

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_ref", "libse_llvmpipe.so")
_lib = None


class LlvmpipeError(Exception):
    pass


def available() -> bool:
    return os.path.exists(LIB_PATH)


def lib(nthreads: int = 0):
    global _lib
    if _lib is None:
        if not available():
            raise LlvmpipeError(f"{LIB_PATH} not built")
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.lp_error.restype = ctypes.c_char_p
        _lib.lp_info.restype = ctypes.c_char_p
        _lib.lp_assemble.restype = ctypes.c_void_p
        _lib.lp_assemble.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
        _lib.lp_free.argtypes = [ctypes.c_void_p]
        _lib.lp_ref_gen.restype = ctypes.c_char_p
        _lib.lp_program.argtypes = [ctypes.c_char_p]
        _lib.lp_upload.argtypes = [ctypes.POINTER(ctypes.c_uint32)]
        _lib.lp_download.argtypes = [ctypes.POINTER(ctypes.c_uint32)]
        _lib.lp_run.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int32), ctypes.c_int]
        _lib.lp_time_run.restype = ctypes.c_double
        _lib.lp_download_color.argtypes = [ctypes.POINTER(ctypes.c_float)]
        _lib.lp_download_light.argtypes = [ctypes.POINTER(ctypes.c_float)]
        if _lib.lp_init(nthreads) != 0:
            raise LlvmpipeError(_lib.lp_error().decode())
    return _lib


def info() -> str:
    return lib().lp_info().decode()


def assemble(gen_materials: Optional[str], gen_rules: Optional[str]) -> str:
    p = lib().lp_assemble(gen_materials.encode() if gen_materials is not None else None,
                          gen_rules.encode() if gen_rules is not None else None)
    if not p:
        raise LlvmpipeError(lib().lp_error().decode())
    s = ctypes.string_at(p).decode()
    lib().lp_free(p)
    return s


def ref_gen(which: int) -> str:
    """0: gen/materials.glsl, 1: gen/rules.glsl, 2: gen/falling_sand.glsl of the reference."""
    return lib().lp_ref_gen(which).decode()


def program_for(res: Optional[L.ParsingResult]) -> str:
    """The assembled compute program for a parsed rule file (None: the reference's own)."""
    if res is None:
        return ref_gen(2)
    m, r = L.create_glsl(res)
    return assemble(m, r)


class World:
    """One Simulation (reference protocol): frame starts at 0; run() = frame += 1 + dispatch."""

    _current_program = None

    def __init__(self, program: str, W: int, H: int):
        if World._current_program != program:
            rc = lib().lp_program(program.encode())
            if rc:
                World._current_program = None
                raise LlvmpipeError(f"program failed ({rc}): {lib().lp_error().decode()}")
            World._current_program = program
        if lib().lp_world(W, H):
            raise LlvmpipeError("world creation failed")
        self.W, self.H = W, H

    def upload(self, grid: np.ndarray):
        g = np.ascontiguousarray(grid, dtype=np.uint32)
        assert g.shape == (self.H, self.W)
        if lib().lp_upload(g.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))):
            raise LlvmpipeError("upload failed")

    def download(self) -> np.ndarray:
        g = np.zeros((self.H, self.W), np.uint32)
        if lib().lp_download(g.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))):
            raise LlvmpipeError("download failed")
        return g

    def render(self):
        """(colour, light) RGBA f32 images of the last step, each (H, W, 4)."""
        out = []
        for fn in (lib().lp_download_color, lib().lp_download_light):
            a = np.zeros((self.H, self.W, 4), np.float32)
            if fn(a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))):
                raise LlvmpipeError("render download failed")
            out.append(a)
        return tuple(out)

    def run(self, nsteps: int, mods=None):
        m = np.zeros((max(1, len(mods or [])), 8), np.int32)
        for i, md in enumerate(mods or []):
            m[i, :5] = md[:5]
        if lib().lp_run(nsteps, m.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(mods or [])):
            raise LlvmpipeError(lib().lp_error().decode())

    @property
    def frame(self) -> int:
        return lib().lp_get_frame()

    @frame.setter
    def frame(self, f: int):
        lib().lp_set_frame(int(f))

    def time_run(self, nsteps: int) -> float:
        return lib().lp_time_run(nsteps)


def seeded_run(program: str, grid: np.ndarray, nsteps: int, mods_at=None, first_frame=2,
               checkpoints=()):
    """Protocol: create (frame 0), one clearing step (frame 1), upload, then frames
    first_frame .. first_frame+N-1 (first_frame defaults to 2).
    mods_at: {step_index: [mods]} applied at that step of the N.
    Returns the final grid, or (final, {k: grid after k steps}) with checkpoints."""
    H, W = grid.shape
    w = World(program, W, H)
    w.run(1)
    w.upload(grid)
    w.frame = first_frame - 1
    mods_at = dict(mods_at or {})
    stops = sorted(set(list(mods_at) + [c for c in checkpoints if c < nsteps]))
    snaps = {}
    s = 0
    while s < nsteps:
        if s in mods_at:
            w.run(1, mods_at[s])
            s += 1
        else:
            nxt = min([k for k in stops if k > s] + [nsteps])
            w.run(nxt - s)
            s = nxt
        if s in checkpoints:
            snaps[s] = w.download()
    final = w.download()
    return (final, snaps) if checkpoints else final
