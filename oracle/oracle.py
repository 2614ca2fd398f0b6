"""Oracle front end: rule bytecode compiler, seeded fills, ctypes binding of step_ref.c.

TEST INFRASTRUCTURE ONLY (see step_ref.c).  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg import this module, as the checker.

* ``compile_ruleset`` turns a lang_ref.ParsingResult into the bytecode blob that
  step_ref.c interprets.  Conditions are parsed from the generated GLSL text
  (the strings sandengine-lang/src/parser/rules.rs:207-334 produces), with GLSL
  operator precedence; the if/else pairing follows rules.rs:47-73 exactly,
  including its list-pairing when an ``else`` level has no ``if``.
* ``fill`` is the seeded initial-grid generator of the benchmark workloads
  (SplitMix64 of (seed, cell index); the reference defines none).
"""
from __future__ import annotations

import ctypes
import os
import re
import struct
import subprocess
from typing import List, Optional, Sequence

import numpy as np

from . import lang_ref as L

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libse_oracle.so")

(OP_CONST, OP_DENS, OP_MATID, OP_TYPEID, OP_ISTYPE, OP_MATEQ_CC, OP_MATEQ_CK, OP_RAND, OP_LT,
 OP_LE, OP_GT, OP_GE, OP_EQ, OP_NE, OP_AND, OP_OR, OP_NOT, OP_NEG) = range(1, 19)
ACT_SWAP, ACT_SET = 0, 1
CELLS = {"self": 0, "right": 1, "down": 2, "downright": 3}
LANES = {"x": 0, "y": 1, "z": 2, "w": 3, "r": 0, "g": 1, "b": 2, "a": 3}


class OracleError(Exception):
    pass


def fbits(x: float) -> int:
    return struct.unpack("<i", struct.pack("<f", float(np.float32(x))))[0]


# ---------------------------------------------------------------- expression compiler
_TOK = re.compile(r"\s*(?:(\d+\.\d*(?:[eE][+-]?\d+)?|\.\d+(?:[eE][+-]?\d+)?|\d+(?:[eE][+-]?\d+)?)[fF]?"
                  r"|([A-Za-z_]\w*)|(&&|\|\||==|!=|<=|>=|[<>!().,\-]))")


def _tokens(s: str):
    pos, out = 0, []
    s = s.rstrip()
    while pos < len(s):
        m = _TOK.match(s, pos)
        if not m or m.end() == pos:
            raise OracleError(f"cannot tokenize {s[pos:]!r}")
        pos = m.end()
        if m.group(1) is not None:
            out.append(("num", m.group(1)))
        elif m.group(2) is not None:
            out.append(("id", m.group(2)))
        else:
            out.append(("op", m.group(3)))
    return out


class _Compiler:
    def __init__(self, toks, ctx):
        self.t, self.i, self.ctx = toks, 0, ctx

    def peek(self, k=0):
        return self.t[self.i + k] if self.i + k < len(self.t) else (None, None)

    def take(self, kind=None, val=None):
        tk = self.peek()
        if tk[0] is None or (kind and tk[0] != kind) or (val and tk[1] != val):
            raise OracleError(f"expected {val or kind}, got {tk}")
        self.i += 1
        return tk

    # each returns (code list, type) with type in {"b", "f", "m"}
    def expr(self):
        code, ty = self.and_()
        while self.peek() == ("op", "||"):
            self.take()
            c2, _ = self.and_()
            code, ty = code + c2 + [OP_OR], "b"
        return code, ty

    def and_(self):
        code, ty = self.eq()
        while self.peek() == ("op", "&&"):
            self.take()
            c2, _ = self.eq()
            code, ty = code + c2 + [OP_AND], "b"
        return code, ty

    def eq(self):
        code, ty = self.rel()
        while self.peek() in (("op", "=="), ("op", "!=")):
            op = self.take()[1]
            c2, t2 = self.rel()
            if ty == "m" or t2 == "m":
                if ty != "m" or t2 != "m":
                    raise OracleError("material compared with non-material")
                code = self._mateq(code, c2)
            else:
                code = code + c2 + [OP_EQ]
            if op == "!=":
                code = code + [OP_NOT]
            ty = "b"
        return code, ty

    def _mateq(self, a, b):
        # a/b are ("cell", idx) or ("const", id) markers encoded as tuples in a list
        (ka, va), (kb, vb) = a[0], b[0]
        if ka == "cell" and kb == "cell":
            return [OP_MATEQ_CC, va, vb]
        if ka == "cell" and kb == "const":
            return [OP_MATEQ_CK, va, vb]
        if ka == "const" and kb == "cell":
            return [OP_MATEQ_CK, vb, va]
        return [OP_CONST, fbits(1.0 if va == vb else 0.0)]

    def rel(self):
        code, ty = self.unary()
        ops = {"<": OP_LT, "<=": OP_LE, ">": OP_GT, ">=": OP_GE}
        while self.peek()[0] == "op" and self.peek()[1] in ops:
            op = self.take()[1]
            c2, _ = self.unary()
            code, ty = code + c2 + [ops[op]], "b"
        return code, ty

    def unary(self):
        if self.peek() == ("op", "!"):
            self.take()
            c, _ = self.unary()
            return c + [OP_NOT], "b"
        if self.peek() == ("op", "-"):
            self.take()
            c, _ = self.unary()
            return c + [OP_NEG], "f"
        return self.primary()

    def primary(self):
        kind, val = self.peek()
        if kind == "num":
            self.take()
            return [OP_CONST, fbits(float(val))], "f"
        if kind == "op" and val == "(":
            self.take()
            c = self.expr()
            self.take("op", ")")
            return c
        if kind != "id":
            raise OracleError(f"unexpected token {val!r}")
        self.take()
        if val in ("true", "false"):
            return [OP_CONST, fbits(1.0 if val == "true" else 0.0)], "b"
        if val.startswith("isType_"):
            tname = val[len("isType_"):]
            if tname not in self.ctx["types"]:
                raise OracleError(f"unknown type {tname}")
            self.take("op", "(")
            cell = self.take("id")[1]
            self.take("op", ")")
            if cell not in CELLS:
                raise OracleError(f"unknown cell {cell}")
            return [OP_ISTYPE, CELLS[cell], self.ctx["types"][tname]], "b"
        if val.startswith("MAT_"):
            name = val[4:]
            if name not in self.ctx["mats"]:
                raise OracleError(f"unknown material {name}")
            return [("const", self.ctx["mats"][name])], "m"
        if val.startswith("TYPE_"):
            name = val[5:]
            if name not in self.ctx["types"]:
                raise OracleError(f"unknown type {name}")
            return [OP_CONST, fbits(float(self.ctx["types"][name]))], "f"
        if val == "rand":
            self.take("op", ".")
            lane = self.take("id")[1]
            if len(lane) != 1 or lane not in LANES:
                raise OracleError(f"rand swizzle {lane!r} unsupported")
            return [OP_RAND, LANES[lane]], "f"
        if val in CELLS:
            self.take("op", ".")
            self.take("id", "mat")
            if self.peek() == ("op", "."):
                self.take()
                fld = self.take("id")[1]
                if fld == "density":
                    return [OP_DENS, CELLS[val]], "f"
                if fld == "id":
                    return [OP_MATID, CELLS[val]], "f"
                if fld == "type":
                    return [OP_TYPEID, CELLS[val]], "f"
                raise OracleError(f"field {fld} unsupported")
            return [("cell", CELLS[val])], "m"
        raise OracleError(f"unknown identifier {val!r}")


def compile_expr(s: str, ctx) -> List[int]:
    c = _Compiler(_tokens(s), ctx)
    code, ty = c.expr()
    if c.i != len(c.t):
        raise OracleError(f"trailing tokens in {s!r}")
    if ty == "m" or not all(isinstance(x, int) for x in code):
        raise OracleError(f"material value used outside a comparison in {s!r}")
    return code


_SWAP = re.compile(r"^swap\((\w+), (\w+)\);$")
_SET = re.compile(r"^(\w+) = newCell\(MAT_(\w+), pos\);$")


def compile_actions(s: str, ctx) -> List[List[int]]:
    out = []
    for line in s.split("\n"):
        line = line.strip()
        if not line:
            continue
        m = _SWAP.match(line)
        if m:
            a, b = m.group(1), m.group(2)
            if a not in CELLS or b not in CELLS:
                raise OracleError(f"unknown cell in {line!r}")
            out.append([ACT_SWAP, CELLS[a], CELLS[b]])
            continue
        m = _SET.match(line)
        if m:
            a, mat = m.group(1), m.group(2)
            if a not in CELLS or mat not in ctx["mats"]:
                raise OracleError(f"unknown name in {line!r}")
            out.append([ACT_SET, CELLS[a], ctx["mats"][mat]])
            continue
        raise OracleError(f"unrecognised action {line!r}")
    return out


def _levels(r: L.SandRule):
    """rules.rs:47-73 pairing: (prob, cond, action) levels, then an optional final action."""
    ifs, dos, ps = list(r.if_conds), list(r.do_actions), list(r.probabilities)
    levels, final = [], None
    while True:
        if not ifs and not dos:
            break
        if not ifs and dos:
            final = dos[0]
            break
        levels.append((ps[0], ifs[0], dos[0]))
        ifs, dos, ps = ifs[1:], dos[1:], ps[1:]
        if not dos:
            break
    return levels, final


def compile_ruleset(res: L.ParsingResult) -> np.ndarray:
    mats = {m.name: m.id for m in res.materials}
    types = {t.name: t.id for t in res.types}
    ctx = {"mats": mats, "types": types}
    blob = [len(res.materials)]
    for m in res.materials:
        blob += [types[m.mattype], fbits(m.density)]
    blob.append(len(res.types))
    for t in res.types:
        members = [t.id] + [types[c] for c in t.children]
        blob += [len(members)] + members
    used = [r for r in res.rules if r.used]
    blob.append(len(used))
    kinds = {"Mirrored": 0, "Left": 1, "Right": 2}
    for r in used:
        blob.append(kinds[r.ruletype])
        if r.precondition is not None:
            pre = compile_expr(r.precondition, ctx)
            blob += [1, len(pre)] + pre
        else:
            blob += [0, 0]
        levels, final = _levels(r)
        blob.append(len(levels))
        for p, cond, act in levels:
            # rules.rs:55-62 pastes "rand.y <= p && " in front of the condition text
            # without parentheses, so GLSL precedence decides what it guards.
            prefix = "" if p == 1.0 else f"rand.y <= {L.fmt_f32(p)} && "
            code = compile_expr(prefix + cond, ctx)
            acts = compile_actions(act, ctx)
            blob += [0, fbits(p), len(code)] + code + [len(acts)]
            for a in acts:
                blob += a
        if final is not None:
            acts = compile_actions(final, ctx)
            blob += [1, len(acts)]
            for a in acts:
                blob += a
        else:
            blob += [0, 0]
    return np.asarray(blob, dtype=np.int32)


# ---------------------------------------------------------------- native library
_lib = None


def build(force: bool = False) -> str:
    src = os.path.join(HERE, "step_ref.c")
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        subprocess.check_call(["make", "-s", "-C", HERE, "libse_oracle.so"])
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        _lib.se_ref_step.argtypes = [P(ctypes.c_int32), ctypes.c_int, P(ctypes.c_uint32),
                                     P(ctypes.c_uint32), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     P(ctypes.c_int32), ctypes.c_int, ctypes.c_int]
        _lib.se_ref_run.argtypes = [P(ctypes.c_int32), ctypes.c_int, P(ctypes.c_uint32),
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    P(ctypes.c_int32), ctypes.c_int]
        _lib.se_ref_hash43.argtypes = [ctypes.c_uint32] * 3 + [P(ctypes.c_uint32), P(ctypes.c_float)]
        _lib.se_ref_validate.argtypes = [P(ctypes.c_int32), ctypes.c_int]
        _lib.se_ref_block_table.argtypes = [P(ctypes.c_int32), ctypes.c_int, ctypes.c_int, P(ctypes.c_uint32),
                                            ctypes.c_int, P(ctypes.c_uint8)]
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def mods_array(mods) -> np.ndarray:
    """mods: iterable of (x, y, shape, size, matID) -> int32[n, 8] (SimModification layout)."""
    out = np.zeros((len(mods), 8), dtype=np.int32)
    for i, m in enumerate(mods):
        out[i, :5] = m[:5]
    return out


def hash43(x: int, y: int, f: int):
    u = np.zeros(4, np.uint32)
    r = np.zeros(4, np.float32)
    lib().se_ref_hash43(x & 0xFFFFFFFF, y & 0xFFFFFFFF, f & 0xFFFFFFFF, _p(u, ctypes.c_uint32),
                        _p(r, ctypes.c_float))
    return u, r


def block_table(blob: np.ndarray, nid: int, hashes: np.ndarray) -> np.ndarray:
    """simulate() on every block of ids in [0, nid) under each row of `hashes`
    (hash43 words, shape (nh, 4)): uint8[nid^4, nh, 4], tuple index
    c0 + nid*c1 + nid^2*c2 + nid^3*c3."""
    h = np.ascontiguousarray(hashes, dtype=np.uint32)
    out = np.zeros((nid ** 4, len(h), 4), np.uint8)
    if lib().se_ref_block_table(_p(blob, ctypes.c_int32), len(blob), nid, _p(h, ctypes.c_uint32), len(h),
                                _p(out, ctypes.c_uint8)):
        raise OracleError("se_ref_block_table failed")
    return out


def step(blob: np.ndarray, grid: np.ndarray, frame: int, mods=None, percell=False) -> np.ndarray:
    g = np.ascontiguousarray(grid, dtype=np.uint32)
    H, W = g.shape
    out = np.empty_like(g)
    m = mods_array(mods) if mods else np.zeros((1, 8), np.int32)
    rc = lib().se_ref_step(_p(blob, ctypes.c_int32), len(blob), _p(g, ctypes.c_uint32),
                           _p(out, ctypes.c_uint32), W, H, frame, _p(m, ctypes.c_int32),
                           len(mods) if mods else 0, 1 if percell else 0)
    if rc:
        raise OracleError("se_ref_step failed (bad ruleset blob?)")
    return out


def run(blob: np.ndarray, grid: np.ndarray, first_frame: int, nsteps: int, mods=None) -> np.ndarray:
    g = np.array(grid, dtype=np.uint32, order="C", copy=True)
    H, W = g.shape
    m = mods_array(mods) if mods else np.zeros((1, 8), np.int32)
    rc = lib().se_ref_run(_p(blob, ctypes.c_int32), len(blob), _p(g, ctypes.c_uint32), W, H,
                          first_frame, nsteps, _p(m, ctypes.c_int32), len(mods) if mods else 0)
    if rc:
        raise OracleError("se_ref_run failed")
    return g


# ---------------------------------------------------------------- seeded fills
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(z: np.ndarray) -> np.ndarray:
    z = (z + np.uint64(0x9E3779B97F4A7C15)) & _M64
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
    return z ^ (z >> np.uint64(31))


def cell_rand(seed: int, idx: np.ndarray) -> np.ndarray:
    """u32 draw for cell index idx: high word of splitmix64((seed << 40) ^ idx)."""
    key = (np.uint64(seed) << np.uint64(40)) ^ idx.astype(np.uint64)
    return (splitmix64(key) >> np.uint64(32)).astype(np.uint32)


def thresholds(percents: Sequence[int]) -> List[int]:
    """Cumulative integer percents -> u32 thresholds floor(cum * 2^32 / 100)."""
    out, cum = [], 0
    for p in percents:
        cum += p
        out.append(min((cum << 32) // 100, 1 << 32))
    return out


def fill(seed: int, W: int, H: int, ids: Sequence[int], percents: Sequence[int],
         stone_rows: int = 0, stone_id: int = 0) -> np.ndarray:
    """i.i.d. categorical fill: cell (x,y) draws r = cell_rand(seed, y*W+x) and takes
    ids[k] for the first k with r < thresholds(percents)[k]; the bottom
    `stone_rows` rows are stone_id."""
    idx = np.arange(W * H, dtype=np.uint64)
    r = cell_rand(seed, idx).astype(np.uint64)
    th = thresholds(percents)
    out = np.full(W * H, ids[-1], dtype=np.uint32)
    for k in range(len(ids) - 1, -1, -1):
        out[r < np.uint64(th[k])] = ids[k]
    g = out.reshape(H, W)
    if stone_rows:
        g[H - stone_rows:, :] = stone_id
    return g
