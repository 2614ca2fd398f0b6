"""The Python mirror's ctypes structures have the C header's layout.

A small C program built with gcc against include/sandengine_hip.h prints the
size of every public struct and the offset of every field; the ctypes
structures in sandengine_amd must agree field for field.
"""
import ctypes
import os
import subprocess
import tempfile

import pytest

import sandengine_amd as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

STRUCTS = {
    "se_modification": S.SeModification,
    "se_serve_config": S.ServeConfig,
    "se_serve_stats": S.ServeStats,
    "se_serve_request": S.ServeRequest,
    "se_grid_step": S.SeGridStep,
}


def c_layout():
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "sandengine_hip.h"', "int main(void) {"]
    for cname, ct in STRUCTS.items():
        lines.append(f'printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for f, _ in ct._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["return 0;", "}"]
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "layout.c"), os.path.join(d, "layout")
        open(src, "w").write("\n".join(lines))
        subprocess.check_call(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), "-o", exe, src])
        out = subprocess.check_output([exe], text=True)
    res = {}
    for ln in out.splitlines():
        s, f, v = ln.split()
        res[(s, f)] = int(v)
    return res


@pytest.mark.parametrize("cname", sorted(STRUCTS))
def test_ctypes_layout_matches_header(cname):
    lay = c_layout()
    ct = STRUCTS[cname]
    assert ctypes.sizeof(ct) == lay[(cname, "sizeof")], cname
    for f, _ in ct._fields_:
        assert getattr(ct, f).offset == lay[(cname, f)], (cname, f)
