"""Writes the golden fixtures under tests/golden/ by running the reference's own
GLSL compute program under Mesa llvmpipe (oracle/_ref/libse_llvmpipe.so).

Run in a container that has the reference tree (the llvmpipe library is built
from it):  python tests/golden/make_golden.py

Outputs (data only: inputs are regenerated from the recorded parameters, and
the input SHA-256 is recorded beside them):
  cases.json   one entry per case: rule set, size, fill, frames, modifications
  grids.npz    expected grids after the run, keyed by case name
  glsl_sha256.json  SHA-256 of the reference's checked-in generated GLSL for
               its data/materials.yaml (pins the GLSL emitters without the tree)
Every grid here was produced by llvmpipe; the oracle is compared against them
in tests/test_oracle.py, the GPU kernels in tests/test_gpu_parity.py.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import lang_ref as L  # noqa: E402
from oracle import llvmpipe as LP  # noqa: E402
from oracle import oracle as O  # noqa: E402
from tests.golden.cases import CASES, RULESETS, make_input  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
REF = os.environ.get("SE_REFERENCE", "/root/reference")


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.uint32).tobytes()).hexdigest()


def main():
    print(LP.info(), flush=True)
    programs = {}
    for name, path in RULESETS.items():
        text = open(os.path.join(ROOT, path)).read() if name != "reference_default" else \
            open(os.path.join(REF, "data", "materials.yaml")).read()
        res = L.parse_string(text)
        # the reference's own assembled program for its own rule file
        programs[name] = LP.program_for(None if name == "reference_default" else res)
    grids, meta = {}, []
    for case in CASES:
        g0 = make_input(case)
        mods_at = {int(k): v for k, v in case.get("mods_at", {}).items()}
        out = LP.seeded_run(programs[case["ruleset"]], g0, case["steps"], mods_at,
                            first_frame=case.get("first_frame", 2),
                            checkpoints=tuple(case.get("checkpoints", ())))
        if case.get("checkpoints"):
            final, snaps = out
            for k, v in snaps.items():
                grids[f"{case['name']}@{k}"] = v
        else:
            final = out
        grids[case["name"]] = final
        meta.append(dict(case, input_sha256=sha(g0), output_sha256=sha(final)))
        print(f"{case['name']}: {g0.shape} {case['steps']} steps -> {sha(final)[:16]}", flush=True)
    np.savez_compressed(os.path.join(OUT, "grids.npz"), **grids)
    with open(os.path.join(OUT, "cases.json"), "w") as f:
        json.dump({"llvmpipe": LP.info(), "cases": meta}, f, indent=1)
    g = {k: hashlib.sha256(LP.ref_gen(i).encode()).hexdigest()
         for i, k in enumerate(["materials.glsl", "rules.glsl", "falling_sand.glsl"])}
    with open(os.path.join(OUT, "glsl_sha256.json"), "w") as f:
        json.dump(g, f, indent=1)


if __name__ == "__main__":
    main()
