"""Loads the committed golden fixtures (data written by make_golden.py)."""
from __future__ import annotations

import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REFERENCE_YAML = os.path.join(os.environ.get("SE_REFERENCE", "/root/reference"), "data",
                              "materials.yaml")


def cases():
    with open(os.path.join(HERE, "cases.json")) as f:
        return json.load(f)["cases"]


_grids = None


def grid(name: str) -> np.ndarray:
    global _grids
    if _grids is None:
        _grids = dict(np.load(os.path.join(HERE, "grids.npz"), allow_pickle=False))
    return _grids[name]


def glsl_sha256():
    with open(os.path.join(HERE, "glsl_sha256.json")) as f:
        return json.load(f)


def ruleset_text(name: str):
    """YAML text of a case's rule set; None when it is the reference's own file and
    the reference tree is absent."""
    from tests.golden.cases import RULESETS
    if name == "reference_default":
        return open(REFERENCE_YAML).read() if os.path.exists(REFERENCE_YAML) else None
    return open(os.path.join(ROOT, RULESETS[name])).read()


def mods_at(case):
    return {int(k): v for k, v in case.get("mods_at", {}).items()}
