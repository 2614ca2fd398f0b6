"""Oracle restatement of sandengine-lang: YAML DSL -> IR -> GLSL text.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``sandengine_amd``) may
import this module; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker.

It restates, step by step, the reference parser and GLSL emitter:

* ``parse_string``            sandengine-lang/src/parser.rs:93-152
* ``_preparse_keys``          sandengine-lang/src/parser.rs:156-169
* ``_check_mapping``          sandengine-lang/src/parser.rs:174-187
* ``_extract_vec4``           sandengine-lang/src/parser.rs:192-250
* ``_parse_rules``            sandengine-lang/src/parser/rules.rs:105-203
* ``_parse_conditionals``     sandengine-lang/src/parser/rules.rs:207-334
* ``_global_scope``           sandengine-lang/src/parser/rules.rs:338-351
* ``_parse_do``               sandengine-lang/src/parser/rules.rs:355-420
* ``_parse_types``            sandengine-lang/src/parser/types.rs:51-182
* ``_add_child_to_type``      sandengine-lang/src/parser/types.rs:186-200
* ``_get_parents_rules``      sandengine-lang/src/parser/types.rs:202-209
* ``_parse_materials``        sandengine-lang/src/parser/materials.rs:51-201
* ``rule_glsl``/``_func_logic`` sandengine-lang/src/parser/rules.rs:47-97
* ``create_glsl``             sandengine-lang/src/lib.rs:17-147

YAML is loaded with PyYAML's safe loader (mapping order preserved), standing
in for serde_yaml 0.9.34; the two agree on every construct the rule files use
(block/flow mappings and sequences, plain/quoted scalars, ints, floats, bools).

Parity pins: ``create_glsl`` on the reference's data/materials.yaml is
compared byte for byte with the reference's checked-in
shaders/compute/gen/{materials,rules}.glsl (their SHA-256 is a committed
fixture), and the error categories are checked against the reference's
tests/test_sandengine-lang.rs inputs.
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field
from typing import Any, List, Optional

import numpy as np
import yaml

# parser.rs:27-34
GLOBAL_CELLNAMES = ["SELF", "LEFT", "RIGHT", "DOWN", "DOWNRIGHT", "DOWNLEFT"]
# parser.rs:16-21
TYPE_HINT_STRING = "string"
TYPE_HINT_BOOL = "bool (true/false)"
TYPE_HINT_FLOAT = "float (0.0 to 1.0)"
TYPE_HINT_SEQUENCE = "sequence (array, '[...]')"
TYPE_HINT_COLOR = ("sequence (array, '[...]') of 3-4 floats (range 0.0-1.0) OR integers "
                   "(range 0-255). (With 3 elements, the alpha channel defaults to 1.0)")
TYPE_HINT_MAPPING = "mapping (dictionary-like)"


class ParsingErr(Exception):
    """One of MissingField / InvalidType / NotFound / NotRecognized (parser.rs:44-73)."""

    def __init__(self, kind: str, message: str):
        super().__init__(message)
        self.kind = kind
        self.message = message


def _dbg(v: Any) -> str:
    # Rust `{:?}` of &str / String prints the quoted string.
    if isinstance(v, _Val):
        return repr(v)
    if isinstance(v, str):
        return '"' + v.replace("\\", "\\\\").replace('"', '\\"') + '"'
    return repr(v)


class _Val:
    """A serde_yaml Value passed as `wrong_type` (Debug-printed as Number(..) etc.)."""

    def __init__(self, v):
        self.v = v

    def __repr__(self):
        v = self.v
        if v is None:
            return "Null"
        if isinstance(v, bool):
            return f"Bool({'true' if v else 'false'})"
        if isinstance(v, (int, float)):
            return f"Number({v})"
        if isinstance(v, str):
            return f'String("{v}")'
        return "Sequence" if isinstance(v, list) else "Mapping"


def _missing(field_name, missing_in):
    return ParsingErr("MissingField",
                      f"(MissingField) Mandatory field '{field_name}' is missing in '{missing_in}'")


def _invalid(wrong, missing_in, expected):
    return ParsingErr("InvalidType",
                      f"(InvalidType) The type of the field '{_dbg(wrong)}' inside of '{missing_in}' "
                      f"is invalid. Expected: '{expected}'")


def _notfound(missing, missing_in):
    return ParsingErr("NotFound",
                      f"(NotFound) The name '{missing}' (in '{missing_in}') was not found. "
                      f"Make sure it was defined before referencing it.")


def _notrecog(unrecog, missing_in):
    return ParsingErr("NotRecognized",
                      f"(NotRecognized) The expression '{unrecog}' (in '{missing_in}') was not "
                      f"recognized as valid syntax. Please check it is valid.")


def f32(x) -> float:
    return float(np.float32(x))


def fmt_f32(x: float) -> str:
    """Rust `{}` Display of an f32: shortest round-trip digits, positional."""
    return np.format_float_positional(np.float32(x), unique=True, trim="-")


# serde_yaml Value accessors --------------------------------------------------------

def _get(v, key):
    # serde_yaml Value::get(&str) -> None unless v is a mapping holding key
    if isinstance(v, dict):
        return v.get(key)
    return None


def _has(v, key):
    return isinstance(v, dict) and key in v


def _as_str(v):
    return v if isinstance(v, str) else None


def _as_bool(v):
    return v if isinstance(v, bool) else None


def _as_f64(v):
    if isinstance(v, bool):
        return None
    if isinstance(v, (int, float)):
        return float(v)
    return None


def _as_u64(v):
    if isinstance(v, bool):
        return None
    if isinstance(v, int) and v >= 0:
        return v
    return None


def _as_seq(v):
    return v if isinstance(v, list) else None


# IR ---------------------------------------------------------------------------------

@dataclass
class SandRule:
    name: str
    ruletype: str                       # "Mirrored" | "Left" | "Right"
    if_conds: List[str]
    do_actions: List[str]
    probabilities: List[float]          # f32 values
    mirror: bool
    precondition: Optional[str]
    used: bool = False


@dataclass
class SandType:
    id: int
    name: str
    inherits: str = ""
    children: List[str] = field(default_factory=list)
    base_rules: List[str] = field(default_factory=list)


@dataclass
class SandMaterial:
    id: int
    name: str
    mattype: str
    color: List[float]
    emission: List[float]
    selectable: bool
    density: float
    extra_rules: List[str] = field(default_factory=list)


@dataclass
class ParsingResult:
    rules: List[SandRule]
    types: List[SandType]
    materials: List[SandMaterial]

    def to_ir(self) -> dict:
        """Plain-data IR, the shape the product's IR dump uses."""
        return {
            "rules": [dict(name=r.name, ruletype=r.ruletype, if_conds=list(r.if_conds),
                           do_actions=list(r.do_actions),
                           probabilities=[f32(p) for p in r.probabilities], mirror=r.mirror,
                           precondition=r.precondition, used=r.used) for r in self.rules],
            "types": [dict(id=t.id, name=t.name, inherits=t.inherits, children=list(t.children),
                           base_rules=list(t.base_rules)) for t in self.types],
            "materials": [dict(id=m.id, name=m.name, mattype=m.mattype,
                               color=[f32(c) for c in m.color],
                               emission=[f32(c) for c in m.emission], selectable=m.selectable,
                               density=f32(m.density), extra_rules=list(m.extra_rules))
                          for m in self.materials],
        }


# parser.rs -------------------------------------------------------------------------

def parse_string(text: str) -> ParsingResult:
    try:
        data = yaml.safe_load(text)
    except yaml.YAMLError as e:  # serde_yaml error -> bail!(err)
        raise ParsingErr("Yaml", str(e))
    raw_rules = _check_mapping(data, "rules")
    raw_types = _check_mapping(data, "types")
    raw_materials = _check_mapping(data, "materials")
    rule_names = _preparse_keys(raw_rules, "rules")
    type_names = _preparse_keys(raw_types, "types") + ["EMPTY"]
    material_names = _preparse_keys(raw_materials, "materials") + ["EMPTY"]
    try:
        rules = _parse_rules(raw_rules, type_names, material_names)
    except ParsingErr as e:
        raise ParsingErr(e.kind, f"Error while parsing rules: '{e.message}'")
    try:
        types = _parse_types(raw_types, rules, rule_names, type_names)
    except ParsingErr as e:
        raise ParsingErr(e.kind, f"Error while parsing types: '{e.message}'")
    try:
        materials = _parse_materials(raw_materials, rules, type_names)
    except ParsingErr as e:
        raise ParsingErr(e.kind, f"Error while parsing materials: '{e.message}'")
    return ParsingResult(rules, types, materials)


def _preparse_keys(m: dict, err_name: str) -> List[str]:
    out = []
    for k in m:
        if not isinstance(k, str):
            raise _invalid(_Val(k), err_name, TYPE_HINT_STRING)
        out.append(k)
    return out


def _check_mapping(d, keyname):
    if not _has(d, keyname):
        raise _missing(keyname, "Root/ Base level of YAML file")
    v = d[keyname]
    if not isinstance(v, dict):
        raise _invalid(_Val(v), "Root/ Base level of YAML file", TYPE_HINT_MAPPING)
    return v


def _extract_vec4(yaml_data, parent_name, field_name, default, mandatory):
    missing_in = f"materials/{parent_name}/{field_name}"
    if not _has(yaml_data, field_name):
        if mandatory:
            raise _missing(field_name, missing_in)
        return list(default)
    val = yaml_data[field_name]
    vec4 = [f32(x) for x in default]
    comps = _as_seq(val)
    if comps is None:
        raise _invalid(field_name, missing_in, TYPE_HINT_COLOR)
    if len(comps) == 0 or len(comps) > 4 or len(comps) < 3:
        raise _invalid(field_name, missing_in, TYPE_HINT_COLOR)
    for idx, comp in enumerate(comps):
        u = _as_u64(comp)
        if u is not None and 0 < u <= 255:
            vec4[idx] = float(np.float32(u) / np.float32(255.0))
            continue
        f = _as_f64(comp)
        if f is not None:
            if 1.0 < f <= 255.0:
                vec4[idx] = float(np.float32(f) / np.float32(255.0))
                continue
            if 0.0 <= f <= 1.0:
                vec4[idx] = f32(f)
                continue
        raise _invalid(field_name, missing_in, TYPE_HINT_COLOR)
    return vec4


# rules.rs --------------------------------------------------------------------------

def _parse_rules(rules: dict, type_names, material_names) -> List[SandRule]:
    out = []
    for name, body in rules.items():
        if not isinstance(name, str):
            raise _invalid(_Val(name), "rules", TYPE_HINT_STRING)
        if_conds, do_actions, probs = [], [], []
        _parse_conditionals(body, False, f"rules/{name}", if_conds, do_actions, probs,
                            type_names, material_names)
        if _has(body, "mirrored"):
            m = _as_bool(body["mirrored"])
            if m is None:
                raise _invalid("mirrored", f"rules/{name}", TYPE_HINT_BOOL)
        else:
            m = True
        if m:
            ruletype = "Mirrored"
        else:
            ruletype = "Left" if "LEFT" in do_actions[0] else "Right"
        if _has(body, "precondition"):
            pre = _as_bool(body["precondition"])
            if pre is None:
                raise _invalid("precondition", f"rules/{name}", TYPE_HINT_BOOL)
        else:
            pre = True
        out.append(SandRule(name, ruletype, if_conds, do_actions, probs, m,
                            "" if pre else None, False))
    return out


_MAT_RE = re.compile(r"\w*.mat\s*(==|!=)\s*(\w*)")
_TYPE_RE = re.compile(r"isType_(\w*)\(\w*\)")
_SWAP_RE = re.compile(r"SWAP (\w+) (\w+)")
_SET_RE = re.compile(r"SET (\w+) (\w+)")


def _parse_conditionals(parent, parent_is_else, parent_path, if_conds, do_actions, probs,
                        type_names, material_names):
    if_cond = _get(parent, "if")
    if not _has(parent, "if") and not parent_is_else:
        raise _missing("if", parent_path)
    elif _has(parent, "if"):
        ipath = f"{parent_path}/if"
        s = _as_str(if_cond)
        if s is None:
            raise _invalid("if", ipath, TYPE_HINT_STRING)
        s = _global_scope(s)
        original = s
        for mm in _MAT_RE.finditer(original):
            cap = mm.group(2)
            if cap in material_names:
                s = s.replace(cap, f"MAT_{cap}")
                continue
            raise _notfound(cap, ipath)
        for mm in _TYPE_RE.finditer(s):
            cap = mm.group(1)
            if cap not in type_names:
                raise _notfound(cap, f"{ipath} -> isType_")
        if_conds.append(s)
    if not _has(parent, "do"):
        raise _missing("do", parent_path)
    do = parent["do"]
    do_path = f"{parent_path}/do"
    do_string = ""
    if isinstance(do, str):
        do_string = _global_scope(_parse_do(do_path, do))
    if isinstance(do, list):
        for a in do:
            if isinstance(a, str):
                do_string += _global_scope(_parse_do(do_path, a))
    do_actions.append(do_string.rstrip())
    if _has(parent, "probability"):
        p = _as_f64(parent["probability"])
        if p is None:
            raise _invalid("probability", parent_path, TYPE_HINT_FLOAT)
        probs.append(f32(p))
    else:
        probs.append(1.0)
    if _has(parent, "else"):
        _parse_conditionals(parent["else"], True, f"{parent_path}/else", if_conds, do_actions,
                            probs, type_names, material_names)


def _global_scope(s: str) -> str:
    s = s.replace(" or ", " || ")
    s = s.replace(" and ", " && ")
    s = s.replace("not ", " !")
    s = s.replace("empty", "MAT_EMPTY")
    s = s.replace("SELF", "self")
    s = s.replace("RIGHT", "right")
    s = s.replace("LEFT", "left")
    s = s.replace("DOWN", "down")
    s = s.replace("DOWNRIGHT", "downright")
    s = s.replace("DOWNLEFT", "downleft")
    return s


def _parse_do(parent: str, do_str: str) -> str:
    out = ""
    found = False
    m = _SWAP_RE.search(do_str)
    if m:
        found = True
        a, b = m.group(1), m.group(2)
        if a not in GLOBAL_CELLNAMES:
            raise _notfound(a, parent)
        if b not in GLOBAL_CELLNAMES:
            raise _notfound(b, parent)
        out += f"swap({a}, {b});\n"
    m = _SET_RE.search(do_str)
    if m:
        found = True
        a, b = m.group(1), m.group(2)
        if a not in GLOBAL_CELLNAMES:
            raise _notfound(a, parent)
        out += f"{a} = newCell(MAT_{b}, pos);\n"
    if not found:
        raise _notrecog(do_str, parent)
    return out


# types.rs --------------------------------------------------------------------------

def _update_pre(rule: SandRule, clause: str):
    if rule.precondition is not None:
        rule.precondition = clause if rule.precondition == "" else f"{rule.precondition} || {clause}"


def _parse_types(types: dict, rules: List[SandRule], rule_names, type_names) -> List[SandType]:
    ts = [SandType(0, "EMPTY"), SandType(1, "NULL"), SandType(2, "WALL")]
    idx = len(ts)
    for name, body in types.items():
        if not isinstance(name, str):
            raise _invalid(_Val(name), "types", TYPE_HINT_STRING)
        parent = ""
        if _has(body, "inherits"):
            p = _as_str(body["inherits"])
            if p is None:
                raise _invalid("inherits", f"types/{name}", TYPE_HINT_STRING)
            parent = p
            for tn in type_names:
                if tn == parent:
                    _add_child_to_type(parent, name, ts)
        base_rules = []
        if _has(body, "base_rules"):
            br = _as_seq(body["base_rules"])
            if br is None:
                raise _invalid("base_rules", f"types/{name}", TYPE_HINT_SEQUENCE)
            for b in br:
                rn = _as_str(b)
                if rn is None:
                    raise _invalid("base_rules", f"types/{name}", TYPE_HINT_STRING)
                if rn not in rule_names:
                    raise _notfound(rn, f"types/{name}/base_rules")
                rule = next(r for r in rules if r.name == rn)
                rule.used = True
                _update_pre(rule, f"isType_{name}(self)")
                base_rules.append(rn)
        ts.append(SandType(idx, name, parent, [], base_rules))
        idx += 1
    for st in [SandType(t.id, t.name, t.inherits, list(t.children), list(t.base_rules)) for t in ts]:
        if st.inherits == "":
            continue
        for rn in _get_parents_rules(ts, st):
            rule = next(r for r in rules if r.name == rn)
            _update_pre(rule, f"isType_{st.name}(self)")
    return ts


def _add_child_to_type(parent_name, childname, ts):
    pp = ""
    for t in ts:
        if t.name == parent_name:
            t.children.append(childname)
            pp = t.inherits
            break
    if pp != "":
        _add_child_to_type(pp, childname, ts)


def _get_parents_rules(all_types, cur):
    if cur.inherits == "":
        return []
    parent = next((t for t in all_types if t.name == cur.inherits), None)
    if parent is None:  # the reference unwraps here and panics
        raise ParsingErr("Panic", f"parent type '{cur.inherits}' of '{cur.name}' not found")
    return list(parent.base_rules) + _get_parents_rules(all_types, parent)


# materials.rs ----------------------------------------------------------------------

def _parse_materials(mats: dict, rules: List[SandRule], type_names) -> List[SandMaterial]:
    ms = [
        SandMaterial(0, "EMPTY", "EMPTY", [0.0, 0.0, 0.0, 0.0], [0.0] * 4, True, 1.0),
        SandMaterial(1, "NULL", "NULL", [1.0, 0.0, 1.0, 1.0], [0.0] * 4, False, 0.0),
        SandMaterial(2, "WALL", "WALL", [f32(0.1), f32(0.2), f32(0.3), 1.0], [0.0] * 4, False,
                     9999.0),
    ]
    idx = len(ms)
    for name, body in mats.items():
        if not isinstance(name, str):
            raise _invalid(_Val(name), "materials", TYPE_HINT_STRING)
        if not _has(body, "type"):
            raise _missing("type", f"materials/{name}")
        mt = _as_str(body["type"])
        if mt is None:
            raise _invalid("type", f"materials/{name}", TYPE_HINT_STRING)
        if mt not in type_names:
            raise _notfound(mt, f"materials/{name}/type")
        color = _extract_vec4(body, name, "color", [1.0, 0.0, 1.0, 1.0], True)
        emission = _extract_vec4(body, name, "emission", [0.0, 0.0, 0.0, 0.0], False)
        if _has(body, "selectable"):
            sel = _as_bool(body["selectable"])
            sel = bool(sel) if sel is not None else False
        else:
            sel = True
        if not _has(body, "density"):
            raise _missing("density", f"materials/{name}")
        d = _as_f64(body["density"])
        if d is None:
            raise _invalid("density", f"materials/{name}/density", TYPE_HINT_FLOAT)
        extra = []
        if _has(body, "extra_rules"):
            er = _as_seq(body["extra_rules"])
            if er is None:
                raise _invalid("extra_rules", f"materials/{name}", TYPE_HINT_SEQUENCE)
            for e in er:
                en = _as_str(e)
                if en is None:
                    raise _invalid("extra_rules", f"materials/{name}", TYPE_HINT_STRING)
                for r in rules:
                    if r.name == en:
                        r.used = True
                        _update_pre(r, f"self.mat == MAT_{name}")
                        extra.append(en)
        ms.append(SandMaterial(idx, name, mt, color, emission, sel, f32(d), extra))
        idx += 1
    return ms


# GLSL emission ---------------------------------------------------------------------

def _func_logic(if_conds, do_actions, probs, indent) -> str:
    if not if_conds and not do_actions:
        return ""
    if not if_conds and do_actions:
        return do_actions[0]
    if if_conds and do_actions:
        ind1, ind2 = " " * (indent * 4), " " * ((indent + 1) * 4)
        p = probs[0]
        prob = "" if p == 1.0 else f"rand.y <= {fmt_f32(p)} && "
        rest = _func_logic(if_conds[1:], do_actions[1:], probs[1:], indent + 1)
        return (f"{ind1}if ({prob}{if_conds[0]}) {{\n{ind2}{do_actions[0]}\n"
                f"{ind1}}} else {{\n{rest}\n{ind1}}}")
    return ""


def rule_glsl(r: SandRule) -> str:
    celldir = "left" if r.ruletype == "Left" else "right"
    pre = "" if r.precondition is None else f"    if (!({r.precondition})) {{\n        return;\n    }}\n"
    body = _func_logic(list(r.if_conds), list(r.do_actions), list(r.probabilities), 1)
    return (f"void rule_{r.name} (inout Cell self, inout Cell {celldir}, inout Cell down, "
            f"inout Cell downright, vec4 rand, ivec2 pos) {{\n{pre}{body}\n}}")


def create_glsl(res: ParsingResult):
    """Returns (materials_glsl, rules_glsl) exactly as sandengine-lang/src/lib.rs:17-147 writes."""
    s = ""
    for t in res.types:
        s += f"#define TYPE_{t.name} {t.id}\n\n"
    for t in res.types:
        tc = f"return cell.mat.type == TYPE_{t.name}"
        for c in t.children:
            tc += f" || cell.mat.type == TYPE_{c}"
        s += f"bool isType_{t.name}(Cell cell) {{\n    {tc};\n}}\n\n"
    s += "\n"
    lst = ""
    n = len(res.materials)
    for i, m in enumerate(res.materials):
        c, e = m.color, m.emission
        s += (f"#define MAT_{m.name} Material({m.id}, vec4({fmt_f32(c[0])}, {fmt_f32(c[1])}, "
              f"{fmt_f32(c[2])}, {fmt_f32(c[3])}), {fmt_f32(m.density)}, vec4({fmt_f32(e[0])}, "
              f"{fmt_f32(e[1])}, {fmt_f32(e[2])}, {fmt_f32(e[3])}), TYPE_{m.mattype})\n")
        lst += f"        MAT_{m.name}" + ("" if i == n - 1 else ",\n")
    s += (f"\nMaterial[{n}] materials() {{\n    Material allMaterials[{n}] = {{\n{lst}\n    }};\n"
          f"    return allMaterials;\n}}\n\nMaterial getMaterialFromID(int id) {{\n"
          f"    for (int i = 0; i < materials().length(); i++) {{\n"
          f"        if (id == materials()[i].id) {{\n            return materials()[i];\n"
          f"        }};\n    }};\n    return MAT_NULL;\n}}\n\n")
    funcs, mir, left, right = "", "", "", ""
    for r in res.rules:
        if not r.used:
            continue
        funcs += rule_glsl(r) + "\n\n"
        if r.ruletype == "Mirrored":
            mir += f"rule_{r.name}(self, right, down, downright, rand, pos);\n"
        elif r.ruletype == "Left":
            left += f"rule_{r.name}(self, left, down, downright, rand, pos);\n"
        else:
            right += f"rule_{r.name}(self, right, down, downright, rand, pos);\n"
    hdr = ("    inout Cell self,\n    inout Cell right,\n    inout Cell down,\n"
           "    inout Cell downright,\n    vec4 rand,\n    ivec2 pos) {\n")
    rules = (f"\n// =============== RULES ===============\n{funcs}\n\n\n"
             f"// =============== CALLERS ===============\nvoid applyMirroredRules(\n{hdr}"
             f"    {mir.rstrip()}\n}}\n\n\nvoid applyLeftRules(\n{hdr}    {left.rstrip()}\n}}\n\n"
             f"void applyRightRules(\n{hdr}    {right.rstrip()}\n}}")
    return s, rules
