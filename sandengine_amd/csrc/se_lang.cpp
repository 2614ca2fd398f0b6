#include "se_lang.hpp"

#include <charconv>
#include <cmath>
#include <functional>
#include <regex>
#include <sstream>

#include "se_yaml.hpp"

namespace se {

namespace {

// parser.rs:27-34
const char* kCellNames[] = {"SELF", "LEFT", "RIGHT", "DOWN", "DOWNRIGHT", "DOWNLEFT"};
// parser.rs:16-21
const char* kHintString = "string";
const char* kHintBool = "bool (true/false)";
const char* kHintFloat = "float (0.0 to 1.0)";
const char* kHintSeq = "sequence (array, '[...]')";
const char* kHintColor =
    "sequence (array, '[...]') of 3-4 floats (range 0.0-1.0) OR integers (range 0-255). (With 3 "
    "elements, the alpha channel defaults to 1.0)";
const char* kHintMapping = "mapping (dictionary-like)";

std::string quoted(const std::string& s) { return "\"" + s + "\""; }

LangError missing(const std::string& field, const std::string& in) {
    return LangError("MissingField", "(MissingField) Mandatory field '" + field + "' is missing in '" + in + "'");
}
LangError invalid(const std::string& wrong_dbg, const std::string& in, const char* expected) {
    return LangError("InvalidType", "(InvalidType) The type of the field '" + wrong_dbg + "' inside of '" + in +
                                        "' is invalid. Expected: '" + expected + "'");
}
LangError notfound(const std::string& m, const std::string& in) {
    return LangError("NotFound", "(NotFound) The name '" + m + "' (in '" + in +
                                     "') was not found. Make sure it was defined before referencing it.");
}
LangError notrecog(const std::string& u, const std::string& in) {
    return LangError("NotRecognized", "(NotRecognized) The expression '" + u + "' (in '" + in +
                                          "') was not recognized as valid syntax. Please check it is valid.");
}

bool contains(const std::vector<std::string>& v, const std::string& s) {
    for (auto& x : v)
        if (x == s) return true;
    return false;
}

std::string replace_all(std::string s, const std::string& from, const std::string& to) {
    if (from.empty()) return s;
    std::string out;
    size_t p = 0;
    for (;;) {
        size_t q = s.find(from, p);
        if (q == std::string::npos) break;
        out.append(s, p, q - p);
        out += to;
        p = q + from.size();
    }
    out.append(s, p, std::string::npos);
    return out;
}

std::string rtrim(const std::string& s) {
    size_t e = s.size();
    while (e > 0 && std::isspace((unsigned char)s[e - 1])) e--;
    return s.substr(0, e);
}

// rules.rs:338-351
std::string global_scope(std::string s) {
    s = replace_all(s, " or ", " || ");
    s = replace_all(s, " and ", " && ");
    s = replace_all(s, "not ", " !");
    s = replace_all(s, "empty", "MAT_EMPTY");
    s = replace_all(s, "SELF", "self");
    s = replace_all(s, "RIGHT", "right");
    s = replace_all(s, "LEFT", "left");
    s = replace_all(s, "DOWN", "down");
    s = replace_all(s, "DOWNRIGHT", "downright");
    s = replace_all(s, "DOWNLEFT", "downleft");
    return s;
}

bool is_cellname(const std::string& s) {
    for (auto* c : kCellNames)
        if (s == c) return true;
    return false;
}

// rules.rs:355-420
std::string parse_do(const std::string& parent, const std::string& d) {
    static const std::regex swap_re("SWAP (\\w+) (\\w+)");
    static const std::regex set_re("SET (\\w+) (\\w+)");
    std::string out;
    bool found = false;
    std::smatch m;
    if (std::regex_search(d, m, swap_re)) {
        found = true;
        std::string a = m[1], b = m[2];
        if (!is_cellname(a)) throw notfound(a, parent);
        if (!is_cellname(b)) throw notfound(b, parent);
        out += "swap(" + a + ", " + b + ");\n";
    }
    if (std::regex_search(d, m, set_re)) {
        found = true;
        std::string a = m[1], b = m[2];
        if (!is_cellname(a)) throw notfound(a, parent);
        out += a + " = newCell(MAT_" + b + ", pos);\n";
    }
    if (!found) throw notrecog(d, parent);
    return out;
}

float to_f32(double d) { return (float)d; }

// rules.rs:207-334
void parse_conditionals(const YNode& parent, bool parent_is_else, const std::string& path, SandRule& r,
                        const std::vector<std::string>& type_names, const std::vector<std::string>& mat_names) {
    const YNode* ifc = parent.get("if");
    if (!ifc && !parent_is_else) throw missing("if", path);
    if (ifc) {
        std::string ipath = path + "/if";
        std::string s;
        if (!ifc->as_str(&s)) throw invalid(quoted("if"), ipath, kHintString);
        s = global_scope(s);
        static const std::regex mat_re("\\w*.mat\\s*(==|!=)\\s*(\\w*)");
        std::string original = s;
        for (std::sregex_iterator it(original.begin(), original.end(), mat_re), end; it != end; ++it) {
            std::string cap = (*it)[2];
            if (contains(mat_names, cap)) {
                s = replace_all(s, cap, "MAT_" + cap);
                continue;
            }
            throw notfound(cap, ipath);
        }
        static const std::regex type_re("isType_(\\w*)\\(\\w*\\)");
        std::string cur = s;
        for (std::sregex_iterator it(cur.begin(), cur.end(), type_re), end; it != end; ++it) {
            std::string cap = (*it)[1];
            if (!contains(type_names, cap)) throw notfound(cap, ipath + " -> isType_");
        }
        r.if_conds.push_back(s);
    }
    const YNode* d = parent.get("do");
    if (!d) throw missing("do", path);
    std::string dpath = path + "/do";
    std::string ds, one;
    if (d->as_str(&one)) ds = global_scope(parse_do(dpath, one));
    if (d->kind == YNode::Seq)
        for (auto& a : d->seq)
            if (a.as_str(&one)) ds += global_scope(parse_do(dpath, one));
    r.do_actions.push_back(rtrim(ds));
    const YNode* pr = parent.get("probability");
    if (pr) {
        double p;
        if (!pr->as_f64(&p)) throw invalid(quoted("probability"), path, kHintFloat);
        r.probabilities.push_back(to_f32(p));
    } else {
        r.probabilities.push_back(1.0f);
    }
    const YNode* e = parent.get("else");
    if (e) parse_conditionals(*e, true, path + "/else", r, type_names, mat_names);
}

std::vector<std::string> preparse_keys(const YNode& m, const std::string& err) {
    std::vector<std::string> out;
    for (size_t i = 0; i < m.map.size(); i++) {
        if (!m.key_is_str[i]) throw invalid(m.key_dbg[i], err, kHintString);
        out.push_back(m.map[i].first);
    }
    return out;
}

const YNode& check_mapping(const YNode& d, const std::string& key) {
    const YNode* v = d.get(key);
    if (!v) throw missing(key, "Root/ Base level of YAML file");
    if (v->kind != YNode::Map) throw invalid(v->debug(), "Root/ Base level of YAML file", kHintMapping);
    return *v;
}

std::vector<SandRule> parse_rules(const YNode& rules, const std::vector<std::string>& type_names,
                                  const std::vector<std::string>& mat_names) {
    std::vector<SandRule> out;
    for (size_t i = 0; i < rules.map.size(); i++) {
        if (!rules.key_is_str[i]) throw invalid(rules.key_dbg[i], "rules", kHintString);
        const std::string& name = rules.map[i].first;
        const YNode& body = rules.map[i].second;
        SandRule r;
        r.name = name;
        parse_conditionals(body, false, "rules/" + name, r, type_names, mat_names);
        bool mirrored = true;
        if (const YNode* m = body.get("mirrored"))
            if (!m->as_bool(&mirrored)) throw invalid(quoted("mirrored"), "rules/" + name, kHintBool);
        r.mirror = mirrored;
        if (mirrored) r.ruletype = RuleType::Mirrored;
        else r.ruletype = r.do_actions[0].find("LEFT") != std::string::npos ? RuleType::Left : RuleType::Right;
        bool pre = true;
        if (const YNode* p = body.get("precondition"))
            if (!p->as_bool(&pre)) throw invalid(quoted("precondition"), "rules/" + name, kHintBool);
        if (pre) r.precondition = std::string();
        out.push_back(std::move(r));
    }
    return out;
}

void update_pre(SandRule& r, const std::string& clause) {
    if (!r.precondition) return;
    if (r.precondition->empty()) r.precondition = clause;
    else r.precondition = *r.precondition + " || " + clause;
}

// types.rs:186-200
void add_child_to_type(const std::string& parent, const std::string& child, std::vector<SandType>& ts) {
    std::string pp;
    for (auto& t : ts)
        if (t.name == parent) {
            t.children.push_back(child);
            pp = t.inherits;
            break;
        }
    if (!pp.empty()) add_child_to_type(pp, child, ts);
}

// types.rs:202-209
std::vector<std::string> parents_rules(const std::vector<SandType>& all, const SandType& cur, int depth = 0) {
    if (cur.inherits.empty()) return {};
    if (depth > 1000) throw LangError("NotFound", "type inheritance cycle at '" + cur.name + "'");
    const SandType* p = nullptr;
    for (auto& t : all)
        if (t.name == cur.inherits) { p = &t; break; }
    if (!p) throw notfound(cur.inherits, "types/" + cur.name + "/inherits");
    std::vector<std::string> rs = p->base_rules;
    auto more = parents_rules(all, *p, depth + 1);
    rs.insert(rs.end(), more.begin(), more.end());
    return rs;
}

SandRule& find_rule(std::vector<SandRule>& rules, const std::string& n) {
    for (auto& r : rules)
        if (r.name == n) return r;
    throw notfound(n, "rules");
}

std::vector<SandType> parse_types(const YNode& types, std::vector<SandRule>& rules,
                                  const std::vector<std::string>& rule_names,
                                  const std::vector<std::string>& type_names) {
    std::vector<SandType> ts(3);
    ts[0].id = 0; ts[0].name = "EMPTY";
    ts[1].id = 1; ts[1].name = "NULL";
    ts[2].id = 2; ts[2].name = "WALL";
    size_t idx = ts.size();
    for (size_t i = 0; i < types.map.size(); i++) {
        if (!types.key_is_str[i]) throw invalid(types.key_dbg[i], "types", kHintString);
        const std::string& name = types.map[i].first;
        const YNode& body = types.map[i].second;
        std::string parent;
        if (const YNode* inh = body.get("inherits")) {
            if (!inh->as_str(&parent)) throw invalid(quoted("inherits"), "types/" + name, kHintString);
            for (auto& tn : type_names)
                if (tn == parent) add_child_to_type(parent, name, ts);
        }
        std::vector<std::string> base;
        if (const YNode* br = body.get("base_rules")) {
            if (br->kind != YNode::Seq) throw invalid(quoted("base_rules"), "types/" + name, kHintSeq);
            for (auto& b : br->seq) {
                std::string rn;
                if (!b.as_str(&rn)) throw invalid(quoted("base_rules"), "types/" + name, kHintString);
                if (!contains(rule_names, rn)) throw notfound(rn, "types/" + name + "/base_rules");
                SandRule& r = find_rule(rules, rn);
                r.used = true;
                update_pre(r, "isType_" + name + "(self)");
                base.push_back(rn);
            }
        }
        SandType t;
        t.id = idx++;
        t.name = name;
        t.inherits = parent;
        t.base_rules = base;
        ts.push_back(t);
    }
    std::vector<SandType> snapshot = ts;
    for (auto& t : snapshot) {
        if (t.inherits.empty()) continue;
        for (auto& rn : parents_rules(ts, t)) update_pre(find_rule(rules, rn), "isType_" + t.name + "(self)");
    }
    return ts;
}

// parser.rs:192-250
void extract_vec4(const YNode& mat, const std::string& parent, const char* field, const float def[4],
                  bool mandatory, float out[4]) {
    std::string in = "materials/" + parent + "/" + field;
    for (int i = 0; i < 4; i++) out[i] = def[i];
    const YNode* v = mat.get(field);
    if (!v) {
        if (mandatory) throw missing(field, in);
        return;
    }
    if (v->kind != YNode::Seq) throw invalid(quoted(field), in, kHintColor);
    if (v->seq.empty() || v->seq.size() > 4 || v->seq.size() < 3) throw invalid(quoted(field), in, kHintColor);
    for (size_t i = 0; i < v->seq.size(); i++) {
        const YNode& c = v->seq[i];
        uint64_t u;
        if (c.as_u64(&u) && u > 0 && u <= 255) {
            out[i] = (float)u / 255.0f;
            continue;
        }
        double d;
        if (c.as_f64(&d)) {
            if (d > 1.0 && d <= 255.0) { out[i] = (float)d / 255.0f; continue; }
            if (d >= 0.0 && d <= 1.0) { out[i] = (float)d; continue; }
        }
        throw invalid(quoted(field), in, kHintColor);
    }
}

std::vector<SandMaterial> parse_materials(const YNode& mats, std::vector<SandRule>& rules,
                                          const std::vector<std::string>& type_names) {
    std::vector<SandMaterial> ms(3);
    ms[0].id = 0; ms[0].name = "EMPTY"; ms[0].mattype = "EMPTY"; ms[0].selectable = true; ms[0].density = 1.0f;
    ms[1].id = 1; ms[1].name = "NULL"; ms[1].mattype = "NULL"; ms[1].selectable = false; ms[1].density = 0.0f;
    ms[1].color[0] = 1.0f; ms[1].color[2] = 1.0f; ms[1].color[3] = 1.0f;
    ms[2].id = 2; ms[2].name = "WALL"; ms[2].mattype = "WALL"; ms[2].selectable = false; ms[2].density = 9999.0f;
    ms[2].color[0] = 0.1f; ms[2].color[1] = 0.2f; ms[2].color[2] = 0.3f; ms[2].color[3] = 1.0f;
    size_t idx = ms.size();
    static const float kColorDef[4] = {1, 0, 1, 1}, kEmisDef[4] = {0, 0, 0, 0};
    for (size_t i = 0; i < mats.map.size(); i++) {
        if (!mats.key_is_str[i]) throw invalid(mats.key_dbg[i], "materials", kHintString);
        const std::string& name = mats.map[i].first;
        const YNode& body = mats.map[i].second;
        SandMaterial m;
        m.id = idx++;
        m.name = name;
        const YNode* ty = body.get("type");
        if (!ty) throw missing("type", "materials/" + name);
        if (!ty->as_str(&m.mattype)) throw invalid(quoted("type"), "materials/" + name, kHintString);
        if (!contains(type_names, m.mattype)) throw notfound(m.mattype, "materials/" + name + "/type");
        extract_vec4(body, name, "color", kColorDef, true, m.color);
        extract_vec4(body, name, "emission", kEmisDef, false, m.emission);
        if (const YNode* sel = body.get("selectable")) {
            bool b;
            m.selectable = sel->as_bool(&b) ? b : false;
        } else {
            m.selectable = true;
        }
        const YNode* de = body.get("density");
        if (!de) throw missing("density", "materials/" + name);
        double d;
        if (!de->as_f64(&d)) throw invalid(quoted("density"), "materials/" + name + "/density", kHintFloat);
        m.density = (float)d;
        if (const YNode* er = body.get("extra_rules")) {
            if (er->kind != YNode::Seq) throw invalid(quoted("extra_rules"), "materials/" + name, kHintSeq);
            for (auto& e : er->seq) {
                std::string en;
                if (!e.as_str(&en)) throw invalid(quoted("extra_rules"), "materials/" + name, kHintString);
                for (auto& r : rules)
                    if (r.name == en) {
                        r.used = true;
                        update_pre(r, "self.mat == MAT_" + name);
                        m.extra_rules.push_back(en);
                    }
            }
        }
        ms.push_back(std::move(m));
    }
    return ms;
}

std::string func_logic(std::vector<std::string> ifs, std::vector<std::string> dos, std::vector<float> ps,
                       size_t indent) {
    if (ifs.empty() && dos.empty()) return "";
    if (ifs.empty()) return dos[0];
    if (dos.empty()) return "";
    std::string ind1(indent * 4, ' '), ind2((indent + 1) * 4, ' ');
    float p = ps[0];
    std::string prob = p == 1.0f ? std::string() : "rand.y <= " + fmt_f32(p) + " && ";
    std::string c0 = ifs[0], a0 = dos[0];
    ifs.erase(ifs.begin());
    dos.erase(dos.begin());
    ps.erase(ps.begin());
    return ind1 + "if (" + prob + c0 + ") {\n" + ind2 + a0 + "\n" + ind1 + "} else {\n" +
           func_logic(ifs, dos, ps, indent + 1) + "\n" + ind1 + "}";
}

std::string rule_glsl(const SandRule& r) {
    std::string celldir = r.ruletype == RuleType::Left ? "left" : "right";
    std::string pre = r.precondition ? "    if (!(" + *r.precondition + ")) {\n        return;\n    }\n" : "";
    return "void rule_" + r.name + " (inout Cell self, inout Cell " + celldir +
           ", inout Cell down, inout Cell downright, vec4 rand, ivec2 pos) {\n" + pre +
           func_logic(r.if_conds, r.do_actions, r.probabilities, 1) + "\n}";
}

std::string json_str(const std::string& s) {
    std::string o = "\"";
    for (char c : s) {
        if (c == '"') o += "\\\"";
        else if (c == '\\') o += "\\\\";
        else if (c == '\n') o += "\\n";
        else if ((unsigned char)c < 0x20) { char b[8]; snprintf(b, sizeof b, "\\u%04x", c); o += b; }
        else o += c;
    }
    return o + "\"";
}
std::string json_f(float v) {
    char b[64];
    snprintf(b, sizeof b, "%.9g", (double)v);
    std::string s = b;
    if (s == "nan" || s == "-nan") return "NaN";
    if (s == "inf") return "Infinity";
    if (s == "-inf") return "-Infinity";
    return s;
}
std::string json_strs(const std::vector<std::string>& v) {
    std::string o = "[";
    for (size_t i = 0; i < v.size(); i++) o += (i ? ", " : "") + json_str(v[i]);
    return o + "]";
}

}  // namespace

std::string fmt_f32(float v) {
    char buf[128];
    auto res = std::to_chars(buf, buf + sizeof buf, v, std::chars_format::fixed);
    std::string s(buf, res.ptr);
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
    return s;
}

ParsingResult parse_string(const std::string& text) {
    YNode data;
    try {
        data = yaml_parse(text);
    } catch (const YamlError& e) {
        throw LangError("Yaml", e.what());
    }
    const YNode& raw_rules = check_mapping(data, "rules");
    const YNode& raw_types = check_mapping(data, "types");
    const YNode& raw_mats = check_mapping(data, "materials");
    auto rule_names = preparse_keys(raw_rules, "rules");
    auto type_names = preparse_keys(raw_types, "types");
    type_names.push_back("EMPTY");
    auto mat_names = preparse_keys(raw_mats, "materials");
    mat_names.push_back("EMPTY");
    ParsingResult res;
    try {
        res.rules = parse_rules(raw_rules, type_names, mat_names);
    } catch (const LangError& e) {
        throw LangError(e.kind, std::string("Error while parsing rules: '") + e.what() + "'");
    }
    try {
        res.types = parse_types(raw_types, res.rules, rule_names, type_names);
    } catch (const LangError& e) {
        throw LangError(e.kind, std::string("Error while parsing types: '") + e.what() + "'");
    }
    try {
        res.materials = parse_materials(raw_mats, res.rules, type_names);
    } catch (const LangError& e) {
        throw LangError(e.kind, std::string("Error while parsing materials: '") + e.what() + "'");
    }
    return res;
}

std::pair<std::string, std::string> create_glsl(const ParsingResult& r) {
    std::string s;
    for (auto& t : r.types) s += "#define TYPE_" + t.name + " " + std::to_string(t.id) + "\n\n";
    for (auto& t : r.types) {
        std::string tc = "return cell.mat.type == TYPE_" + t.name;
        for (auto& c : t.children) tc += " || cell.mat.type == TYPE_" + c;
        s += "bool isType_" + t.name + "(Cell cell) {\n    " + tc + ";\n}\n\n";
    }
    s += "\n";
    std::string list;
    size_t n = r.materials.size();
    for (size_t i = 0; i < n; i++) {
        const auto& m = r.materials[i];
        s += "#define MAT_" + m.name + " Material(" + std::to_string(m.id) + ", vec4(" + fmt_f32(m.color[0]) + ", " +
             fmt_f32(m.color[1]) + ", " + fmt_f32(m.color[2]) + ", " + fmt_f32(m.color[3]) + "), " +
             fmt_f32(m.density) + ", vec4(" + fmt_f32(m.emission[0]) + ", " + fmt_f32(m.emission[1]) + ", " +
             fmt_f32(m.emission[2]) + ", " + fmt_f32(m.emission[3]) + "), TYPE_" + m.mattype + ")\n";
        list += "        MAT_" + m.name + (i == n - 1 ? "" : ",\n");
    }
    std::string ns = std::to_string(n);
    s += "\nMaterial[" + ns + "] materials() {\n    Material allMaterials[" + ns + "] = {\n" + list +
         "\n    };\n    return allMaterials;\n}\n\nMaterial getMaterialFromID(int id) {\n"
         "    for (int i = 0; i < materials().length(); i++) {\n        if (id == materials()[i].id) {\n"
         "            return materials()[i];\n        };\n    };\n    return MAT_NULL;\n}\n\n";
    std::string funcs, mir, left, right;
    for (auto& rl : r.rules) {
        if (!rl.used) continue;
        funcs += rule_glsl(rl) + "\n\n";
        if (rl.ruletype == RuleType::Mirrored) mir += "rule_" + rl.name + "(self, right, down, downright, rand, pos);\n";
        else if (rl.ruletype == RuleType::Left) left += "rule_" + rl.name + "(self, left, down, downright, rand, pos);\n";
        else right += "rule_" + rl.name + "(self, right, down, downright, rand, pos);\n";
    }
    std::string hdr =
        "    inout Cell self,\n    inout Cell right,\n    inout Cell down,\n    inout Cell downright,\n"
        "    vec4 rand,\n    ivec2 pos) {\n";
    std::string rules = "\n// =============== RULES ===============\n" + funcs +
                        "\n\n\n// =============== CALLERS ===============\nvoid applyMirroredRules(\n" + hdr + "    " +
                        rtrim(mir) + "\n}\n\n\nvoid applyLeftRules(\n" + hdr + "    " + rtrim(left) +
                        "\n}\n\nvoid applyRightRules(\n" + hdr + "    " + rtrim(right) + "\n}";
    return {s, rules};
}

std::string ir_json(const ParsingResult& r) {
    std::ostringstream o;
    o << "{\"rules\": [";
    for (size_t i = 0; i < r.rules.size(); i++) {
        const auto& x = r.rules[i];
        const char* rt = x.ruletype == RuleType::Mirrored ? "Mirrored" : x.ruletype == RuleType::Left ? "Left" : "Right";
        o << (i ? ", " : "") << "{\"name\": " << json_str(x.name) << ", \"ruletype\": \"" << rt
          << "\", \"if_conds\": " << json_strs(x.if_conds) << ", \"do_actions\": " << json_strs(x.do_actions)
          << ", \"probabilities\": [";
        for (size_t k = 0; k < x.probabilities.size(); k++) o << (k ? ", " : "") << json_f(x.probabilities[k]);
        o << "], \"mirror\": " << (x.mirror ? "true" : "false") << ", \"precondition\": "
          << (x.precondition ? json_str(*x.precondition) : "null") << ", \"used\": " << (x.used ? "true" : "false")
          << "}";
    }
    o << "], \"types\": [";
    for (size_t i = 0; i < r.types.size(); i++) {
        const auto& t = r.types[i];
        o << (i ? ", " : "") << "{\"id\": " << t.id << ", \"name\": " << json_str(t.name)
          << ", \"inherits\": " << json_str(t.inherits) << ", \"children\": " << json_strs(t.children)
          << ", \"base_rules\": " << json_strs(t.base_rules) << "}";
    }
    o << "], \"materials\": [";
    for (size_t i = 0; i < r.materials.size(); i++) {
        const auto& m = r.materials[i];
        o << (i ? ", " : "") << "{\"id\": " << m.id << ", \"name\": " << json_str(m.name)
          << ", \"mattype\": " << json_str(m.mattype) << ", \"color\": [";
        for (int k = 0; k < 4; k++) o << (k ? ", " : "") << json_f(m.color[k]);
        o << "], \"emission\": [";
        for (int k = 0; k < 4; k++) o << (k ? ", " : "") << json_f(m.emission[k]);
        o << "], \"selectable\": " << (m.selectable ? "true" : "false") << ", \"density\": " << json_f(m.density)
          << ", \"extra_rules\": " << json_strs(m.extra_rules) << "}";
    }
    o << "]}";
    return o.str();
}

}  // namespace se
