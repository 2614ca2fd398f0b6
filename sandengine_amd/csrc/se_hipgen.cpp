#include "se_hipgen.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <sstream>

namespace se {

static const char* kKernelTemplates =
#include "se_kernels_src.inc"
    ;

namespace {

constexpr uint64_t kTableMaxBytes = 48 * 1024;  // transition table staged in LDS
constexpr uint64_t kCTableMaxBytes = 64 * 1024;  // coded table: D^4 u8 entries staged in LDS

[[noreturn]] void gen_error(const std::string& m) { throw LangError("NotRecognized", "(HIP codegen) " + m); }

// ------------------------------------------------------------------ expressions
enum class Ty { Bool, Num, Mat };

struct Ex {
    enum K { Num, Bool, CellMat, MatConst, Dens, MatId, TypeId, IsType, Rand, Cmp, And, Or, Not, Neg } k;
    Ty ty = Ty::Bool;
    int op = 0;  // Cmp: 0 <, 1 <=, 2 >, 3 >=, 4 ==, 5 !=
    float num = 0;
    bool bval = false;
    int cell = 0, mat = 0, type = 0, lane = 0;
    std::unique_ptr<Ex> a, b;
};
using ExP = std::unique_ptr<Ex>;

struct Ctx {
    std::map<std::string, int> mats;   // name -> id
    std::map<std::string, int> types;  // name -> id
    std::vector<int> mat_type;         // material id -> type id
    std::vector<float> dens;           // material id -> density
    std::vector<std::array<float, 4>> color, emission;  // material id -> RGBA
    std::vector<std::vector<int>> type_members;  // type id -> {type id} + children type ids
    int nmat = 0;
};

int cell_index(const std::string& n) {
    if (n == "self") return 0;
    if (n == "right") return 1;
    if (n == "down") return 2;
    if (n == "downright") return 3;
    if (n == "left" || n == "downleft")
        gen_error("cell '" + n + "' is not declared in the generated rule functions (only self, right, down, "
                  "downright are)");
    gen_error("unknown cell '" + n + "'");
}

struct Tok {
    int kind;  // 0 num, 1 id, 2 op
    std::string s;
};

std::vector<Tok> tokenize(const std::string& s) {
    std::vector<Tok> out;
    size_t i = 0;
    while (i < s.size()) {
        char c = s[i];
        if (std::isspace((unsigned char)c)) { i++; continue; }
        if (std::isdigit((unsigned char)c) || (c == '.' && i + 1 < s.size() && std::isdigit((unsigned char)s[i + 1]))) {
            size_t j = i;
            while (j < s.size() && std::isdigit((unsigned char)s[j])) j++;
            if (j < s.size() && s[j] == '.') { j++; while (j < s.size() && std::isdigit((unsigned char)s[j])) j++; }
            if (j < s.size() && (s[j] == 'e' || s[j] == 'E')) {
                size_t k = j + 1;
                if (k < s.size() && (s[k] == '+' || s[k] == '-')) k++;
                if (k < s.size() && std::isdigit((unsigned char)s[k])) {
                    j = k;
                    while (j < s.size() && std::isdigit((unsigned char)s[j])) j++;
                }
            }
            std::string num = s.substr(i, j - i);
            if (j < s.size() && (s[j] == 'f' || s[j] == 'F')) j++;
            out.push_back({0, num});
            i = j;
            continue;
        }
        if (std::isalpha((unsigned char)c) || c == '_') {
            size_t j = i;
            while (j < s.size() && (std::isalnum((unsigned char)s[j]) || s[j] == '_')) j++;
            out.push_back({1, s.substr(i, j - i)});
            i = j;
            continue;
        }
        static const char* two[] = {"&&", "||", "==", "!=", "<=", ">="};
        bool got = false;
        for (auto* t : two)
            if (s.compare(i, 2, t) == 0) { out.push_back({2, t}); i += 2; got = true; break; }
        if (got) continue;
        if (std::strchr("<>!().-", c)) { out.push_back({2, std::string(1, c)}); i++; continue; }
        gen_error("unsupported character '" + std::string(1, c) + "' in condition '" + s + "'");
    }
    return out;
}

class ExprParser {
public:
    ExprParser(const std::string& src, const Ctx& ctx) : src_(src), t_(tokenize(src)), ctx_(ctx) {}

    ExP parse() {
        ExP e = orx();
        if (i_ != t_.size()) gen_error("unexpected '" + t_[i_].s + "' in condition '" + src_ + "'");
        if (e->ty != Ty::Bool) gen_error("condition is not boolean: '" + src_ + "'");
        return e;
    }

private:
    std::string src_;
    std::vector<Tok> t_;
    size_t i_ = 0;
    const Ctx& ctx_;

    bool peek_op(const char* o) const { return i_ < t_.size() && t_[i_].kind == 2 && t_[i_].s == o; }
    Tok take() {
        if (i_ >= t_.size()) gen_error("unexpected end of condition '" + src_ + "'");
        return t_[i_++];
    }
    void expect_op(const char* o) {
        if (!peek_op(o)) gen_error(std::string("expected '") + o + "' in condition '" + src_ + "'");
        i_++;
    }
    std::string expect_id() {
        Tok t = take();
        if (t.kind != 1) gen_error("expected identifier in condition '" + src_ + "'");
        return t.s;
    }
    static ExP mk(Ex::K k, Ty ty) {
        ExP e(new Ex());
        e->k = k;
        e->ty = ty;
        return e;
    }
    ExP bin(Ex::K k, ExP a, ExP b) {
        if (a->ty != Ty::Bool || b->ty != Ty::Bool) gen_error("&&/|| on non-boolean in '" + src_ + "'");
        ExP e = mk(k, Ty::Bool);
        e->a = std::move(a);
        e->b = std::move(b);
        return e;
    }
    ExP orx() {
        ExP e = andx();
        while (peek_op("||")) { i_++; e = bin(Ex::Or, std::move(e), andx()); }
        return e;
    }
    ExP andx() {
        ExP e = eqx();
        while (peek_op("&&")) { i_++; e = bin(Ex::And, std::move(e), eqx()); }
        return e;
    }
    ExP cmp(int op, ExP a, ExP b) {
        if (op >= 4) {
            if (a->ty != b->ty) gen_error("== / != between different types in '" + src_ + "'");
        } else if (a->ty != Ty::Num || b->ty != Ty::Num) {
            gen_error("relational comparison of non-numbers in '" + src_ + "'");
        }
        ExP e = mk(Ex::Cmp, Ty::Bool);
        e->op = op;
        e->a = std::move(a);
        e->b = std::move(b);
        return e;
    }
    ExP eqx() {
        ExP e = relx();
        for (;;) {
            if (peek_op("==")) { i_++; e = cmp(4, std::move(e), relx()); }
            else if (peek_op("!=")) { i_++; e = cmp(5, std::move(e), relx()); }
            else return e;
        }
    }
    ExP relx() {
        ExP e = unary();
        for (;;) {
            int op = -1;
            if (peek_op("<")) op = 0;
            else if (peek_op("<=")) op = 1;
            else if (peek_op(">")) op = 2;
            else if (peek_op(">=")) op = 3;
            if (op < 0) return e;
            i_++;
            e = cmp(op, std::move(e), unary());
        }
    }
    ExP unary() {
        if (peek_op("!")) {
            i_++;
            ExP a = unary();
            if (a->ty != Ty::Bool) gen_error("! on non-boolean in '" + src_ + "'");
            ExP e = mk(Ex::Not, Ty::Bool);
            e->a = std::move(a);
            return e;
        }
        if (peek_op("-")) {
            i_++;
            ExP a = unary();
            if (a->ty != Ty::Num) gen_error("unary - on non-number in '" + src_ + "'");
            ExP e = mk(Ex::Neg, Ty::Num);
            e->a = std::move(a);
            return e;
        }
        return primary();
    }
    ExP primary() {
        Tok t = take();
        if (t.kind == 0) {
            ExP e = mk(Ex::Num, Ty::Num);
            e->num = std::strtof(t.s.c_str(), nullptr);
            return e;
        }
        if (t.kind == 2 && t.s == "(") {
            ExP e = orx();
            expect_op(")");
            return e;
        }
        if (t.kind != 1) gen_error("unexpected '" + t.s + "' in condition '" + src_ + "'");
        const std::string& id = t.s;
        if (id == "true" || id == "false") {
            ExP e = mk(Ex::Bool, Ty::Bool);
            e->bval = id == "true";
            return e;
        }
        if (id.rfind("isType_", 0) == 0) {
            std::string tn = id.substr(7);
            auto it = ctx_.types.find(tn);
            if (it == ctx_.types.end()) gen_error("unknown type '" + tn + "'");
            expect_op("(");
            std::string c = expect_id();
            expect_op(")");
            ExP e = mk(Ex::IsType, Ty::Bool);
            e->cell = cell_index(c);
            e->type = it->second;
            return e;
        }
        if (id.rfind("MAT_", 0) == 0) {
            auto it = ctx_.mats.find(id.substr(4));
            if (it == ctx_.mats.end()) gen_error("unknown material '" + id + "'");
            ExP e = mk(Ex::MatConst, Ty::Mat);
            e->mat = it->second;
            return e;
        }
        if (id.rfind("TYPE_", 0) == 0) {
            auto it = ctx_.types.find(id.substr(5));
            if (it == ctx_.types.end()) gen_error("unknown type '" + id + "'");
            ExP e = mk(Ex::Num, Ty::Num);
            e->num = (float)it->second;
            return e;
        }
        if (id == "rand") {
            expect_op(".");
            std::string sw = expect_id();
            static const char* lanes = "xyzw";
            static const char* lanes2 = "rgba";
            if (sw.size() != 1) gen_error("rand swizzle '" + sw + "' unsupported");
            const char* p = std::strchr(lanes, sw[0]);
            const char* q = std::strchr(lanes2, sw[0]);
            if (!p && !q) gen_error("rand swizzle '" + sw + "' unsupported");
            ExP e = mk(Ex::Rand, Ty::Num);
            e->lane = p ? (int)(p - lanes) : (int)(q - lanes2);
            return e;
        }
        int cell = cell_index(id);
        expect_op(".");
        std::string f = expect_id();
        if (f != "mat") gen_error("only '.mat' is supported on cells (got '" + f + "')");
        if (peek_op(".")) {
            i_++;
            std::string fld = expect_id();
            ExP e;
            if (fld == "density") e = mk(Ex::Dens, Ty::Num);
            else if (fld == "id") e = mk(Ex::MatId, Ty::Num);
            else if (fld == "type") e = mk(Ex::TypeId, Ty::Num);
            else gen_error("material field '" + fld + "' unsupported in conditions");
            e->cell = cell;
            return e;
        }
        ExP e = mk(Ex::CellMat, Ty::Mat);
        e->cell = cell;
        return e;
    }
};

// ------------------------------------------------------------------ analysis and folding
unsigned cells_used(const Ex& e) {
    unsigned m = 0;
    switch (e.k) {
        case Ex::CellMat: case Ex::Dens: case Ex::MatId: case Ex::TypeId: case Ex::IsType: m = 1u << e.cell; break;
        default: break;
    }
    if (e.a) m |= cells_used(*e.a);
    if (e.b) m |= cells_used(*e.b);
    return m;
}
bool uses_rand(const Ex& e) {
    if (e.k == Ex::Rand) return true;
    return (e.a && uses_rand(*e.a)) || (e.b && uses_rand(*e.b));
}

struct Env {
    int c[4];
    float r[4];
};

float eval_num(const Ex& e, const Env& v, const Ctx& ctx);
int eval_mat(const Ex& e, const Env& v) { return e.k == Ex::CellMat ? v.c[e.cell] : e.mat; }
bool eval_bool(const Ex& e, const Env& v, const Ctx& ctx) {
    switch (e.k) {
        case Ex::Bool: return e.bval;
        case Ex::IsType: {
            int ty = ctx.mat_type[v.c[e.cell]];
            for (int t : ctx.type_members[e.type])
                if (t == ty) return true;
            return false;
        }
        case Ex::And: return eval_bool(*e.a, v, ctx) && eval_bool(*e.b, v, ctx);
        case Ex::Or: return eval_bool(*e.a, v, ctx) || eval_bool(*e.b, v, ctx);
        case Ex::Not: return !eval_bool(*e.a, v, ctx);
        case Ex::Cmp: {
            if (e.a->ty == Ty::Mat) {
                bool eq = eval_mat(*e.a, v) == eval_mat(*e.b, v);
                return e.op == 4 ? eq : !eq;
            }
            if (e.a->ty == Ty::Bool) {
                bool eq = eval_bool(*e.a, v, ctx) == eval_bool(*e.b, v, ctx);
                return e.op == 4 ? eq : !eq;
            }
            float x = eval_num(*e.a, v, ctx), y = eval_num(*e.b, v, ctx);
            switch (e.op) {
                case 0: return x < y;
                case 1: return x <= y;
                case 2: return x > y;
                case 3: return x >= y;
                case 4: return x == y;
                default: return x != y;
            }
        }
        default: gen_error("internal: not a boolean node");
    }
}
float eval_num(const Ex& e, const Env& v, const Ctx& ctx) {
    switch (e.k) {
        case Ex::Num: return e.num;
        case Ex::Dens: return ctx.dens[v.c[e.cell]];
        case Ex::MatId: return (float)v.c[e.cell];
        case Ex::TypeId: return (float)ctx.mat_type[v.c[e.cell]];
        case Ex::Rand: return v.r[e.lane];
        case Ex::Neg: return -eval_num(*e.a, v, ctx);
        default: gen_error("internal: not a numeric node");
    }
}

// rand = RNE_f32(u) * 2^-32, monotone in u.  thr(c, le): max u with rand(u) (<=|<) c, or -1.
float rnd_of(uint64_t u) { return (float)(uint32_t)u * 0x1p-32f; }
int64_t rand_threshold(float c, bool le) {
    auto ok = [&](uint64_t u) { return le ? rnd_of(u) <= c : rnd_of(u) < c; };
    if (!ok(0)) return -1;
    uint64_t lo = 0, hi = 0xFFFFFFFFull;
    while (lo < hi) {
        uint64_t mid = (lo + hi + 1) / 2;
        if (ok(mid)) lo = mid;
        else hi = mid - 1;
    }
    return (int64_t)lo;
}

std::string hex_u32(uint64_t v) {
    char b[32];
    snprintf(b, sizeof b, "0x%08llxu", (unsigned long long)(v & 0xFFFFFFFFull));
    return b;
}
std::string hex_u64(uint64_t v) {
    char b[40];
    snprintf(b, sizeof b, "0x%016llxull", (unsigned long long)v);
    return b;
}
std::string flit(float f) {
    if (std::isnan(f)) return "__builtin_nanf(\"\")";
    if (std::isinf(f)) return f > 0 ? "__builtin_inff()" : "(-__builtin_inff())";
    char b[64];
    snprintf(b, sizeof b, "%af", (double)f);
    return b;
}
const char* kCell[4] = {"c0", "c1", "c2", "c3"};

class Emitter {
public:
    explicit Emitter(const Ctx& ctx) : ctx_(ctx) {}

    // material bitmask as a device expression testing id `c`
    std::string bit(const std::vector<bool>& m, const std::string& c) const {
        int n = ctx_.nmat;
        bool all = true, none = true;
        for (int i = 0; i < n; i++) { all &= (bool)m[i]; none &= !m[i]; }
        if (all) return "true";
        if (none) return "false";
        int ones = 0, one = -1;
        for (int i = 0; i < n; i++)
            if (m[i]) { ones++; one = i; }
        if (ones == 1) return "(" + c + " == " + std::to_string(one) + "u)";
        if (n <= 32) {
            uint64_t w = 0;
            for (int i = 0; i < n; i++)
                if (m[i]) w |= 1ull << i;
            return "(((" + hex_u32(w) + " >> " + c + ") & 1u) != 0u)";
        }
        std::string words[4];
        int nw = (n + 63) / 64;
        for (int k = 0; k < nw; k++) {
            uint64_t w = 0;
            for (int i = 0; i < 64 && k * 64 + i < n; i++)
                if (m[k * 64 + i]) w |= 1ull << i;
            words[k] = hex_u64(w);
        }
        if (nw == 1) return "(((" + words[0] + " >> " + c + ") & 1ull) != 0ull)";
        std::string sel = words[nw - 1];
        for (int k = nw - 2; k >= 0; k--) sel = "(" + c + " < " + std::to_string((k + 1) * 64) + "u ? " + words[k] + " : " + sel + ")";
        return "(((" + sel + " >> (" + c + " & 63u)) & 1ull) != 0ull)";
    }

    std::string emit_bool(const Ex& e) {
        unsigned cm = cells_used(e);
        bool r = uses_rand(e);
        if (!r && cm == 0) {
            Env v{};
            return eval_bool(e, v, ctx_) ? "true" : "false";
        }
        if (!r && (cm & (cm - 1)) == 0) {
            int k = __builtin_ctz(cm);
            std::vector<bool> m(ctx_.nmat);
            for (int mat = 0; mat < ctx_.nmat; mat++) {
                Env v{};
                v.c[k] = mat;
                m[mat] = eval_bool(e, v, ctx_);
            }
            return bit(m, kCell[k]);
        }
        switch (e.k) {
            case Ex::And: return "(" + emit_bool(*e.a) + " && " + emit_bool(*e.b) + ")";
            case Ex::Or: return "(" + emit_bool(*e.a) + " || " + emit_bool(*e.b) + ")";
            case Ex::Not: return "(!" + emit_bool(*e.a) + ")";
            case Ex::Cmp: return emit_cmp(e);
            default: gen_error("internal: unexpected boolean node");
        }
    }

    std::set<int> lanes_used;
    std::map<int, std::set<uint32_t>> thresholds;  // lane -> t for every "h <= t" / "h > t" emitted
    bool rand_float_used = false;                  // rand read as a float (no threshold form)
    bool rank_used = false, dens_used = false, type_used = false;

private:
    const Ctx& ctx_;

    static const char* opstr(int op) {
        static const char* s[] = {"<", "<=", ">", ">=", "==", "!="};
        return s[op];
    }

    std::string emit_cmp(const Ex& e) {
        if (e.a->ty == Ty::Mat) {
            std::string x = e.a->k == Ex::CellMat ? kCell[e.a->cell] : std::to_string(e.a->mat) + "u";
            std::string y = e.b->k == Ex::CellMat ? kCell[e.b->cell] : std::to_string(e.b->mat) + "u";
            return "(" + x + (e.op == 4 ? " == " : " != ") + y + ")";
        }
        if (e.a->ty == Ty::Bool)
            return "(" + emit_bool(*e.a) + (e.op == 4 ? " == " : " != ") + emit_bool(*e.b) + ")";
        // density of one cell against density of another: compare ranks
        if (e.a->k == Ex::Dens && e.b->k == Ex::Dens) {
            rank_used = true;
            return "(se_rank(" + std::string(kCell[e.a->cell]) + ") " + opstr(e.op) + " se_rank(" + kCell[e.b->cell] + "))";
        }
        // rand lane against a constant: integer threshold on the hash lane
        const Ex* rn = nullptr;
        const Ex* cst = nullptr;
        int op = e.op;
        if (e.a->k == Ex::Rand && !uses_rand(*e.b) && cells_used(*e.b) == 0) { rn = e.a.get(); cst = e.b.get(); }
        else if (e.b->k == Ex::Rand && !uses_rand(*e.a) && cells_used(*e.a) == 0) {
            rn = e.b.get(); cst = e.a.get();
            static const int flip[] = {2, 3, 0, 1, 4, 5};
            op = flip[op];
        }
        if (rn) {
            Env v{};
            float c = eval_num(*cst, v, ctx_);
            lanes_used.insert(rn->lane);
            std::string h = "h" + std::to_string(rn->lane);
            if (std::isnan(c)) return op == 5 ? "true" : "false";
            int64_t tle = rand_threshold(c, true), tlt = rand_threshold(c, false);
            const int lane = rn->lane;
            auto le = [&](int64_t t) -> std::string {  // h <= t
                if (t < 0) return "false";
                if (t >= 0xFFFFFFFFll) return "true";
                thresholds[lane].insert((uint32_t)t);
                return "(" + h + " <= " + hex_u32((uint64_t)t) + ")";
            };
            auto gt = [&](int64_t t) -> std::string {  // h > t
                if (t < 0) return "true";
                if (t >= 0xFFFFFFFFll) return "false";
                thresholds[lane].insert((uint32_t)t);
                return "(" + h + " > " + hex_u32((uint64_t)t) + ")";
            };
            switch (op) {
                case 0: return le(tlt);
                case 1: return le(tle);
                case 2: return gt(tle);
                case 3: return gt(tlt);
                case 4: return "(" + gt(tlt) + " && " + le(tle) + ")";
                default: return "(!(" + gt(tlt) + " && " + le(tle) + "))";
            }
        }
        return "(" + emit_num(*e.a) + " " + opstr(e.op) + " " + emit_num(*e.b) + ")";
    }

    std::string emit_num(const Ex& e) {
        switch (e.k) {
            case Ex::Num: return flit(e.num);
            case Ex::Dens: dens_used = true; return "se_dens(" + std::string(kCell[e.cell]) + ")";
            case Ex::MatId: return "((float)" + std::string(kCell[e.cell]) + ")";
            case Ex::TypeId: type_used = true; return "((float)se_type(" + std::string(kCell[e.cell]) + "))";
            case Ex::Rand:
                lanes_used.insert(e.lane);
                rand_float_used = true;
                return "((float)h" + std::to_string(e.lane) + " * 0x1p-32f)";
            case Ex::Neg: return "(-" + emit_num(*e.a) + ")";
            default: gen_error("internal: unexpected numeric node");
        }
    }
};

// ------------------------------------------------------------------ actions
std::string emit_actions(const std::string& text, const Ctx& ctx, const std::string& ind) {
    std::string out;
    std::istringstream is(text);
    std::string line;
    while (std::getline(is, line)) {
        size_t b = line.find_first_not_of(" \t"), e = line.find_last_not_of(" \t\r");
        if (b == std::string::npos) continue;
        line = line.substr(b, e - b + 1);
        if (line.rfind("swap(", 0) == 0) {
            size_t comma = line.find(", "), close = line.find(");");
            if (comma == std::string::npos || close == std::string::npos || close + 2 != line.size())
                gen_error("unrecognised action '" + line + "'");
            int a = cell_index(line.substr(5, comma - 5)), c = cell_index(line.substr(comma + 2, close - comma - 2));
            out += ind + "SE_GSWAP(" + kCell[a] + ", " + kCell[c] + ");\n";
            continue;
        }
        size_t eq = line.find(" = newCell(MAT_");
        if (eq != std::string::npos) {
            int a = cell_index(line.substr(0, eq));
            size_t s = eq + std::strlen(" = newCell(MAT_");
            size_t comma = line.find(", pos);", s);
            if (comma == std::string::npos || comma + 7 != line.size()) gen_error("unrecognised action '" + line + "'");
            std::string mat = line.substr(s, comma - s);
            auto it = ctx.mats.find(mat);
            if (it == ctx.mats.end())
                gen_error("SET to unknown material 'MAT_" + mat + "' (not declared in the generated GLSL either)");
            out += ind + kCell[a] + " = " + std::to_string(it->second) + "u;\n";
            continue;
        }
        gen_error("unrecognised action '" + line + "'");
    }
    return out;
}

Ctx build_ctx(const ParsingResult& r) {
    Ctx ctx;
    std::set<std::string> tnames, mnames;
    for (auto& t : r.types) {
        if (!tnames.insert(t.name).second) gen_error("type '" + t.name + "' is defined twice (TYPE_ redefinition)");
        ctx.types[t.name] = (int)t.id;
    }
    for (auto& m : r.materials) {
        if (!mnames.insert(m.name).second) gen_error("material '" + m.name + "' is defined twice (MAT_ redefinition)");
        ctx.mats[m.name] = (int)m.id;
    }
    ctx.nmat = (int)r.materials.size();
    if (ctx.nmat > 255) gen_error("more than 255 materials are not supported");
    ctx.mat_type.resize(ctx.nmat);
    ctx.dens.resize(ctx.nmat);
    ctx.color.resize(ctx.nmat);
    ctx.emission.resize(ctx.nmat);
    for (auto& m : r.materials) {
        auto it = ctx.types.find(m.mattype);
        if (it == ctx.types.end()) gen_error("material '" + m.name + "' has unknown type");
        ctx.mat_type[m.id] = it->second;
        ctx.dens[m.id] = m.density;
        for (int k = 0; k < 4; k++) {
            ctx.color[m.id][k] = m.color[k];
            ctx.emission[m.id][k] = m.emission[k];
        }
    }
    ctx.type_members.resize(r.types.size());
    for (auto& t : r.types) {
        ctx.type_members[t.id].push_back((int)t.id);
        for (auto& c : t.children) {
            auto it = ctx.types.find(c);
            if (it != ctx.types.end()) ctx.type_members[t.id].push_back(it->second);
        }
    }
    return ctx;
}

// rules.rs:47-73 pairing of the flattened if/do/probability lists.
struct Level {
    float p;
    std::string cond, act;
};
void levels_of(const SandRule& r, std::vector<Level>& lv, std::string* final_act, bool* has_final) {
    std::vector<std::string> ifs = r.if_conds, dos = r.do_actions;
    std::vector<float> ps = r.probabilities;
    *has_final = false;
    for (;;) {
        if (ifs.empty() && dos.empty()) break;
        if (ifs.empty()) { *final_act = dos[0]; *has_final = true; break; }
        if (dos.empty()) break;
        lv.push_back({ps[0], ifs[0], dos[0]});
        ifs.erase(ifs.begin());
        dos.erase(dos.begin());
        ps.erase(ps.begin());
    }
}

std::string emit_rule(const SandRule& r, const Ctx& ctx, Emitter& em) {
    std::ostringstream o;
    o << "    // rule " << r.name << "\n";
    std::string ind = "    ";
    bool guarded = false;
    {
        // every condition and action must translate (the GLSL would fail to compile
        // otherwise), even where folding later removes the code
        Emitter probe(ctx);
        std::vector<Level> lv0;
        std::string fin0;
        bool hf0;
        levels_of(r, lv0, &fin0, &hf0);
        for (auto& l : lv0) {
            std::string text = (l.p == 1.0f ? std::string() : "rand.y <= " + fmt_f32(l.p) + " && ") + l.cond;
            probe.emit_bool(*ExprParser(text, ctx).parse());
            emit_actions(l.act, ctx, "");
        }
        if (hf0) emit_actions(fin0, ctx, "");
    }
    if (r.precondition) {
        ExP pre = ExprParser(*r.precondition, ctx).parse();
        std::string c = em.emit_bool(*pre);
        if (c == "false") return o.str() + "    // (precondition never holds)\n";
        if (c != "true") {
            o << "    if (" << c << ") {\n";
            ind = "        ";
            guarded = true;
        }
    }
    std::vector<Level> lv;
    std::string fin;
    bool has_fin;
    levels_of(r, lv, &fin, &has_fin);
    int open = 0;
    for (size_t i = 0; i < lv.size(); i++) {
        std::string text = (lv[i].p == 1.0f ? std::string() : "rand.y <= " + fmt_f32(lv[i].p) + " && ") + lv[i].cond;
        ExP ce = ExprParser(text, ctx).parse();
        std::string c = em.emit_bool(*ce);
        std::string pad = ind + std::string(4 * open, ' ');
        o << pad << "if (" << c << ") {\n" << emit_actions(lv[i].act, ctx, pad + "    ") << pad << "} else {\n";
        open++;
    }
    if (has_fin) o << emit_actions(fin, ctx, ind + std::string(4 * open, ' '));
    for (int k = open - 1; k >= 0; k--) o << ind << std::string(4 * k, ' ') << "}\n";
    if (guarded) o << "    }\n";
    return o.str();
}

}  // namespace

std::string hip_rules_source(const ParsingResult& r, HipGenInfo* info, bool allow_ctable, uint64_t ct_x_words) {
    Ctx ctx = build_ctx(r);
    Emitter em(ctx);
    std::string mirrored, right;
    int nrules = 0;
    for (auto& rl : r.rules) {
        if (!rl.used) continue;
        nrules++;
        if (rl.ruletype == RuleType::Left)
            gen_error("rule '" + rl.name + "' is a Left rule; its generated call passes an undeclared 'left' cell");
        std::string body = emit_rule(rl, ctx, em);
        if (rl.ruletype == RuleType::Mirrored) mirrored += body;
        else right += body;
    }
    std::ostringstream o;
    const int n = ctx.nmat;
    o << "// ---- generated by the sandengine HIP backend: " << n << " materials, " << nrules << " rules\n";
    o << "#define SE_NMAT " << n << "u\n";
    // guarded swap refuses WALL/NULL-typed cells (operations.glsl:16-23)
    std::vector<bool> fixed(n);
    for (int i = 0; i < n; i++) fixed[i] = ctx.mat_type[i] == 1 || ctx.mat_type[i] == 2;
    o << "__device__ __forceinline__ bool se_fixed(unsigned int c) { return " << em.bit(fixed, "c") << "; }\n";
    // density ranks (equal densities share a rank; NaN densities are refused)
    {
        std::vector<float> d = ctx.dens;
        for (float x : d)
            if (std::isnan(x)) gen_error("NaN density is not supported by the HIP backend");
        std::vector<float> u = d;
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end(), [](float a, float b) { return a == b; }), u.end());
        std::vector<int> rank(n);
        for (int i = 0; i < n; i++)
            rank[i] = (int)(std::lower_bound(u.begin(), u.end(), d[i]) - u.begin());
        int bits = 1;
        while ((1 << bits) < (int)u.size()) bits++;
        if (bits == 3) bits = 4;
        if (bits > 4 && bits < 8) bits = 8;
        o << "__device__ __forceinline__ unsigned int se_rank(unsigned int c) {\n";
        if (n * bits <= 32) {
            uint64_t w = 0;
            for (int i = 0; i < n; i++) w |= (uint64_t)rank[i] << (i * bits);
            o << "    return __builtin_amdgcn_ubfe(" << hex_u32(w) << ", c * " << bits << "u, " << bits << "u);\n";
        } else if (n * bits <= 64) {
            uint64_t w = 0;
            for (int i = 0; i < n; i++) w |= (uint64_t)rank[i] << (i * bits);
            o << "    return (unsigned int)(" << hex_u64(w) << " >> (c * " << bits << "u)) & " << ((1u << bits) - 1)
              << "u;\n";
        } else {
            o << "    const unsigned char t[" << n << "] = {";
            for (int i = 0; i < n; i++) o << (i ? ", " : "") << rank[i];
            o << "};\n    return t[c];\n";
        }
        o << "}\n";
    }
    // per-material tables in constant memory (a local array indexed by a
    // runtime id would be rebuilt in scratch memory at every use)
    o << "__constant__ float se_dens_tab[" << n << "] = {";
    for (int i = 0; i < n; i++) o << (i ? ", " : "") << flit(ctx.dens[i]);
    o << "};\n__device__ __forceinline__ float se_dens(unsigned int c) { return se_dens_tab[c]; }\n";
    o << "__constant__ unsigned int se_type_tab[" << n << "] = {";
    for (int i = 0; i < n; i++) o << (i ? ", " : "") << ctx.mat_type[i];
    o << "};\n__device__ __forceinline__ unsigned int se_type(unsigned int c) { return se_type_tab[c]; }\n";
    // render side-outputs (operations.glsl:99-170): material colour and emission,
    // and isLightObstacle (operations.glsl:68-70): no RGB emission and not of type EMPTY
    for (int k = 0; k < 2; k++) {
        const auto& tab = k == 0 ? ctx.color : ctx.emission;
        const char* nm = k == 0 ? "se_color" : "se_emission";
        o << "__constant__ float4 " << nm << "_tab[" << n << "] = {";
        for (int i = 0; i < n; i++)
            o << (i ? ", " : "") << "{" << flit(tab[i][0]) << ", " << flit(tab[i][1]) << ", " << flit(tab[i][2]) << ", "
              << flit(tab[i][3]) << "}";
        o << "};\n__device__ __forceinline__ float4 " << nm << "(unsigned int c) { return " << nm << "_tab[c]; }\n";
    }
    {
        int empty_type = -1;
        auto it = ctx.types.find("EMPTY");
        if (it != ctx.types.end()) empty_type = it->second;
        std::vector<bool> obst(n);
        for (int i = 0; i < n; i++)
            obst[i] = ctx.emission[i][0] == 0.0f && ctx.emission[i][1] == 0.0f && ctx.emission[i][2] == 0.0f &&
                      ctx.mat_type[i] != empty_type;
        o << "__device__ __forceinline__ bool se_light_obstacle(unsigned int c) { return " << em.bit(obst, "c")
          << "; }\n";
    }
    o << "__device__ __forceinline__ unsigned int se_lowbias(unsigned int h) {\n"
         "    h ^= h >> 16; h *= 0x7feb352du; h ^= h >> 15; h *= 0x846ca68bu; h ^= h >> 16;\n    return h;\n}\n";
    // guarded swap on register values (copies first: a conditional over lvalues
    // would select between addresses and keep the cells out of registers)
    o << "#define SE_GSWAP(a, b) do { const unsigned int _a = (a), _b = (b); "
         "const bool _ok = !se_fixed(_a) && !se_fixed(_b); (a) = _ok ? _b : _a; (b) = _ok ? _a : _b; } while (0)\n";
    // se_block_h: simulate() for one block (falling_sand.glsl:61-118) given the
    // four hash lanes of the block; se_block computes them from (origin, frame).
    int64_t tmirror = rand_threshold(0.5f, false);
    em.thresholds[0].insert((uint32_t)tmirror);
    static const unsigned mult[4] = {213u, 2131u, 21313u, 213132u};
    // The rule chain works on local copies, written back once at the end: stores
    // through the reference parameters inside the branches would be merged into
    // stores through pointer phis, which keeps the cells in scratch memory.
    o << "__device__ __forceinline__ void se_block_h(unsigned int& r0, unsigned int& r1, unsigned int& r2, "
         "unsigned int& r3, unsigned int h0, unsigned int h1, unsigned int h2, unsigned int h3) {\n";
    o << "    (void)h1; (void)h2; (void)h3;\n";
    o << "    unsigned int c0 = r0, c1 = r1, c2 = r2, c3 = r3;\n";
    o << "    if ((c0 | c1 | c2 | c3) == 0u) return;\n";
    o << "    const bool mirror = h0 <= " << hex_u32((uint64_t)tmirror) << ";\n";
    o << "    if (mirror) { SE_GSWAP(c0, c1); SE_GSWAP(c2, c3); }\n";
    o << mirrored;
    o << "    if (mirror) {\n        SE_GSWAP(c0, c1);\n        SE_GSWAP(c2, c3);\n    } else {\n";
    if (!right.empty()) {
        std::istringstream is(right);
        std::string ln;
        while (std::getline(is, ln)) o << "    " << ln << "\n";
    }
    o << "    }\n    r0 = c0; r1 = c1; r2 = c2; r3 = c3;\n}\n";
    // x = ox*461 + oy*2131 + frame*2131*2131 (math.glsl:74-79) and lane k hashes
    // x*m_k; the products are distributed so that the per-block terms use 24-bit
    // multiplies where the constant fits (|ox|, |oy| < 2^23; -1 stays -1).
    auto hash_expr = [&](int l) {
        uint32_t A = 461u * mult[l], B = 2131u * mult[l], Cf = 4541161u * mult[l];
        auto term = [](uint32_t k, const char* v) -> std::string {
            if (k < (1u << 23)) return "(unsigned int)__mul24((int)" + std::string(v) + ", " + std::to_string(k) + ")";
            return std::string(v) + " * " + std::to_string(k) + "u";
        };
        return "se_lowbias(" + term(A, "ox") + " + " + term(B, "oy") + " + frame * " + std::to_string(Cf) + "u)";
    };
    o << "__device__ __forceinline__ unsigned int se_hash_lane(int l, unsigned int ox, unsigned int oy, "
         "unsigned int frame) {\n";
    for (int l = 0; l < 4; l++) o << "    if (l == " << l << ") return " << hash_expr(l) << ";\n";
    o << "    return 0u;\n}\n";
    // the same hash split into a column term gx*ka, a row term gy*kb and a frame
    // term frame*kf (all mod 2^32), for kernels that keep the terms across blocks
    {
        const char* nm[3] = {"ka", "kb", "kf"};
        const uint32_t base[3] = {461u, 2131u, 4541161u};
        for (int k = 0; k < 3; k++) {
            o << "__device__ __forceinline__ unsigned int se_hash_" << nm[k] << "(int l) {\n";
            for (int l = 0; l < 4; l++) o << "    if (l == " << l << ") return " << (uint32_t)(base[k] * mult[l]) << "u;\n";
            o << "    return 0u;\n}\n";
        }
        o << "__device__ __forceinline__ bool se_hash_used(int l) {\n    return l == 0";
        for (int l = 1; l < 4; l++)
            if (em.lanes_used.count(l)) o << " || l == " << l;
        o << ";\n}\n";
    }
    o << "__device__ __forceinline__ void se_block(unsigned int& r0, unsigned int& r1, unsigned int& r2, "
         "unsigned int& r3, unsigned int ox, unsigned int oy, unsigned int frame) {\n";
    o << "    if ((r0 | r1 | r2 | r3) == 0u) return;\n";
    for (int l = 0; l < 4; l++) {
        if (l > 0 && !em.lanes_used.count(l)) o << "    const unsigned int h" << l << " = 0u;\n";
        else o << "    const unsigned int h" << l << " = " << hash_expr(l) << ";\n";
    }
    o << "    se_block_h(r0, r1, r2, r3, h0, h1, h2, h3);\n}\n";
    // Transition table: the block result is a function of the four ids and, per
    // hash lane, of which threshold interval the lane falls in (every use of a
    // lane is a compare against one of its thresholds).  Enumerated on the
    // device by se_build_table from se_block_h itself; used when it fits LDS.
    {
        uint64_t R = 1;
        std::vector<std::vector<uint32_t>> th(4);
        for (int l = 0; l < 4; l++) {
            th[l].assign(em.thresholds[l].begin(), em.thresholds[l].end());
            R *= th[l].size() + 1;
        }
        uint64_t tuples = (uint64_t)n * n * n * n;
        uint64_t bytes = tuples * R * 4;
        bool ok = !em.rand_float_used && bytes <= kTableMaxBytes;
        // Coded table, for rule sets whose full table does not fit: tuples of the
        // D = n - 2 materials other than NULL and WALL (blocks holding those run
        // the chain), the mirror factored out as a byte permutation (lane 0 only
        // decides the mirror and there are no Right rules), and per tuple a u8
        // naming its outcome vector over the classes of lanes 1..3.  Whether the
        // rule set fits that form is checked on the enumerated table when the rule
        // set is loaded; it is generated without it otherwise.
        const uint64_t D = (uint64_t)n - 2, Rc = R / 2;
        const uint64_t ct_bytes = D * D * D * D, t1_words = (ct_bytes + 15) / 16 * 4;
        const bool ct = allow_ctable && !ok && !em.rand_float_used && right.empty() && th[0].size() == 1 &&
                        n >= 3 && D <= 16 && ct_bytes <= kCTableMaxBytes;
        o << "#define SE_TBL " << (ok ? 1 : ct ? 2 : 0) << "\n";
        o << "#define SE_TBL_R " << R << "u\n";
        o << "#define SE_TBL_ENTRIES " << (ok ? tuples * R : 1) << "u\n";
        // u32 words of the table staged into LDS (the coded table's outcome vectors
        // are counted when the enumerated table has been compacted)
        o << "#define SE_TBL_WORDS " << (ok ? tuples * R : ct ? t1_words + ct_x_words : 0) << "u\n";
        if (ct) {
            o << "#define SE_CT_D " << D << "u\n";
            o << "#define SE_CT_R " << Rc << "u\n";
            o << "#define SE_CT_T1_WORDS " << t1_words << "u\n";
            o << "#define SE_CT_MIRROR_T " << hex_u32((uint64_t)tmirror) << "\n";
            o << "__device__ __forceinline__ unsigned int se_rclass_ct(unsigned int h1, unsigned int h2, "
                 "unsigned int h3) {\n    (void)h1; (void)h2; (void)h3;\n    unsigned int r = 0u;\n";
            uint64_t st = 1;
            for (int l = 1; l < 4; l++) {
                for (uint32_t t : th[l]) o << "    r += (h" << l << " > " << hex_u32(t) << ") ? " << st << "u : 0u;\n";
                st *= th[l].size() + 1;
            }
            o << "    return r;\n}\n";
        }
        // class index: lane l contributes (#thresholds below h) * stride_l
        o << "__device__ __forceinline__ unsigned int se_rclass(unsigned int h0, unsigned int h1, unsigned int h2, "
             "unsigned int h3) {\n    (void)h0; (void)h1; (void)h2; (void)h3;\n    unsigned int r = 0u;\n";
        uint64_t stride = 1;
        for (int l = 0; l < 4; l++) {
            for (uint32_t t : th[l]) o << "    r += (h" << l << " > " << hex_u32(t) << ") ? " << stride << "u : 0u;\n";
            stride *= th[l].size() + 1;
        }
        o << "    return r;\n}\n";
        // representative hash value of each interval (its upper end)
        o << "__device__ __forceinline__ unsigned int se_rclass_rep(int l, unsigned int idx) {\n";
        for (int l = 0; l < 4; l++) {
            o << "    if (l == " << l << ") {\n";
            for (size_t j = 0; j < th[l].size(); j++)
                o << "        if (idx == " << j << "u) return " << hex_u32(th[l][j]) << ";\n";
            o << "        return 0xffffffffu;\n    }\n";
        }
        o << "    return 0u;\n}\n";
        o << "__device__ __forceinline__ unsigned int se_rclass_count(int l) {\n";
        for (int l = 0; l < 4; l++) o << "    if (l == " << l << ") return " << th[l].size() + 1 << "u;\n";
        o << "    return 1u;\n}\n";
        if (info) {
            info->table = ok;
            info->table_R = (int)R;
            info->ctable = ct;
            info->ct_D = ct ? (int)D : 0;
            info->ct_R = ct ? (int)Rc : 0;
        }
    }
    if (info) {
        info->nmat = n;
        info->nrules = nrules;
        info->rand_lanes.assign(em.lanes_used.begin(), em.lanes_used.end());
    }
    return o.str();
}

std::string lds_kernel_name(const LdsVariant& v) {
    return "se_step_lds_" + std::to_string(v.tw) + "x" + std::to_string(v.th) + "_h" + std::to_string(v.hb) + "_t" +
           std::to_string(v.nt);
}

std::string hip_program_source(const ParsingResult& r, const std::vector<LdsVariant>& lds, bool allow_ctable,
                               uint64_t ct_x_words) {
    std::string s = hip_rules_source(r, nullptr, allow_ctable, ct_x_words);
    s += "\n";
    s += kKernelTemplates;
    s += "\n";
    for (auto& v : lds)
        s += "SE_LDS_KERNEL(" + std::to_string(v.tw) + ", " + std::to_string(v.th) + ", " + std::to_string(v.hb) + ", " +
             std::to_string(v.nt) + ")\n";
    return s;
}

}  // namespace se
