#include "se_yaml.hpp"

#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <regex>

namespace se {

namespace {

struct Line {
    int indent;
    std::string s;  // content without indentation and comment, right-trimmed
    int no;
};

std::string rtrim(const std::string& s) {
    size_t e = s.size();
    while (e > 0 && (s[e - 1] == ' ' || s[e - 1] == '\t' || s[e - 1] == '\r')) e--;
    return s.substr(0, e);
}
std::string ltrim(const std::string& s) {
    size_t b = 0;
    while (b < s.size() && (s[b] == ' ' || s[b] == '\t')) b++;
    return s.substr(b);
}
std::string trim(const std::string& s) { return ltrim(rtrim(s)); }

[[noreturn]] void fail(int line, const std::string& m) {
    throw YamlError("YAML error at line " + std::to_string(line) + ": " + m);
}

// Strips a trailing comment: '#' at the start or after whitespace, outside quotes.
std::string strip_comment(const std::string& s) {
    bool sq = false, dq = false;
    for (size_t i = 0; i < s.size(); i++) {
        char c = s[i];
        if (dq) {
            if (c == '\\') i++;
            else if (c == '"') dq = false;
        } else if (sq) {
            if (c == '\'') {
                if (i + 1 < s.size() && s[i + 1] == '\'') i++;
                else sq = false;
            }
        } else if (c == '"') {
            dq = true;
        } else if (c == '\'') {
            if (i == 0 || s[i - 1] == ' ' || s[i - 1] == '[' || s[i - 1] == '{' || s[i - 1] == ',' ||
                s[i - 1] == ':' || s[i - 1] == '-')
                sq = true;
        } else if (c == '#' && (i == 0 || s[i - 1] == ' ' || s[i - 1] == '\t')) {
            return s.substr(0, i);
        }
    }
    return s;
}

// Position of the ':' that ends a mapping key on this line, or npos.
size_t key_colon(const std::string& s) {
    if (s.empty()) return std::string::npos;
    if (s[0] == '[' || s[0] == '{') return std::string::npos;
    size_t i = 0;
    if (s[0] == '"' || s[0] == '\'') {
        char q = s[0];
        for (i = 1; i < s.size(); i++) {
            if (q == '"' && s[i] == '\\') { i++; continue; }
            if (s[i] == q) {
                if (q == '\'' && i + 1 < s.size() && s[i + 1] == '\'') { i++; continue; }
                break;
            }
        }
        size_t j = i + 1;
        while (j < s.size() && s[j] == ' ') j++;
        if (j < s.size() && s[j] == ':' && (j + 1 == s.size() || s[j + 1] == ' ')) return j;
        return std::string::npos;
    }
    for (i = 0; i < s.size(); i++)
        if (s[i] == ':' && (i + 1 == s.size() || s[i + 1] == ' ' || s[i + 1] == '\t')) return i;
    return std::string::npos;
}

std::string unquote(const std::string& s, int line) {
    std::string out;
    if (s.size() < 2 || s.back() != s[0]) fail(line, "unterminated quoted scalar");
    char q = s[0];
    for (size_t i = 1; i + 1 < s.size(); i++) {
        char c = s[i];
        if (q == '\'') {
            if (c == '\'' && i + 2 < s.size() && s[i + 1] == '\'') { out += '\''; i++; }
            else out += c;
        } else {
            if (c == '\\' && i + 2 < s.size()) {
                char e = s[++i];
                switch (e) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case '\\': out += '\\'; break;
                    case '"': out += '"'; break;
                    case '0': out += '\0'; break;
                    default: out += '\\'; out += e;
                }
            } else {
                out += c;
            }
        }
    }
    return out;
}

YNode scalar(const std::string& raw, int line) {
    YNode n;
    n.line = line;
    std::string t = trim(raw);
    if (!t.empty() && (t[0] == '"' || t[0] == '\'')) {
        n.kind = YNode::Scalar;
        n.quoted = true;
        n.text = unquote(t, line);
        return n;
    }
    if (t.empty() || t == "~" || t == "null" || t == "Null" || t == "NULL") {
        n.kind = YNode::Null;
        return n;
    }
    n.kind = YNode::Scalar;
    n.text = t;
    return n;
}

class Parser {
public:
    explicit Parser(const std::string& text) {
        size_t p = 0;
        int no = 0;
        while (p <= text.size()) {
            size_t e = text.find('\n', p);
            if (e == std::string::npos) e = text.size();
            std::string raw = text.substr(p, e - p);
            no++;
            p = e + 1;
            int ind = 0;
            while (ind < (int)raw.size() && raw[ind] == ' ') ind++;
            if (ind < (int)raw.size() && raw[ind] == '\t') {
                std::string rest = trim(raw);
                if (!rest.empty() && rest[0] != '#') fail(no, "tab in indentation");
            }
            std::string s = rtrim(strip_comment(raw.substr(ind)));
            if (s.empty()) continue;
            if (ind == 0 && (s == "---" || s == "...")) continue;
            lines_.push_back({ind, s, no});
            if (e == text.size()) break;
        }
    }

    YNode document() {
        if (lines_.empty()) return YNode();
        YNode n = block(lines_[0].indent);
        if (i_ < lines_.size()) fail(lines_[i_].no, "unexpected content");
        return n;
    }

private:
    std::vector<Line> lines_;
    size_t i_ = 0;

    static bool is_dash(const std::string& s) { return s == "-" || (s.size() > 1 && s[0] == '-' && s[1] == ' '); }

    YNode block(int ind) {
        const Line& L = lines_[i_];
        if (is_dash(L.s)) return seq(ind);
        if (key_colon(L.s) != std::string::npos) return map(ind);
        return inline_value(L.s, L.no, ind, true);
    }

    YNode map(int ind) {
        YNode n;
        n.kind = YNode::Map;
        n.line = lines_[i_].no;
        while (i_ < lines_.size() && lines_[i_].indent == ind) {
            const Line L = lines_[i_];
            if (is_dash(L.s)) break;
            size_t c = key_colon(L.s);
            if (c == std::string::npos) fail(L.no, "expected a mapping key");
            std::string kraw = trim(L.s.substr(0, c));
            YNode k = scalar(kraw, L.no);
            bool kstr = k.as_str(nullptr);
            std::string key = k.kind == YNode::Null ? std::string("~") : k.text;
            for (auto& kv : n.map)
                if (kv.first == key) fail(L.no, "duplicate entry with key \"" + key + "\"");
            std::string rest = ltrim(L.s.substr(c + 1));
            i_++;
            YNode v;
            if (rest.empty()) {
                if (i_ < lines_.size() && lines_[i_].indent > ind) v = block(lines_[i_].indent);
                else if (i_ < lines_.size() && lines_[i_].indent == ind && is_dash(lines_[i_].s)) v = seq(ind);
                else { v.kind = YNode::Null; v.line = L.no; }
            } else {
                v = inline_value(rest, L.no, ind, false);
            }
            n.map.emplace_back(key, std::move(v));
            n.key_is_str.push_back(kstr);
            n.key_dbg.push_back(k.debug());
        }
        if (i_ < lines_.size() && lines_[i_].indent > ind) fail(lines_[i_].no, "bad indentation");
        return n;
    }

    YNode seq(int ind) {
        YNode n;
        n.kind = YNode::Seq;
        n.line = lines_[i_].no;
        while (i_ < lines_.size() && lines_[i_].indent == ind && is_dash(lines_[i_].s)) {
            Line L = lines_[i_];
            std::string rest = L.s.size() > 1 ? L.s.substr(2) : std::string();
            int extra = 2;
            while (!rest.empty() && rest[0] == ' ') { rest.erase(0, 1); extra++; }
            if (rest.empty()) {
                i_++;
                if (i_ < lines_.size() && lines_[i_].indent > ind) n.seq.push_back(block(lines_[i_].indent));
                else { YNode z; z.line = L.no; n.seq.push_back(z); }
                continue;
            }
            if (is_dash(rest) || key_colon(rest) != std::string::npos) {
                lines_[i_] = Line{ind + extra, rest, L.no};
                n.seq.push_back(block(ind + extra));
                continue;
            }
            i_++;
            n.seq.push_back(inline_value(rest, L.no, ind, false));
        }
        return n;
    }

    // A value written on the current line (already consumed unless `at_line`).
    YNode inline_value(const std::string& rest, int no, int ind, bool at_line) {
        if (at_line) i_++;
        if (rest[0] == '[' || rest[0] == '{') {
            std::string buf = rest;
            while (!balanced(buf)) {
                if (i_ >= lines_.size()) fail(no, "unterminated flow collection");
                buf += " " + lines_[i_].s;
                i_++;
            }
            size_t p = 0;
            YNode v = flow(buf, p, no);
            while (p < buf.size() && buf[p] == ' ') p++;
            if (p != buf.size()) fail(no, "trailing characters after flow collection");
            return v;
        }
        if (rest[0] == '"' || rest[0] == '\'') return scalar(rest, no);
        std::string s = rest;
        while (i_ < lines_.size() && lines_[i_].indent > ind && key_colon(lines_[i_].s) == std::string::npos &&
               !is_dash(lines_[i_].s)) {
            s += " " + lines_[i_].s;  // folded plain continuation line
            i_++;
        }
        return scalar(s, no);
    }

    static bool balanced(const std::string& s) {
        int d = 0;
        bool sq = false, dq = false;
        for (size_t i = 0; i < s.size(); i++) {
            char c = s[i];
            if (dq) { if (c == '\\') i++; else if (c == '"') dq = false; continue; }
            if (sq) { if (c == '\'') sq = false; continue; }
            if (c == '"') dq = true;
            else if (c == '\'') sq = true;
            else if (c == '[' || c == '{') d++;
            else if (c == ']' || c == '}') d--;
        }
        return d <= 0;
    }

    static void skipws(const std::string& s, size_t& p) {
        while (p < s.size() && (s[p] == ' ' || s[p] == '\t')) p++;
    }

    YNode flow(const std::string& s, size_t& p, int no) {
        skipws(s, p);
        if (p >= s.size()) fail(no, "unexpected end of flow collection");
        if (s[p] == '[') {
            YNode n;
            n.kind = YNode::Seq;
            n.line = no;
            p++;
            for (;;) {
                skipws(s, p);
                if (p >= s.size()) fail(no, "unterminated [");
                if (s[p] == ']') { p++; break; }
                n.seq.push_back(flow(s, p, no));
                skipws(s, p);
                if (p < s.size() && s[p] == ',') { p++; continue; }
                if (p < s.size() && s[p] == ']') { p++; break; }
                fail(no, "expected , or ] in flow sequence");
            }
            return n;
        }
        if (s[p] == '{') {
            YNode n;
            n.kind = YNode::Map;
            n.line = no;
            p++;
            for (;;) {
                skipws(s, p);
                if (p >= s.size()) fail(no, "unterminated {");
                if (s[p] == '}') { p++; break; }
                YNode k = flow_scalar(s, p, no, true);
                skipws(s, p);
                YNode v;
                if (p < s.size() && s[p] == ':') {
                    p++;
                    skipws(s, p);
                    if (p < s.size() && (s[p] == ',' || s[p] == '}')) v.kind = YNode::Null;
                    else v = flow(s, p, no);
                }
                std::string key = k.kind == YNode::Null ? std::string("~") : k.text;
                for (auto& kv : n.map)
                    if (kv.first == key) fail(no, "duplicate entry with key \"" + key + "\"");
                n.map.emplace_back(key, std::move(v));
                n.key_is_str.push_back(k.as_str(nullptr));
                n.key_dbg.push_back(k.debug());
                skipws(s, p);
                if (p < s.size() && s[p] == ',') { p++; continue; }
                if (p < s.size() && s[p] == '}') { p++; break; }
                fail(no, "expected , or } in flow mapping");
            }
            return n;
        }
        return flow_scalar(s, p, no, false);
    }

    YNode flow_scalar(const std::string& s, size_t& p, int no, bool is_key) {
        skipws(s, p);
        if (p < s.size() && (s[p] == '"' || s[p] == '\'')) {
            char q = s[p];
            size_t b = p++;
            while (p < s.size()) {
                if (q == '"' && s[p] == '\\') { p += 2; continue; }
                if (s[p] == q) {
                    if (q == '\'' && p + 1 < s.size() && s[p + 1] == '\'') { p += 2; continue; }
                    break;
                }
                p++;
            }
            if (p >= s.size()) fail(no, "unterminated quoted scalar");
            p++;
            return scalar(s.substr(b, p - b), no);
        }
        size_t b = p;
        while (p < s.size()) {
            char c = s[p];
            if (c == ',' || c == ']' || c == '}') break;
            if (is_key && c == ':' && (p + 1 == s.size() || s[p + 1] == ' ')) break;
            if (c == ':' && p + 1 < s.size() && s[p + 1] == ' ' && is_key) break;
            p++;
        }
        return scalar(s.substr(b, p - b), no);
    }
};

const std::regex& int_re() {
    static const std::regex r("^[-+]?[0-9]+$|^0o[0-7]+$|^0x[0-9a-fA-F]+$");
    return r;
}
const std::regex& float_re() {
    static const std::regex r(
        "^[-+]?(\\.[0-9]+|[0-9]+(\\.[0-9]*)?)([eE][-+]?[0-9]+)?$|^[-+]?\\.(inf|Inf|INF)$|^\\.(nan|NaN|NAN)$");
    return r;
}

}  // namespace

bool YNode::is_null() const { return kind == Null; }

bool YNode::is_int() const {
    return kind == Scalar && !quoted && std::regex_match(text, int_re());
}
bool YNode::is_float() const {
    return kind == Scalar && !quoted && !is_int() && std::regex_match(text, float_re());
}

bool YNode::as_str(std::string* out) const {
    if (kind != Scalar) return false;
    if (!quoted) {
        if (is_int() || is_float()) return false;
        if (text == "true" || text == "True" || text == "TRUE" || text == "false" || text == "False" ||
            text == "FALSE")
            return false;
    }
    if (out) *out = text;
    return true;
}

bool YNode::as_bool(bool* out) const {
    if (kind != Scalar || quoted) return false;
    if (text == "true" || text == "True" || text == "TRUE") { if (out) *out = true; return true; }
    if (text == "false" || text == "False" || text == "FALSE") { if (out) *out = false; return true; }
    return false;
}

bool YNode::as_f64(double* out) const {
    if (kind != Scalar || quoted) return false;
    if (is_int()) {
        if (text.rfind("0o", 0) == 0) { if (out) *out = (double)std::strtoull(text.c_str() + 2, nullptr, 8); return true; }
        if (text.rfind("0x", 0) == 0) { if (out) *out = (double)std::strtoull(text.c_str() + 2, nullptr, 16); return true; }
        if (out) *out = std::strtod(text.c_str(), nullptr);
        return true;
    }
    if (is_float()) {
        std::string t = text;
        bool neg = !t.empty() && t[0] == '-';
        std::string u = (!t.empty() && (t[0] == '-' || t[0] == '+')) ? t.substr(1) : t;
        if (u == ".inf" || u == ".Inf" || u == ".INF") { if (out) *out = neg ? -INFINITY : INFINITY; return true; }
        if (u == ".nan" || u == ".NaN" || u == ".NAN") { if (out) *out = NAN; return true; }
        if (out) *out = std::strtod(text.c_str(), nullptr);
        return true;
    }
    return false;
}

bool YNode::as_u64(uint64_t* out) const {
    if (kind != Scalar || quoted || !is_int()) return false;
    if (text[0] == '-') return false;
    errno = 0;
    uint64_t v;
    if (text.rfind("0o", 0) == 0) v = std::strtoull(text.c_str() + 2, nullptr, 8);
    else if (text.rfind("0x", 0) == 0) v = std::strtoull(text.c_str() + 2, nullptr, 16);
    else v = std::strtoull(text.c_str() + (text[0] == '+' ? 1 : 0), nullptr, 10);
    if (errno) return false;
    if (out) *out = v;
    return true;
}

const YNode* YNode::get(const std::string& key) const {
    if (kind != Map) return nullptr;
    for (size_t i = 0; i < map.size(); i++)
        if (map[i].first == key && key_is_str[i]) return &map[i].second;
    return nullptr;
}

std::string YNode::debug() const {
    switch (kind) {
        case Null: return "Null";
        case Scalar: {
            bool b;
            if (as_bool(&b)) return std::string("Bool(") + (b ? "true" : "false") + ")";
            if (is_int() || is_float()) return "Number(" + text + ")";
            return "String(\"" + text + "\")";
        }
        case Seq: return "Sequence";
        case Map: return "Mapping";
    }
    return "?";
}

YNode yaml_parse(const std::string& text) {
    Parser p(text);
    return p.document();
}

}  // namespace se
