// YAML subset reader for sandengine rule files.
//
// Covers what the rule language uses (sandengine-lang reads its input with
// serde_yaml 0.9, sandengine-lang/src/parser.rs:95): block mappings and
// sequences, flow sequences/mappings (also spanning lines), plain, single- and
// double-quoted scalars, comments, `---`.  Mapping order is preserved because
// it fixes material and type ids.  Scalars keep their text; typing follows the
// YAML 1.2 core schema on access, as serde_yaml does.
#pragma once
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace se {

struct YamlError : std::runtime_error {
    explicit YamlError(const std::string& m) : std::runtime_error(m) {}
};

struct YNode {
    enum Kind { Null, Scalar, Seq, Map } kind = Null;
    std::string text;      // scalar text
    bool quoted = false;   // quoted scalars are always strings
    std::vector<YNode> seq;
    std::vector<std::pair<std::string, YNode>> map;  // keys are scalar texts
    std::vector<bool> key_is_str;                    // key typed as a string?
    std::vector<std::string> key_dbg;                // key rendered like a serde_yaml Value ({:?})
    int line = 0;

    bool is_null() const;
    // serde_yaml Value accessors: return false when the value has another type.
    bool as_str(std::string* out) const;
    bool as_bool(bool* out) const;
    bool as_f64(double* out) const;   // ints convert too
    bool as_u64(uint64_t* out) const;
    bool is_int() const;
    bool is_float() const;
    const YNode* get(const std::string& key) const;  // mapping lookup, else nullptr
    std::string debug() const;                       // Rust {:?}-like rendering
};

YNode yaml_parse(const std::string& text);

}  // namespace se
