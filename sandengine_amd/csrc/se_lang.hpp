// sandengine rule language: YAML -> IR -> GLSL text (and the IR the HIP
// backend consumes).  Behaviour follows sandengine-lang:
//   parse_string        sandengine-lang/src/parser.rs:93-152
//   SandRule & parsing  sandengine-lang/src/parser/rules.rs:25-420
//   SandType & parsing  sandengine-lang/src/parser/types.rs:12-209
//   SandMaterial        sandengine-lang/src/parser/materials.rs:11-201
//   GLSL emission       sandengine-lang/src/lib.rs:17-147
// Errors carry the reference's categories (parser.rs:44-73): MissingField,
// InvalidType, NotFound, NotRecognized; YAML syntax errors are "Yaml".
#pragma once
#include <cstdint>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

namespace se {

struct LangError : std::runtime_error {
    std::string kind;
    LangError(std::string k, const std::string& m) : std::runtime_error(m), kind(std::move(k)) {}
};

enum class RuleType { Mirrored, Left, Right };

struct SandRule {
    std::string name;
    RuleType ruletype = RuleType::Mirrored;
    std::vector<std::string> if_conds;
    std::vector<std::string> do_actions;
    std::vector<float> probabilities;
    bool mirror = true;
    std::optional<std::string> precondition;
    bool used = false;
};

struct SandType {
    size_t id = 0;
    std::string name;
    std::string inherits;
    std::vector<std::string> children;
    std::vector<std::string> base_rules;
};

struct SandMaterial {
    size_t id = 0;
    std::string name;
    std::string mattype;
    float color[4] = {0, 0, 0, 0};
    float emission[4] = {0, 0, 0, 0};
    bool selectable = true;
    float density = 0;
    std::vector<std::string> extra_rules;
};

struct ParsingResult {
    std::vector<SandRule> rules;
    std::vector<SandType> types;
    std::vector<SandMaterial> materials;
};

ParsingResult parse_string(const std::string& yaml);

// Rust `{}` Display of an f32 (shortest round-trip digits, positional).
std::string fmt_f32(float v);

// (materials.glsl, rules.glsl) exactly as the reference writes them.
std::pair<std::string, std::string> create_glsl(const ParsingResult& r);

// JSON dump of the IR (for tests and tooling).
std::string ir_json(const ParsingResult& r);

}  // namespace se
