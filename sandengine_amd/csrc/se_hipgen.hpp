// HIP backend of the rule language: IR -> CDNA4 kernel source for hiprtc.
//
// The rule chain the GLSL backend expresses as rule_* functions
// (sandengine-lang/src/parser/rules.rs:47-97, lib.rs:78-147) becomes one
// straight-line device function se_block() over four u32 material ids:
//   * conditions are parsed from the same GLSL text the reference emits
//     (including the unparenthesised "rand.y <= p && " prefix, rules.rs:55-62);
//   * every sub-expression that depends on one cell only is folded at code
//     generation time into a material bitmask test;
//   * density comparisons between cells become comparisons of density ranks;
//   * rand comparisons against constants become integer thresholds on the
//     hash lane (exact: rand = RNE_f32(h) * 2^-32, math.glsl:74-79).
#pragma once
#include <array>
#include <string>
#include <vector>

#include "se_lang.hpp"

namespace se {

struct LdsVariant {
    int tw, th, hb, nt;  // tile width/height, halo (= max sub-steps), threads per workgroup
};

struct HipGenInfo {
    int nmat = 0;
    std::vector<int> rand_lanes;  // hash lanes the rule chain reads besides x
    int nrules = 0;
    bool table = false;  // transition table mode (fits LDS, thresholds only)
    int table_R = 1;     // rand classes per 4-id tuple
    bool ctable = false; // coded-table candidate (SE_TBL 2): fit checked when the rule set is loaded
    int ct_D = 0;        // materials the coded table covers (all but NULL and WALL)
    int ct_R = 0;        // classes of hash lanes 1..3
};

// Generated device code for the rule set (SE_NMAT, se_fixed, se_block, ...).
// ct_x_words: u32 words of the coded table's outcome vectors, known once the
// enumerated table has been compacted (0 for the first, enumerating build).
std::string hip_rules_source(const ParsingResult& r, HipGenInfo* info = nullptr, bool allow_ctable = true,
                             uint64_t ct_x_words = 0);

// Full translation unit: rules + kernel templates + LDS instantiations.
std::string hip_program_source(const ParsingResult& r, const std::vector<LdsVariant>& lds, bool allow_ctable = true,
                               uint64_t ct_x_words = 0);

// Name of an LDS kernel instantiation.
std::string lds_kernel_name(const LdsVariant& v);

}  // namespace se
