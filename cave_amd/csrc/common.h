// common.h -- shared host/device definitions of the CAVE MI355X engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/cave.h"

namespace cave {

// ---- error reporting (thread-local last error, no exceptions across the C-ABI) ----
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define CAVE_HIP(expr)                                                                     \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess)                                                              \
            return ::cave::fail(CAVE_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

constexpr uint64_t kPermMask = (uint64_t(1) << 60) - 1;

// Rule in the form every kernel consumes.  bits: bit v of born_lut / surv_lut
// is set when a cell with v live neighbours is born / survives, i.e. the
// reference's "n > 0 && mask >> (n-1) & 1" (searchSimulationCube.slang:181-196)
// tabulated over n = 0..31 (bit 0 is always clear).
struct DevRule {
    uint32_t M;          // live state, 1..16 (16: no representable live state)
    uint32_t flags;      // F=1 E=2 C=4
    uint32_t born_lut;
    uint32_t surv_lut;
};

inline DevRule make_rule(uint32_t M, uint32_t flags, uint64_t birth, uint64_t survival)
{
    DevRule r{};
    r.M = M;
    r.flags = flags & 7u;
    // counts can reach 26; bits 26..59 of a mask can never match and are dropped
    r.born_lut = uint32_t((birth & 0x3FFFFFFull) << 1);
    r.surv_lut = uint32_t((survival & 0x3FFFFFFull) << 1);
    return r;
}

// Per-block partial statistics of one tick.
struct StatPart {
    uint32_t dead, dying, born, surv;
    int32_t bmin[3];
    int32_t bmax[3];
};

// Reference packed layout helpers (searchSimulationCube.slang:52-53).
inline uint32_t ref_words_per_row(uint32_t X) { return (X + 7u) / 8u; }

}  // namespace cave
