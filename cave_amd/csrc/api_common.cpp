// api_common.cpp -- library-wide C-ABI entry points: errors, device info and
// the host helpers (rule strings, bitmask enumeration, synthetic RNG).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <sstream>
#include <string>
#include <vector>

#include "common.h"
#include "host_util.h"

namespace cave {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

// src/common/permutationMath.h:101-128 -- sorted bitmasks with 1..K bits of N.
std::vector<uint64_t> enumerate_valid_bitmasks(uint32_t n, uint32_t k)
{
    std::vector<uint64_t> out;
    if (k == 0 || n == 0) return out;
    const uint32_t kk = std::min(k, n);
    // iterative combination walk per popcount p, then one sort
    for (uint32_t p = 1; p <= kk; p++) {
        std::vector<uint32_t> idx(p);
        for (uint32_t i = 0; i < p; i++) idx[i] = i;
        while (true) {
            uint64_t m = 0;
            for (uint32_t i : idx) m |= uint64_t(1) << i;
            out.push_back(m);
            int i = (int)p - 1;
            while (i >= 0 && idx[i] == n - p + (uint32_t)i) i--;
            if (i < 0) break;
            idx[i]++;
            for (uint32_t j = (uint32_t)i + 1; j < p; j++) idx[j] = idx[j - 1] + 1;
        }
    }
    std::sort(out.begin(), out.end());
    return out;
}

uint32_t neighbor_count(uint32_t flags)
{
    return ((flags & 1u) ? 6u : 0u) + ((flags & 2u) ? 12u : 0u) + ((flags & 4u) ? 8u : 0u);
}

// src/vulkan/systems/guiSystem.cpp:133-163
uint64_t rule_string_to_bitmask(const std::string &rule)
{
    uint64_t mask = 0;
    std::stringstream ss(rule);
    std::string tok;
    while (std::getline(ss, tok, ',')) {
        tok.erase(std::remove_if(tok.begin(), tok.end(), [](unsigned char c) { return std::isspace(c); }),
                  tok.end());
        if (tok.empty()) continue;
        const size_t dash = tok.find('-');
        if (dash != std::string::npos) {
            const int a = std::atoi(tok.substr(0, dash).c_str());
            const int b = std::atoi(tok.substr(dash + 1).c_str());
            for (int i = a; i <= b && i <= 60; i++)
                if (i >= 1) mask |= uint64_t(1) << (i - 1);
        } else {
            const int v = std::atoi(tok.c_str());
            if (v >= 1 && v <= 60) mask |= uint64_t(1) << (v - 1);
        }
    }
    return mask;
}

// src/vulkan/systems/guiSystem.cpp:92-131
std::string bitmask_to_rule_string(uint64_t mask)
{
    std::vector<int> bits;
    for (int b = 0; b < 60; b++)
        if ((mask >> b) & 1u) bits.push_back(b + 1);
    if (bits.empty()) return "-";
    std::string out;
    size_t i = 0;
    while (i < bits.size()) {
        const size_t start = i;
        while (i + 1 < bits.size() && bits[i + 1] == bits[i] + 1) i++;
        if (!out.empty()) out += ",";
        if (i - start >= 2) {
            out += std::to_string(bits[start]) + "-" + std::to_string(bits[i]);
        } else {
            for (size_t j = start; j <= i; j++) {
                if (j > start) out += ",";
                out += std::to_string(bits[j]);
            }
        }
        i++;
    }
    return out;
}

uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double uniform01(uint64_t &s) { return double(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

uint32_t density_threshold(double density)
{   // searchSystem.cpp:260-262, evaluated in double
    return density >= 1.0 ? 0xFFFFFFFFu : (uint32_t)(density * 4294967296.0);
}

uint32_t render_density_threshold(double density, bool *all_alive)
{   // the density arrives as the std::to_string ("%f") text of a float
    // (simulation.h:19, simulationRenderer.cpp:201); cellInit.slang:57-58
    char text[64];
    std::snprintf(text, sizeof text, "%f", (double)(float)density);
    const float d = std::strtof(text, nullptr);
    *all_alive = d >= 1.0f;
    if (*all_alive) return 0u;
    const float t = d * 4294967296.0f;
    return t <= 0.0f ? 0u : (uint32_t)t;
}

DevInit make_init(const uint32_t dims[3], const uint32_t spawn[3], uint32_t spawn_mode, double density,
                  uint32_t seed)
{   // searchSystem.cpp:264-279: box [(G-S)/2, (G+S)/2) per axis, S clamped to G
    DevInit in{};
    for (int a = 0; a < 3; a++) {
        const uint32_t s = std::min(spawn[a], dims[a]);
        in.box_lo[a] = (dims[a] - s) / 2;
        in.box_hi[a] = (dims[a] + s) / 2;
    }
    in.random = spawn_mode == CAVE_SPAWN_RANDOM || spawn_mode == CAVE_SPAWN_RENDER_RANDOM ? 1u : 0u;
    if (spawn_mode == CAVE_SPAWN_RENDER_RANDOM) {
        // cellInit.slang:44-60: the search hash with seed 0, a float threshold
        bool all = false;
        in.threshold = render_density_threshold(density, &all);
        in.all_in = all ? 1u : 0u;
        in.seed = 0;
        return in;
    }
    in.all_in = density >= 1.0 ? 1u : 0u;
    in.threshold = density_threshold(density);
    in.seed = seed;
    return in;
}

}  // namespace cave

using namespace cave;

extern "C" {

int cave_abi_version(void) { return CAVE_ABI_VERSION; }

const char *cave_last_error(void) { return g_last_error.c_str(); }

int cave_device_info(char *name, size_t cap, int *cu_count, uint64_t *hbm_bytes)
{
    hipDeviceProp_t prop;
    CAVE_HIP(hipGetDeviceProperties(&prop, 0));
    if (name && cap) {
        std::snprintf(name, cap, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (cu_count) *cu_count = prop.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = prop.totalGlobalMem;
    return CAVE_OK;
}

uint64_t cave_rule_string_to_bitmask(const char *rule) { return rule ? rule_string_to_bitmask(rule) : 0; }

size_t cave_bitmask_to_rule_string(uint64_t bitmask, char *out, size_t cap)
{
    const std::string s = bitmask_to_rule_string(bitmask);
    if (out && cap) {
        const size_t n = std::min(cap - 1, s.size());
        std::memcpy(out, s.data(), n);
        out[n] = 0;
    }
    return s.size();
}

size_t cave_enumerate_valid_bitmasks(uint32_t neighbor_count, uint32_t max_bits, uint64_t *out, size_t cap)
{
    const std::vector<uint64_t> v = enumerate_valid_bitmasks(neighbor_count, max_bits);
    if (out) std::copy_n(v.begin(), std::min(cap, v.size()), out);
    return v.size();
}

size_t cave_cell_words(uint32_t X, uint32_t Y, uint32_t Z) { return (size_t)((X + 7u) / 8u) * Y * Z; }

uint64_t cave_splitmix64(uint64_t *state) { return state ? splitmix64(*state) : 0; }

uint64_t cave_random_rule(uint64_t *state, uint32_t n_counts, double p)
{
    uint64_t m = 0;
    if (!state) return 0;
    for (uint32_t b = 0; b < n_counts && b < 60; b++)
        if (uniform01(*state) < p) m |= uint64_t(1) << b;
    return m;
}

}  // extern "C"

// ---- render-path storage (cellStateUtils.slang:18-46) ----
namespace {

uint32_t spread3(uint32_t v)
{   // bits b0 b1 b2 -> positions 0, 3, 6
    return (v & 1u) | ((v & 2u) << 2) | ((v & 4u) << 4);
}

struct RenderGeom {
    uint32_t iw, nbx, nby, nbz;
    RenderGeom(uint32_t X, uint32_t Y, uint32_t Z)
        : iw((X + 1u) / 2u), nbx((iw + 7u) / 8u), nby((Y + 7u) / 8u), nbz((Z + 7u) / 8u) {}
    size_t tiled_index(uint32_t tx, uint32_t y, uint32_t z) const
    {
        const size_t block = ((size_t)(z >> 3) * nby + (y >> 3)) * nbx + (tx >> 3);
        return block * 512u + (spread3(tx & 7u) | (spread3(y & 7u) << 1) | (spread3(z & 7u) << 2));
    }
};

}  // namespace

extern "C" size_t cave_erd_cell_words(uint32_t X, uint32_t Y, uint32_t Z)
{
    return ((size_t)X * Y * Z + 7u) / 8u;
}

extern "C" size_t cave_render_bytes(uint32_t layout, uint32_t X, uint32_t Y, uint32_t Z)
{
    const RenderGeom r(X, Y, Z);
    if (layout == CAVE_RENDER_IMAGE) return (size_t)r.iw * Y * Z;
    if (layout == CAVE_RENDER_TILED) return (size_t)r.nbx * r.nby * r.nbz * 512u * 4u;
    return 0;
}

extern "C" int cave_render_to_packed(uint32_t layout, const void *render, uint32_t X, uint32_t Y, uint32_t Z,
                                     uint32_t *packed)
{
    if (!render || !packed || !X || !Y || !Z) return fail(CAVE_E_INVALID, "null argument or empty grid");
    if (layout > CAVE_RENDER_TILED) return fail(CAVE_E_INVALID, "bad render layout");
    const RenderGeom r(X, Y, Z);
    const uint32_t upr = (X + 7u) / 8u;
    std::memset(packed, 0, sizeof(uint32_t) * (size_t)upr * Y * Z);
    const uint8_t *img = static_cast<const uint8_t *>(render);
    const uint32_t *buf = static_cast<const uint32_t *>(render);
    for (uint32_t z = 0; z < Z; z++)
        for (uint32_t y = 0; y < Y; y++)
            for (uint32_t tx = 0; tx < r.iw; tx++) {
                const uint32_t byte = layout == CAVE_RENDER_IMAGE ? img[((size_t)z * Y + y) * r.iw + tx]
                                                                  : buf[r.tiled_index(tx, y, z)] & 0xFFu;
                for (uint32_t k = 0; k < 2; k++) {
                    const uint32_t x = 2u * tx + k;
                    if (x >= X) break;
                    packed[(size_t)(z * Y + y) * upr + x / 8u] |= ((byte >> (4u * k)) & 15u) << (4u * (x % 8u));
                }
            }
    return CAVE_OK;
}

extern "C" int cave_packed_to_render(uint32_t layout, const uint32_t *packed, uint32_t X, uint32_t Y, uint32_t Z,
                                     void *render)
{
    if (!render || !packed || !X || !Y || !Z) return fail(CAVE_E_INVALID, "null argument or empty grid");
    if (layout > CAVE_RENDER_TILED) return fail(CAVE_E_INVALID, "bad render layout");
    const RenderGeom r(X, Y, Z);
    const uint32_t upr = (X + 7u) / 8u;
    std::memset(render, 0, cave_render_bytes(layout, X, Y, Z));
    uint8_t *img = static_cast<uint8_t *>(render);
    uint32_t *buf = static_cast<uint32_t *>(render);
    for (uint32_t z = 0; z < Z; z++)
        for (uint32_t y = 0; y < Y; y++)
            for (uint32_t tx = 0; tx < r.iw; tx++) {
                uint32_t byte = 0;
                for (uint32_t k = 0; k < 2; k++) {
                    const uint32_t x = 2u * tx + k;
                    if (x >= X) break;
                    byte |= ((packed[(size_t)(z * Y + y) * upr + x / 8u] >> (4u * (x % 8u))) & 15u) << (4u * k);
                }
                if (layout == CAVE_RENDER_IMAGE) img[((size_t)z * Y + y) * r.iw + tx] = (uint8_t)byte;
                else buf[r.tiled_index(tx, y, z)] = byte;
            }
    return CAVE_OK;
}
