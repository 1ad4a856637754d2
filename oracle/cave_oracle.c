/*
 * cave_oracle.c -- CPU restatement of CAVE's search-mode step semantics.
 *
 * TEST INFRASTRUCTURE ONLY (see cave_oracle.h).  Literal per-cell code; each
 * function cites the reference file:line it restates.
 */
#include "cave_oracle.h"

#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define PERM_MASK ((((uint64_t)1) << 60) - 1)

/* ------------------------------------------------------------------ */
/* packed layout: searchSimulationCube.slang:40, 52-53, 88-92          */
/* ------------------------------------------------------------------ */
uint32_t orc_words_per_row(uint32_t X) { return (X + 7u) / 8u; }

size_t orc_cell_words(uint32_t X, uint32_t Y, uint32_t Z)
{
    return (size_t)orc_words_per_row(X) * Y * Z;
}

static size_t word_index(uint32_t X, uint32_t Y, uint32_t x, uint32_t y, uint32_t z)
{
    size_t upr = orc_words_per_row(X);
    return x / 8u + (size_t)y * upr + (size_t)z * upr * Y;
}

uint32_t orc_get_cell(const uint32_t *cells, uint32_t X, uint32_t Y, uint32_t x, uint32_t y, uint32_t z)
{
    return (cells[word_index(X, Y, x, y, z)] >> ((x % 8u) * 4u)) & 0xFu;
}

void orc_set_cell(uint32_t *cells, uint32_t X, uint32_t Y, uint32_t x, uint32_t y, uint32_t z, uint32_t v)
{
    size_t w = word_index(X, Y, x, y, z);
    uint32_t sh = (x % 8u) * 4u;
    cells[w] = (cells[w] & ~(0xFu << sh)) | ((v & 0xFu) << sh);
}

uint64_t orc_canonical_hash(const uint32_t *cells, uint32_t X, uint32_t Y, uint32_t Z)
{
    uint64_t h = 1469598103934665603ULL;
    uint64_t n = (uint64_t)X * Y * Z;
    uint32_t byte = 0;
    uint64_t i = 0;
    for (uint32_t z = 0; z < Z; z++)
        for (uint32_t y = 0; y < Y; y++)
            for (uint32_t x = 0; x < X; x++, i++) {
                uint32_t v = orc_get_cell(cells, X, Y, x, y, z);
                if ((i & 1u) == 0) {
                    byte = v;
                    if (i + 1 == n) { h ^= byte; h *= 1099511628211ULL; }
                } else {
                    byte |= v << 4;
                    h ^= byte;
                    h *= 1099511628211ULL;
                }
            }
    return h;
}

/* ------------------------------------------------------------------ */
/* init: src/vulkan/systems/searchSystem.cpp:224-306                   */
/* ------------------------------------------------------------------ */
uint32_t orc_wang_hash(uint32_t seed)
{   /* searchSystem.cpp:244-252 (Thomas Wang 32-bit integer hash) */
    seed = (seed ^ 61u) ^ (seed >> 16);
    seed *= 9u;
    seed = seed ^ (seed >> 4);
    seed *= 0x27d4eb2du;
    seed = seed ^ (seed >> 15);
    return seed;
}

uint32_t orc_cell_hash(uint32_t x, uint32_t y, uint32_t z, uint32_t seed)
{   /* searchSystem.cpp:253-259 */
    uint32_t h = orc_wang_hash(x * 73856093u + seed);
    h = orc_wang_hash(h ^ (y * 19349663u));
    h = orc_wang_hash(h ^ (z * 83492791u));
    return h;
}

uint32_t orc_density_threshold(double density)
{   /* searchSystem.cpp:260-262, evaluated in double */
    return (density >= 1.0) ? 0xFFFFFFFFu : (uint32_t)(density * 4294967296.0);
}

uint32_t orc_render_density_threshold(double density, int *all_alive)
{   /* cellInit.slang:53-60: the density is the float Simulation keeps
     * (simulation.h:19), passed as the std::to_string text of that float
     * (simulationRenderer.cpp:201), i.e. "%f"; threshold = uint(d * 2^32) in float */
    char text[64];
    snprintf(text, sizeof text, "%f", (double)(float)density);
    const float d = strtof(text, NULL);
    *all_alive = d >= 1.0f;
    if (*all_alive) return 0u;
    const float t = d * 4294967296.0f;
    return t <= 0.0f ? 0u : (uint32_t)t;
}

void orc_init_cells_render(uint32_t *cells, uint32_t X, uint32_t Y, uint32_t Z, uint32_t max_cell_state,
                           uint32_t sx, uint32_t sy, uint32_t sz, uint32_t spawn_mode, double density)
{   /* cellInit.slang:33-60, 105-144: the spawn box of simulationRenderer.cpp:118-128,
     * the search hash without a seed, the float threshold above */
    memset(cells, 0, orc_cell_words(X, Y, Z) * sizeof(uint32_t));
    int all = 0;
    const uint32_t threshold = orc_render_density_threshold(density, &all);
    for (uint32_t z = 0; z < Z; z++)
        for (uint32_t y = 0; y < Y; y++)
            for (uint32_t x = 0; x < X; x++) {
                int in_box = x >= (X - sx) / 2 && x < (X + sx) / 2 && y >= (Y - sy) / 2 && y < (Y + sy) / 2 &&
                             z >= (Z - sz) / 2 && z < (Z + sz) / 2;
                if (!in_box) continue;
                int alive = spawn_mode == 0 ? (all || orc_cell_hash(x, y, z, 0) < threshold) : 1;
                if (alive) cells[word_index(X, Y, x, y, z)] |= (max_cell_state & 0xFu) << ((x % 8u) * 4u);
            }
}

void orc_init_cells(uint32_t *cells, uint32_t X, uint32_t Y, uint32_t Z,
                    uint32_t max_cell_state, uint32_t sx, uint32_t sy, uint32_t sz,
                    uint32_t spawn_mode, double density, uint32_t seed)
{   /* searchSystem.cpp:221-306 with spawnSizeIndex = 0 (full spawn area) */
    memset(cells, 0, orc_cell_words(X, Y, Z) * sizeof(uint32_t));
    uint32_t threshold = orc_density_threshold(density);
    int random = (spawn_mode == 0);
    for (uint32_t x = 0; x < X; x++)
        for (uint32_t y = 0; y < Y; y++)
            for (uint32_t z = 0; z < Z; z++) {
                int in_box = x >= (X - sx) / 2 && x < (X + sx) / 2 &&
                             y >= (Y - sy) / 2 && y < (Y + sy) / 2 &&
                             z >= (Z - sz) / 2 && z < (Z + sz) / 2;
                if (!in_box) continue;
                int alive = random ? (density >= 1.0 || orc_cell_hash(x, y, z, seed) < threshold) : 1;
                if (alive) {
                    size_t w = word_index(X, Y, x, y, z);
                    cells[w] |= (max_cell_state & 0xFu) << ((x % 8u) * 4u);
                }
            }
}

/* ------------------------------------------------------------------ */
/* step: shaders/slang/search/searchSimulationCube.slang               */
/* ------------------------------------------------------------------ */
/* shaders/slang/common/cube.slang:7-21 */
static const int kFace[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
static const int kEdge[12][3] = {{1, 1, 0}, {1, -1, 0}, {-1, 1, 0}, {-1, -1, 0},
                                 {1, 0, 1}, {1, 0, -1}, {-1, 0, 1}, {-1, 0, -1},
                                 {0, 1, 1}, {0, 1, -1}, {0, -1, 1}, {0, -1, -1}};
static const int kCorner[8][3] = {{1, 1, 1}, {-1, 1, 1}, {1, -1, 1}, {-1, -1, 1},
                                  {1, 1, -1}, {-1, 1, -1}, {1, -1, -1}, {-1, -1, -1}};

static int wrap_coord(int c, int n) { return ((c % n) + n) % n; }

/* searchSimulationCube.slang:78-94; with wrap, rayMarchCompute.slang:80-108
 * (every axis taken modulo its size: WRAP_AT_BOUNDARY, one domain) */
static int is_alive_neighbor(const uint32_t *in, int cx, int cy, int cz, uint32_t M,
                             uint32_t X, uint32_t Y, uint32_t Z, int wrap)
{
    if (wrap) {
        cx = wrap_coord(cx, (int)X);
        cy = wrap_coord(cy, (int)Y);
        cz = wrap_coord(cz, (int)Z);
    } else if (cx < 0 || cx >= (int)X || cy < 0 || cy >= (int)Y || cz < 0 || cz >= (int)Z) {
        return 0;
    }
    return orc_get_cell(in, X, Y, (uint32_t)cx, (uint32_t)cy, (uint32_t)cz) == M;
}

/* searchSimulationCube.slang:96-115 (rayMarchCompute.slang:115-148 for cubes):
 * one term per offset, so on a wrapped grid smaller than 3 cells along an
 * axis the same cell (or the cell itself) is counted once per offset */
static int count_alive(const uint32_t *in, int x, int y, int z, uint32_t M, uint32_t X,
                       uint32_t Y, uint32_t Z, uint32_t flags, int wrap)
{
    int c = 0;
    if (flags & 1u)
        for (int i = 0; i < 6; i++)
            c += is_alive_neighbor(in, x + kFace[i][0], y + kFace[i][1], z + kFace[i][2], M, X, Y, Z, wrap);
    if (flags & 2u)
        for (int i = 0; i < 12; i++)
            c += is_alive_neighbor(in, x + kEdge[i][0], y + kEdge[i][1], z + kEdge[i][2], M, X, Y, Z, wrap);
    if (flags & 4u)
        for (int i = 0; i < 8; i++)
            c += is_alive_neighbor(in, x + kCorner[i][0], y + kCorner[i][1], z + kCorner[i][2], M, X, Y, Z, wrap);
    return c;
}

static void step_impl(const uint32_t *in, uint32_t *out, uint32_t X, uint32_t Y, uint32_t Z,
                      uint64_t word0, uint64_t word1, const uint64_t *valid_bitmasks,
                      uint32_t max_rule_bits, orc_tick_stats *st, int wrap)
{
    /* searchSimulationCube.slang:139-157 */
    uint32_t M = (uint32_t)(word0 >> 60) + 1u;
    uint32_t flags = (uint32_t)(word1 >> 60);
    uint64_t B, S;
    if (max_rule_bits > 0) {
        uint32_t bi = (uint32_t)(word0 & PERM_MASK) - 1u;
        uint32_t si = (uint32_t)(word1 & PERM_MASK) - 1u;
        B = valid_bitmasks[bi];
        S = valid_bitmasks[si];
    } else {
        B = word0 & PERM_MASK;
        S = word1 & PERM_MASK;
    }
    memset(st, 0, sizeof(*st));
    for (int a = 0; a < 3; a++) { st->bbox_min[a] = 0x7fffffff; st->bbox_max[a] = -1; }
    size_t nw = orc_cell_words(X, Y, Z);
    /* out words start at zero: padding nibbles are written as 0 (:161, :213-216) */
    memset(out, 0, nw * sizeof(uint32_t));
    for (uint32_t z = 0; z < Z; z++)
        for (uint32_t y = 0; y < Y; y++)
            for (uint32_t x = 0; x < X; x++) {
                uint32_t cur = orc_get_cell(in, X, Y, x, y, z);
                int n = count_alive(in, (int)x, (int)y, (int)z, M, X, Y, Z, flags, wrap);
                uint32_t ns;
                /* searchSimulationCube.slang:178-211 */
                if (cur == M) {
                    if (n > 0 && ((S >> (uint64_t)(n - 1)) & 1u)) { ns = M; st->snap.numberOfSurvivingCells++; }
                    else { ns = cur - 1u; st->snap.numberOfDyingCells++; }
                } else if (cur == 0u) {
                    if (n > 0 && ((B >> (uint64_t)(n - 1)) & 1u)) { ns = M; st->snap.numberOfCellsBorn++; }
                    else { ns = 0u; st->snap.numberOfDeadCells++; }
                } else {
                    ns = cur - 1u;
                    st->snap.numberOfDyingCells++;
                }
                orc_set_cell(out, X, Y, x, y, z, ns & 0xFu);
                if ((ns & 0xFu) == M && ns == M) {
                    int c[3] = {(int)x, (int)y, (int)z};
                    for (int a = 0; a < 3; a++) {
                        if (c[a] < st->bbox_min[a]) st->bbox_min[a] = c[a];
                        if (c[a] > st->bbox_max[a]) st->bbox_max[a] = c[a];
                    }
                }
            }
    st->live = st->snap.numberOfCellsBorn + st->snap.numberOfSurvivingCells;
    st->changed = st->snap.numberOfCellsBorn + st->snap.numberOfDyingCells;
}

void orc_step(const uint32_t *in, uint32_t *out, uint32_t X, uint32_t Y, uint32_t Z,
              uint64_t word0, uint64_t word1, const uint64_t *valid_bitmasks,
              uint32_t max_rule_bits, orc_snapshot *snap)
{
    orc_tick_stats st;
    step_impl(in, out, X, Y, Z, word0, word1, valid_bitmasks, max_rule_bits, &st, 0);
    *snap = st.snap;
}

void orc_step_stats(const uint32_t *in, uint32_t *out, uint32_t X, uint32_t Y, uint32_t Z,
                    uint64_t word0, uint64_t word1, const uint64_t *valid_bitmasks,
                    uint32_t max_rule_bits, orc_tick_stats *stats)
{
    step_impl(in, out, X, Y, Z, word0, word1, valid_bitmasks, max_rule_bits, stats, 0);
}

void orc_step_render(const uint32_t *in, uint32_t *out, uint32_t X, uint32_t Y, uint32_t Z,
                     uint64_t word0, uint64_t word1, orc_tick_stats *stats)
{   /* rayMarchCompute.slang:150-178, 222-281: raw bitmask rules; the W flag
     * (bit 3 of the neighbourhood nibble, simulationRenderer.cpp:595) wraps */
    step_impl(in, out, X, Y, Z, word0, word1, NULL, 0, stats, (int)((word1 >> 60) & 8u) != 0);
}

/* ------------------------------------------------------------------ */
/* ERD step: shaders/slang/search/searchSimulationERD.slang             */
/* ------------------------------------------------------------------ */
/* linear nibble layout (searchSystem.cpp:296-300): cell x + X*(y + Y*z), 8 per word */
size_t orc_erd_words(uint32_t X, uint32_t Y, uint32_t Z)
{
    return ((size_t)X * Y * Z + 7u) / 8u;
}

static uint32_t erd_get(const uint32_t *c, size_t i) { return (c[i / 8u] >> ((i % 8u) * 4u)) & 0xFu; }

/* shaders/slang/common/elongatedRhombicDodecahedron.slang:6-27 */
static const int kErdFace[2][12][3] = {
    {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1},
     {-1, -1, 0}, {-1, -1, -1}, {0, -1, -1}, {-1, 1, 0}, {-1, 1, -1}, {0, 1, -1}},
    {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1},
     {1, -1, 0}, {1, -1, 1}, {0, -1, 1}, {1, 1, 0}, {1, 1, 1}, {0, 1, 1}},
};
static const int kErdEdge[4][3] = {{1, 0, 1}, {1, 0, -1}, {-1, 0, 1}, {-1, 0, -1}};

uint32_t orc_erd_neighbor_count(uint32_t flags) { return ((flags & 1u) ? 12u : 0u) + ((flags & 2u) ? 4u : 0u); }

/* searchSimulationERD.slang:59-80: wrap (W) or bounds check, linear read */
static int erd_alive(const uint32_t *in, int x, int y, int z, uint32_t M, uint32_t X, uint32_t Y, uint32_t Z,
                     int wrap)
{
    if (wrap) {
        x = (x + (int)X) % (int)X;
        y = (y + (int)Y) % (int)Y;
        z = (z + (int)Z) % (int)Z;
    } else if (x < 0 || x >= (int)X || y < 0 || y >= (int)Y || z < 0 || z >= (int)Z) {
        return 0;
    }
    return erd_get(in, (size_t)x + (size_t)X * ((size_t)y + (size_t)Y * (size_t)z)) == M;
}

void orc_step_erd(const uint32_t *in, uint32_t *out, uint32_t X, uint32_t Y, uint32_t Z, uint64_t word0,
                  uint64_t word1, const uint64_t *valid_bitmasks, uint32_t max_rule_bits, orc_tick_stats *st)
{   /* searchSimulationERD.slang:99-217 */
    const uint32_t M = (uint32_t)(word0 >> 60) + 1u;
    const uint32_t flags = (uint32_t)(word1 >> 60);
    const int wrap = (flags & 8u) != 0;
    uint64_t B, S;
    if (max_rule_bits > 0) {
        B = valid_bitmasks[(uint32_t)(word0 & PERM_MASK) - 1u];
        S = valid_bitmasks[(uint32_t)(word1 & PERM_MASK) - 1u];
    } else {
        B = word0 & PERM_MASK;
        S = word1 & PERM_MASK;
    }
    memset(st, 0, sizeof(*st));
    for (int a = 0; a < 3; a++) { st->bbox_min[a] = 0x7fffffff; st->bbox_max[a] = -1; }
    const size_t nw = orc_erd_words(X, Y, Z);
    memset(out, 0, nw * sizeof(uint32_t));
    for (uint32_t z = 0; z < Z; z++)
        for (uint32_t y = 0; y < Y; y++)
            for (uint32_t x = 0; x < X; x++) {
                const size_t i = (size_t)x + (size_t)X * ((size_t)y + (size_t)Y * z);
                const uint32_t cur = erd_get(in, i);
                const int cfg = (int)(y & 1u);
                int n = 0;
                if (flags & 1u)
                    for (int k = 0; k < 12; k++)
                        n += erd_alive(in, (int)x + kErdFace[cfg][k][0], (int)y + kErdFace[cfg][k][1],
                                       (int)z + kErdFace[cfg][k][2], M, X, Y, Z, wrap);
                if (flags & 2u)
                    for (int k = 0; k < 4; k++)
                        n += erd_alive(in, (int)x + kErdEdge[k][0], (int)y + kErdEdge[k][1], (int)z + kErdEdge[k][2],
                                       M, X, Y, Z, wrap);
                uint32_t ns;
                if (cur == M) {
                    if (n > 0 && ((S >> (uint64_t)(n - 1)) & 1u)) { ns = M; st->snap.numberOfSurvivingCells++; }
                    else { ns = cur - 1u; st->snap.numberOfDyingCells++; }
                } else if (cur == 0u) {
                    if (n > 0 && ((B >> (uint64_t)(n - 1)) & 1u)) { ns = M; st->snap.numberOfCellsBorn++; }
                    else { ns = 0u; st->snap.numberOfDeadCells++; }
                } else {
                    ns = cur - 1u;
                    st->snap.numberOfDyingCells++;
                }
                out[i / 8u] |= (ns & 0xFu) << ((i % 8u) * 4u);
                if (ns == M) {
                    const int c[3] = {(int)x, (int)y, (int)z};
                    for (int a = 0; a < 3; a++) {
                        if (c[a] < st->bbox_min[a]) st->bbox_min[a] = c[a];
                        if (c[a] > st->bbox_max[a]) st->bbox_max[a] = c[a];
                    }
                }
            }
    st->live = st->snap.numberOfCellsBorn + st->snap.numberOfSurvivingCells;
    st->changed = st->snap.numberOfCellsBorn + st->snap.numberOfDyingCells;
}

void orc_init_cells_erd(uint32_t *cells, uint32_t X, uint32_t Y, uint32_t Z, uint32_t max_cell_state, uint32_t sx,
                        uint32_t sy, uint32_t sz, uint32_t spawn_mode, double density, uint32_t seed)
{   /* searchSystem.cpp:264-303, linear branch */
    memset(cells, 0, orc_erd_words(X, Y, Z) * sizeof(uint32_t));
    const uint32_t threshold = orc_density_threshold(density);
    for (uint32_t z = 0; z < Z; z++)
        for (uint32_t y = 0; y < Y; y++)
            for (uint32_t x = 0; x < X; x++) {
                int in_box = x >= (X - sx) / 2 && x < (X + sx) / 2 && y >= (Y - sy) / 2 && y < (Y + sy) / 2 &&
                             z >= (Z - sz) / 2 && z < (Z + sz) / 2;
                if (!in_box) continue;
                int alive = spawn_mode == 0 ? (density >= 1.0 || orc_cell_hash(x, y, z, seed) < threshold) : 1;
                if (alive) {
                    const size_t i = (size_t)x + (size_t)X * ((size_t)y + (size_t)Y * z);
                    cells[i / 8u] |= (max_cell_state & 0xFu) << ((i % 8u) * 4u);
                }
            }
}

uint64_t orc_erd_canonical_hash(const uint32_t *cells, uint32_t X, uint32_t Y, uint32_t Z)
{   /* the canonical hash's byte stream is the linear layout itself */
    const uint64_t n = (uint64_t)X * Y * Z;
    uint64_t h = 1469598103934665603ull;
    for (uint64_t i = 0; i < n; i += 2) {
        uint32_t byte = erd_get(cells, i);
        if (i + 1 < n) byte |= erd_get(cells, i + 1) << 4;
        h ^= byte;
        h *= 1099511628211ull;
    }
    return h;
}

/* ------------------------------------------------------------------ */
/* manager: shaders/slang/search/searchManager.slang                   */
/* ------------------------------------------------------------------ */
static void next_permutation(orc_grid_info *info, orc_permutation *p, uint64_t max_permutation,
                             uint64_t survival_end)
{   /* searchManager.slang:56-96 */
    if (p->survivalAndNeighborhoodRules == 0 || p->birthAndMaxCellStateRules == 0) return;
    if ((p->birthAndMaxCellStateRules & PERM_MASK) < max_permutation) {
        p->birthAndMaxCellStateRules++;
    } else if ((p->survivalAndNeighborhoodRules & PERM_MASK) < survival_end) {
        p->survivalAndNeighborhoodRules++;
        p->birthAndMaxCellStateRules = (p->birthAndMaxCellStateRules & ~PERM_MASK) + 1u;
    } else {
        p->survivalAndNeighborhoodRules = 0;
        p->birthAndMaxCellStateRules = 0;
        return;
    }
    memset(info, 0, sizeof(*info));
}

static uint32_t manager_tick(orc_grid_info *info, const orc_snapshot *snap, orc_permutation *perm,
                             uint64_t max_permutation, uint64_t survival_end,
                             uint32_t max_ticks, orc_grid_info *copy_info,
                             orc_permutation *copy_perm, int legacy_death, uint64_t total_cells,
                             int null_marker);

uint32_t orc_manager_tick(orc_grid_info *info, const orc_snapshot *snap, orc_permutation *perm,
                          uint64_t max_permutation, uint64_t survival_end,
                          uint32_t max_ticks, orc_grid_info *copy_info,
                          orc_permutation *copy_perm)
{
    return manager_tick(info, snap, perm, max_permutation, survival_end, max_ticks, copy_info, copy_perm, 0, 0, 1);
}

/* legacy_death selects the superseded death test of
 * shaders/source/search/manager.comp:115 (every cell stayed dead this tick)
 * instead of searchManager.slang:128 (no survivor and no birth).
 * null_marker: an all-zero rule word is the exhausted-slot marker
 * (searchManager.slang:103-109) and the manager does nothing; without it the
 * words are evaluated as the rule they spell (M = 1 with no birth count packs
 * to an all-zero first word). */
static uint32_t manager_tick(orc_grid_info *info, const orc_snapshot *snap, orc_permutation *perm,
                             uint64_t max_permutation, uint64_t survival_end,
                             uint32_t max_ticks, orc_grid_info *copy_info,
                             orc_permutation *copy_perm, int legacy_death, uint64_t total_cells,
                             int null_marker)
{   /* searchManager.slang:98-149 */
    if (null_marker && (perm->survivalAndNeighborhoodRules == 0 || perm->birthAndMaxCellStateRules == 0)) {
        *copy_info = *info;
        *copy_perm = *perm;
        return copy_info->status;
    }
    orc_grid_info g = *info;
    g.numberOfSimulationTicksSurvived++;
    g.totalNumberOfCellsThatDied += snap->numberOfDeadCells;
    g.totalNumberOfCellsThatWereBorn += snap->numberOfCellsBorn;
    g.finalAliveCount = snap->numberOfSurvivingCells;
    if (snap->numberOfSurvivingCells > g.maxAliveCount) g.maxAliveCount = snap->numberOfSurvivingCells;
    int died = legacy_death ? ((uint64_t)snap->numberOfDeadCells == total_cells)
                            : (snap->numberOfSurvivingCells == 0u && snap->numberOfCellsBorn == 0u);
    if (died) g.status = 1;
    else if (g.numberOfSimulationTicksSurvived >= max_ticks) g.status = 2;
    else g.status = 0;
    *info = g;
    *copy_info = g;
    *copy_perm = *perm;
    if (g.status > 0) next_permutation(info, perm, max_permutation, survival_end);
    return g.status;
}

uint32_t orc_classify(const orc_grid_info *info, uint32_t max_cell_state)
{   /* searchSystem.cpp:695-711 */
    if (info->status == 2) return 2;
    if ((uint64_t)info->numberOfSimulationTicksSurvived > (uint64_t)max_cell_state + 1u) return 1;
    return 0;
}

/* ------------------------------------------------------------------ */
/* enumeration: src/common/permutationMath.h                           */
/* ------------------------------------------------------------------ */
static int cmp_u64(const void *a, const void *b)
{
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return (x > y) - (x < y);
}

typedef struct { uint64_t *out; size_t cap; size_t n; uint32_t N; } emit_ctx;

static void emit(emit_ctx *c, uint64_t acc, uint32_t start, uint32_t remaining)
{   /* permutationMath.h:109-121 */
    if (remaining == 0) {
        if (acc != 0) { if (c->n < c->cap && c->out) c->out[c->n] = acc; c->n++; }
        return;
    }
    for (uint32_t i = start; i + remaining <= c->N + 1 - 1; i++)
        emit(c, acc | (((uint64_t)1) << i), i + 1, remaining - 1);
}

size_t orc_enumerate_valid_bitmasks(uint32_t neighbor_count, uint32_t max_bits, uint64_t *out, size_t cap)
{   /* permutationMath.h:101-128 */
    if (max_bits == 0 || neighbor_count == 0) return 0;
    uint32_t eff = max_bits < neighbor_count ? max_bits : neighbor_count;
    emit_ctx c = {out, cap, 0, neighbor_count};
    for (uint32_t p = 1; p <= eff; p++) emit(&c, 0, 0, p);
    if (out && c.n <= cap) qsort(out, c.n, sizeof(uint64_t), cmp_u64);
    return c.n;
}

uint64_t orc_rule_permutation_count(uint32_t n)
{   /* permutationMath.h:16-22 */
    if (n == 0) return 0;
    if (n >= 60) n = 60;
    uint64_t r = (((uint64_t)1) << n) - 1;
    return r * r;
}

uint32_t orc_neighbor_count(uint32_t flags)
{   /* shape.h:117-127 with cube.h:13-46 counts */
    return ((flags & 1u) ? 6u : 0u) + ((flags & 2u) ? 12u : 0u) + ((flags & 4u) ? 8u : 0u);
}

uint64_t orc_search_total_permutations(int min_m, int max_m, int faces, int edges, int corners, int wrap)
{   /* permutationMath.h:56-81 */
    int cnt = max_m - min_m + 1;
    if (cnt <= 0) return 0;
    uint64_t total = 0;
    for (int f = 0; f <= (faces ? 1 : 0); f++)
        for (int e = 0; e <= (edges ? 1 : 0); e++)
            for (int c = 0; c <= (corners ? 1 : 0); c++) {
                if (!f && !e && !c) continue;
                total += orc_rule_permutation_count(orc_neighbor_count((f ? 1u : 0u) | (e ? 2u : 0u) | (c ? 4u : 0u)));
            }
    total *= (uint64_t)cnt;
    if (wrap) total *= 2;
    return total;
}

/* ------------------------------------------------------------------ */
/* rule strings: src/vulkan/systems/guiSystem.cpp:92-163               */
/* ------------------------------------------------------------------ */
uint64_t orc_rule_string_to_bitmask(const char *s)
{   /* guiSystem.cpp:133-163: comma tokens, whitespace removed, "a-b" ranges,
       values outside 1..60 ignored */
    uint64_t mask = 0;
    char tok[256];
    const char *p = s;
    while (1) {
        size_t n = 0;
        while (*p && *p != ',') {
            if (*p != ' ' && *p != '\t' && *p != '\n' && *p != '\r' && *p != '\v' && *p != '\f' && n < sizeof(tok) - 1)
                tok[n++] = *p;
            p++;
        }
        tok[n] = 0;
        if (n > 0) {
            char *dash = strchr(tok, '-');
            if (dash) {
                *dash = 0;
                int a = atoi(tok), b = atoi(dash + 1);
                for (int i = a; i <= b && i <= 60; i++)
                    if (i >= 1) mask |= ((uint64_t)1) << (i - 1);
            } else {
                int v = atoi(tok);
                if (v >= 1 && v <= 60) mask |= ((uint64_t)1) << (v - 1);
            }
        }
        if (!*p) break;
        p++;
    }
    return mask;
}

static size_t put(char *out, size_t cap, size_t n, const char *s)
{
    while (*s) { if (n + 1 < cap) out[n] = *s; n++; s++; }
    if (cap) out[n < cap ? n : cap - 1] = 0;
    return n;
}

size_t orc_bitmask_to_rule_string(uint64_t bitmask, char *out, size_t cap)
{   /* guiSystem.cpp:92-131: runs of >= 3 consecutive values become "a-b" */
    int bits[60], nb = 0;
    for (int b = 0; b < 60; b++)
        if ((bitmask >> b) & 1u) bits[nb++] = b + 1;
    size_t n = 0;
    if (cap) out[0] = 0;
    if (nb == 0) return put(out, cap, 0, "-");
    int i = 0;
    char buf[32];
    while (i < nb) {
        int start = i;
        while (i + 1 < nb && bits[i + 1] == bits[i] + 1) i++;
        if (n > 0) n = put(out, cap, n, ",");
        if (i - start >= 2) {
            snprintf(buf, sizeof buf, "%d-%d", bits[start], bits[i]);
            n = put(out, cap, n, buf);
        } else {
            for (int j = start; j <= i; j++) {
                if (j > start) n = put(out, cap, n, ",");
                snprintf(buf, sizeof buf, "%d", bits[j]);
                n = put(out, cap, n, buf);
            }
        }
        i++;
    }
    return n;
}

/* ------------------------------------------------------------------ */
/* one rule evaluation (SearchSystem per-slot lifecycle)               */
/* ------------------------------------------------------------------ */
static uint32_t run_rule(uint32_t X, uint32_t Y, uint32_t Z, uint32_t max_cell_state,
                         uint32_t flags, uint64_t birth, uint64_t survival,
                         uint32_t max_ticks, const uint32_t *initial_cells,
                         orc_grid_info *info_out, orc_tick_stats *per_tick,
                         uint32_t *final_cells, int legacy_death, int erd);

uint32_t orc_run_rule_erd(uint32_t X, uint32_t Y, uint32_t Z, uint32_t max_cell_state, uint32_t flags,
                          uint64_t birth, uint64_t survival, uint32_t max_ticks, const uint32_t *initial_cells,
                          orc_grid_info *info_out, orc_tick_stats *per_tick, uint32_t *final_cells)
{
    return run_rule(X, Y, Z, max_cell_state, flags, birth, survival, max_ticks, initial_cells, info_out, per_tick,
                    final_cells, 0, 1);
}

uint32_t orc_run_rule(uint32_t X, uint32_t Y, uint32_t Z, uint32_t max_cell_state,
                      uint32_t flags, uint64_t birth, uint64_t survival,
                      uint32_t max_ticks, const uint32_t *initial_cells,
                      orc_grid_info *info_out, orc_tick_stats *per_tick,
                      uint32_t *final_cells)
{
    return run_rule(X, Y, Z, max_cell_state, flags, birth, survival, max_ticks, initial_cells,
                    info_out, per_tick, final_cells, 0, 0);
}

static uint32_t run_rule(uint32_t X, uint32_t Y, uint32_t Z, uint32_t max_cell_state,
                         uint32_t flags, uint64_t birth, uint64_t survival,
                         uint32_t max_ticks, const uint32_t *initial_cells,
                         orc_grid_info *info_out, orc_tick_stats *per_tick,
                         uint32_t *final_cells, int legacy_death, int erd)
{
    size_t nw = erd ? orc_erd_words(X, Y, Z) : orc_cell_words(X, Y, Z);
    uint32_t *a = (uint32_t *)malloc(nw * sizeof(uint32_t));
    uint32_t *b = (uint32_t *)malloc(nw * sizeof(uint32_t));
    memcpy(a, initial_cells, nw * sizeof(uint32_t));
    /* starting permutation: searchSystem.cpp:191-194, raw masks (K = 0 form) */
    orc_permutation perm;
    perm.birthAndMaxCellStateRules = (((uint64_t)(max_cell_state - 1u) & 0xFu) << 60) | birth;
    perm.survivalAndNeighborhoodRules = (((uint64_t)flags) << 60) | survival;
    orc_grid_info info, ci;
    orc_permutation cp;
    memset(&info, 0, sizeof info);
    uint32_t t = 0;
    while (1) {
        orc_tick_stats st;
        if (erd)
            orc_step_erd(a, b, X, Y, Z, perm.birthAndMaxCellStateRules, perm.survivalAndNeighborhoodRules, NULL, 0,
                         &st);
        else
            step_impl(a, b, X, Y, Z, perm.birthAndMaxCellStateRules, perm.survivalAndNeighborhoodRules,
                      NULL, 0, &st, 0);
        if (per_tick && t < max_ticks) per_tick[t] = st;
        t++;
        /* the copy-based ping-pong of searchSystem.cpp:599-616 */
        uint32_t *tmp = a; a = b; b = tmp;
        /* max_permutation / survival_end equal to the current values: the
           manager exhausts the slot right after this rule completes. */
        uint32_t status = manager_tick(&info, &st.snap, &perm, birth, survival, max_ticks, &ci, &cp,
                                       legacy_death, (uint64_t)X * Y * Z, 0);
        if (status > 0 || t >= 0x7fffffffu) break;
    }
    if (info_out) *info_out = ci;
    if (final_cells) memcpy(final_cells, a, nw * sizeof(uint32_t));
    free(a);
    free(b);
    return t;
}

/* ------------------------------------------------------------------ */
/* a whole search chunk                                                 */
/* ------------------------------------------------------------------ */
typedef struct {
    uint32_t M, flags;
    uint64_t birth, survival;
} orc_job;

typedef struct {
    const orc_job *jobs;
    size_t n_jobs;
    size_t next;
    pthread_mutex_t mu;
    uint32_t X;
    uint32_t max_ticks;
    uint32_t **initial; /* indexed by M */
    int legacy_death;
    int erd;
    uint64_t counts[3];
} search_ctx;

static void *search_worker(void *arg)
{
    search_ctx *c = (search_ctx *)arg;
    uint64_t local[3] = {0, 0, 0};
    while (1) {
        pthread_mutex_lock(&c->mu);
        size_t i = c->next;
        if (i < c->n_jobs) c->next = i + 64;
        pthread_mutex_unlock(&c->mu);
        if (i >= c->n_jobs) break;
        size_t e = i + 64 < c->n_jobs ? i + 64 : c->n_jobs;
        for (; i < e; i++) {
            const orc_job *j = &c->jobs[i];
            orc_grid_info info;
            run_rule(c->X, c->X, c->X, j->M, j->flags, j->birth, j->survival, c->max_ticks,
                     c->initial[j->M], &info, NULL, NULL, c->legacy_death, c->erd);
            local[orc_classify(&info, j->M)]++;
        }
    }
    pthread_mutex_lock(&c->mu);
    for (int k = 0; k < 3; k++) c->counts[k] += local[k];
    pthread_mutex_unlock(&c->mu);
    return NULL;
}

uint64_t orc_run_search(uint32_t grid, uint32_t sel_flags, uint32_t max_rule_bits,
                        int min_m, int max_m, uint32_t max_ticks,
                        uint32_t spawn, uint32_t spawn_mode, double density,
                        uint32_t seed, int n_threads, uint64_t counts[3])
{
    return orc_run_search_ex(grid, sel_flags, max_rule_bits, min_m, max_m, max_ticks, spawn, spawn_mode,
                             density, seed, n_threads, counts, 0);
}

uint64_t orc_run_search_ex(uint32_t grid, uint32_t sel_flags, uint32_t max_rule_bits,
                           int min_m, int max_m, uint32_t max_ticks,
                           uint32_t spawn, uint32_t spawn_mode, double density,
                           uint32_t seed, int n_threads, uint64_t counts[3], int legacy_death)
{
    return orc_run_search_shape(grid, 0, sel_flags, max_rule_bits, min_m, max_m, max_ticks, spawn, spawn_mode,
                                density, seed, n_threads, counts, legacy_death);
}

uint64_t orc_run_search_shape(uint32_t grid, uint32_t shape, uint32_t sel_flags, uint32_t max_rule_bits,
                              int min_m, int max_m, uint32_t max_ticks, uint32_t spawn, uint32_t spawn_mode,
                              double density, uint32_t seed, int n_threads, uint64_t counts[3], int legacy_death)
{   /* shape 0: cube (3D-tiled, padded rows); shape 1: ERD (linear layout), searchMode.cpp:71-76 */
    const int erd = shape == 1u;
    /* slots: searchMode.cpp:49-102 (subsets of the selected F/E/C flags) */
    size_t cap = 0, n = 0;
    orc_job *jobs = NULL;
    uint32_t *initial[17] = {0};
    size_t nw = erd ? orc_erd_words(grid, grid, grid) : orc_cell_words(grid, grid, grid);
    for (int M = min_m; M <= max_m; M++) {
        initial[M] = (uint32_t *)malloc(nw * sizeof(uint32_t));
        uint32_t s = spawn < grid ? spawn : grid; /* searchMode.cpp:177-179 */
        if (erd) orc_init_cells_erd(initial[M], grid, grid, grid, (uint32_t)M, s, s, s, spawn_mode, density, seed);
        else orc_init_cells(initial[M], grid, grid, grid, (uint32_t)M, s, s, s, spawn_mode, density, seed);
        for (uint32_t subset = 1; subset <= sel_flags; subset++) {
            if ((subset & sel_flags) != subset) continue;
            uint32_t N = erd ? orc_erd_neighbor_count(subset) : orc_neighbor_count(subset);
            if (N == 0) continue; /* no neighbours (corners only on ERD): no rule space */
            uint64_t *vb = NULL;
            uint64_t maxp;
            if (max_rule_bits > 0) {
                size_t nv = orc_enumerate_valid_bitmasks(N, max_rule_bits, NULL, 0);
                vb = (uint64_t *)malloc(nv * sizeof(uint64_t));
                orc_enumerate_valid_bitmasks(N, max_rule_bits, vb, nv);
                maxp = nv;
            } else {
                maxp = (((uint64_t)1) << N) - 1;
            }
            /* iteration order of searchManager.slang:56-96 over the full range */
            for (uint64_t si = 1; si <= maxp; si++)
                for (uint64_t bi = 1; bi <= maxp; bi++) {
                    if (n == cap) { cap = cap ? cap * 2 : 4096; jobs = (orc_job *)realloc(jobs, cap * sizeof(orc_job)); }
                    jobs[n].M = (uint32_t)M;
                    jobs[n].flags = subset;
                    jobs[n].birth = vb ? vb[bi - 1] : bi;
                    jobs[n].survival = vb ? vb[si - 1] : si;
                    n++;
                }
            free(vb);
        }
    }
    search_ctx c;
    memset(&c, 0, sizeof c);
    c.jobs = jobs;
    c.n_jobs = n;
    c.X = grid;
    c.max_ticks = max_ticks;
    c.initial = initial;
    c.legacy_death = legacy_death;
    c.erd = erd;
    pthread_mutex_init(&c.mu, NULL);
    if (n_threads < 1) n_threads = 1;
    pthread_t th[256];
    if (n_threads > 256) n_threads = 256;
    for (int i = 0; i < n_threads; i++) pthread_create(&th[i], NULL, search_worker, &c);
    for (int i = 0; i < n_threads; i++) pthread_join(th[i], NULL);
    pthread_mutex_destroy(&c.mu);
    for (int k = 0; k < 3; k++) counts[k] = c.counts[k];
    for (int M = 0; M <= 16; M++) free(initial[M]);
    free(jobs);
    return (uint64_t)n;
}
