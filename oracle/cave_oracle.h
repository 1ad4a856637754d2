/*
 * cave_oracle.h -- CPU restatement of CAVE's search-mode step semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (cave_amd/, the C-ABI
 * library, the CLI) includes, links or calls this code.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, and only
 * as the checker / the timed CPU baseline.
 *
 * Every function restates one piece of the reference and cites the file:line
 * it follows (paths relative to the reference repository root).  The code is
 * a plain, per-cell, literal restatement: no SWAR, no tiling, no shortcuts, so
 * that it stays easy to audit against the Slang/C++ text.
 *
 * Pinning (what checks each function; tests/test_oracle.py):
 *   orc_rule_string_to_bitmask / orc_bitmask_to_rule_string
 *       -- cli-rules KAT (src/testing/testRunner.cpp:231-260).
 *   orc_enumerate_valid_bitmasks
 *       -- sizes quoted at src/common/permutationMath.h:94-98 (21/105/351/17901),
 *          closed form sum_k C(N,k), sortedness, brute-force popcount filter.
 *   orc_search_total_permutations
 *       -- 71,532 = CSV Total (benchmarks/reference/debug/bench_results.csv:2-9).
 *   orc_step (transition + snapshot)
 *       -- an independent numpy model (27-offset formulation), brute force on
 *          tiny grids, invariants (bucket sum, dead grid stays dead, dying
 *          decrements, padding stays 0, single-cell birth of 6/18/26 neighbours).
 *   orc_run_search (manager + classification + init + step together)
 *       -- the reference's 49^3 integration counts (bench_results.csv:6-9):
 *          under the superseded death test (manager.comp:115) the unviable
 *          count (24,732) and the viable + possible total match exactly.
 *          Parity unpinned: the viable / possible split of 248 rules, and the
 *          25^3 CSV row, which neither death test reproduces.
 *   orc_init_cells
 *       -- textbook Thomas Wang 32-bit hash vectors + spawn-box closed form.
 */
#ifndef CAVE_ORACLE_H
#define CAVE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* src/common/structs.h:79-87 */
typedef struct {
    uint32_t totalNumberOfCellsThatDied;
    uint32_t totalNumberOfCellsThatWereBorn;
    uint32_t numberOfSimulationTicksSurvived;
    uint32_t status;
    uint32_t maxAliveCount;
    uint32_t finalAliveCount;
} orc_grid_info;

/* src/common/structs.h:130-136 */
typedef struct {
    uint32_t numberOfDeadCells;
    uint32_t numberOfDyingCells;
    uint32_t numberOfCellsBorn;
    uint32_t numberOfSurvivingCells;
} orc_snapshot;

/* shaders/slang/search/searchManager.slang:40-44 */
typedef struct {
    uint64_t birthAndMaxCellStateRules;
    uint64_t survivalAndNeighborhoodRules;
} orc_permutation;

/* Per-tick statistics: the reference snapshot plus the derived live /
 * changed counts and the bounding box of live (state == M) cells after the
 * tick.  bbox_min > bbox_max on an axis means "no live cell". */
typedef struct {
    orc_snapshot snap;
    uint32_t live;     /* born + surviving */
    uint32_t changed;  /* born + dying      */
    int32_t bbox_min[3];
    int32_t bbox_max[3];
} orc_tick_stats;

/* ---- packed layout (searchSimulationCube.slang:52-53, 88-92) ---- */
uint32_t orc_words_per_row(uint32_t X);
size_t orc_cell_words(uint32_t X, uint32_t Y, uint32_t Z);
uint32_t orc_get_cell(const uint32_t *cells, uint32_t X, uint32_t Y, uint32_t x, uint32_t y, uint32_t z);
void orc_set_cell(uint32_t *cells, uint32_t X, uint32_t Y, uint32_t x, uint32_t y, uint32_t z, uint32_t v);
/* FNV-1a-64 over the canonical unpadded packing: linear index x+X(y+Yz),
 * two cells per byte, even index in the low nibble. */
uint64_t orc_canonical_hash(const uint32_t *cells, uint32_t X, uint32_t Y, uint32_t Z);

/* ---- init (src/vulkan/systems/searchSystem.cpp:224-306) ---- */
uint32_t orc_wang_hash(uint32_t seed);
uint32_t orc_cell_hash(uint32_t x, uint32_t y, uint32_t z, uint32_t seed);
uint32_t orc_density_threshold(double density);
/* spawn_mode: 0 = random (hash < threshold), 1 = filled.  threshold as
 * returned by orc_density_threshold; density >= 1 means every box cell. */
void orc_init_cells(uint32_t *cells, uint32_t X, uint32_t Y, uint32_t Z,
                    uint32_t max_cell_state, uint32_t sx, uint32_t sy, uint32_t sz,
                    uint32_t spawn_mode, double density, uint32_t seed);

/* ---- step (shaders/slang/search/searchSimulationCube.slang:78-247) ----
 * word0/word1: rule words.  When max_rule_bits > 0 the low 60 bits are
 * 1-based indices into valid_bitmasks, otherwise raw bitmasks. */
/* the headless render-path step (rayMarchCompute.slang): wraps every axis
 * when the W flag (word1 bit 63) is set; rules are raw bitmasks */
void orc_step_render(const uint32_t *in, uint32_t *out, uint32_t X, uint32_t Y, uint32_t Z,
                     uint64_t word0, uint64_t word1, orc_tick_stats *stats);
/* render-path initial cells (cellInit.slang): unsalted hash, float threshold */
uint32_t orc_render_density_threshold(double density, int *all_alive);
void orc_init_cells_render(uint32_t *cells, uint32_t X, uint32_t Y, uint32_t Z, uint32_t max_cell_state,
                           uint32_t sx, uint32_t sy, uint32_t sz, uint32_t spawn_mode, double density);

/* ERD shape (searchSimulationERD.slang): linear layout, 12 face / 4 edge deltas */
size_t orc_erd_words(uint32_t X, uint32_t Y, uint32_t Z);
uint32_t orc_erd_neighbor_count(uint32_t flags);
void orc_step_erd(const uint32_t *in, uint32_t *out, uint32_t X, uint32_t Y, uint32_t Z, uint64_t word0,
                  uint64_t word1, const uint64_t *valid_bitmasks, uint32_t max_rule_bits, orc_tick_stats *st);
void orc_init_cells_erd(uint32_t *cells, uint32_t X, uint32_t Y, uint32_t Z, uint32_t max_cell_state, uint32_t sx,
                        uint32_t sy, uint32_t sz, uint32_t spawn_mode, double density, uint32_t seed);
uint64_t orc_erd_canonical_hash(const uint32_t *cells, uint32_t X, uint32_t Y, uint32_t Z);
uint32_t orc_run_rule_erd(uint32_t X, uint32_t Y, uint32_t Z, uint32_t max_cell_state, uint32_t flags,
                          uint64_t birth, uint64_t survival, uint32_t max_ticks, const uint32_t *initial_cells,
                          orc_grid_info *info_out, orc_tick_stats *per_tick, uint32_t *final_cells);
uint64_t orc_run_search_shape(uint32_t grid, uint32_t shape, uint32_t sel_flags, uint32_t max_rule_bits,
                              int min_m, int max_m, uint32_t max_ticks, uint32_t spawn, uint32_t spawn_mode,
                              double density, uint32_t seed, int n_threads, uint64_t counts[3], int legacy_death);

void orc_step(const uint32_t *in, uint32_t *out, uint32_t X, uint32_t Y, uint32_t Z,
              uint64_t word0, uint64_t word1, const uint64_t *valid_bitmasks,
              uint32_t max_rule_bits, orc_snapshot *snap);
/* Same step, also filling the derived per-tick stats. */
void orc_step_stats(const uint32_t *in, uint32_t *out, uint32_t X, uint32_t Y, uint32_t Z,
                    uint64_t word0, uint64_t word1, const uint64_t *valid_bitmasks,
                    uint32_t max_rule_bits, orc_tick_stats *stats);

/* ---- manager (shaders/slang/search/searchManager.slang:56-149) ----
 * Returns the status written.  *copy_info / *copy_perm receive what the
 * manager writes to its host-visible copy buffers. */
uint32_t orc_manager_tick(orc_grid_info *info, const orc_snapshot *snap, orc_permutation *perm,
                          uint64_t max_permutation, uint64_t survival_end,
                          uint32_t max_ticks, orc_grid_info *copy_info,
                          orc_permutation *copy_perm);

/* ---- classification (src/vulkan/systems/searchSystem.cpp:695-711) ----
 * 2 = viable, 1 = possible, 0 = unviable. */
uint32_t orc_classify(const orc_grid_info *info, uint32_t max_cell_state);

/* ---- enumeration (src/common/permutationMath.h:16-22, 56-81, 101-128) ---- */
size_t orc_enumerate_valid_bitmasks(uint32_t neighbor_count, uint32_t max_bits,
                                    uint64_t *out, size_t cap);
uint64_t orc_rule_permutation_count(uint32_t neighbor_count);
uint32_t orc_neighbor_count(uint32_t flags);
uint64_t orc_search_total_permutations(int min_m, int max_m, int faces, int edges,
                                       int corners, int wrap);

/* ---- rule strings (src/vulkan/systems/guiSystem.cpp:92-163) ---- */
uint64_t orc_rule_string_to_bitmask(const char *s);
size_t orc_bitmask_to_rule_string(uint64_t bitmask, char *out, size_t cap);

/* ---- one rule evaluation ----
 * Runs a fresh simulation from the initial cells until the manager sets a
 * non-zero status (reference per-slot behaviour: searchSystem.cpp:496-526
 * resets the cells from initialCells whenever a permutation completes).
 * birth/survival are raw bitmasks.  per_tick (nullable) receives up to
 * max_ticks entries.  final_cells (nullable) receives the last state.
 * Returns the number of ticks run. */
uint32_t orc_run_rule(uint32_t X, uint32_t Y, uint32_t Z, uint32_t max_cell_state,
                      uint32_t flags, uint64_t birth, uint64_t survival,
                      uint32_t max_ticks, const uint32_t *initial_cells,
                      orc_grid_info *info_out, orc_tick_stats *per_tick,
                      uint32_t *final_cells);

/* ---- a whole search chunk (SearchMode + SearchSystem semantics) ----
 * Enumerates every slot (maxCS in [min_m, max_m] x non-empty neighbourhood
 * subsets of sel_flags, searchMode.cpp:49-102) and every permutation of the
 * slot's (birth, survival) index space, runs each from the chunk's initial
 * cells and counts the classification.  n_threads > 1 splits rules across
 * pthreads (results are order independent).  Returns the number of
 * evaluations; counts[0..2] = {unviable, possible, viable}. */
uint64_t orc_run_search(uint32_t grid, uint32_t sel_flags, uint32_t max_rule_bits,
                        int min_m, int max_m, uint32_t max_ticks,
                        uint32_t spawn, uint32_t spawn_mode, double density,
                        uint32_t seed, int n_threads, uint64_t counts[3]);

/* Same sweep with a selectable death test: legacy_death = 1 uses the
 * superseded rule of shaders/source/search/manager.comp:115 (status 1 when
 * every cell stayed dead during the tick). */
uint64_t orc_run_search_ex(uint32_t grid, uint32_t sel_flags, uint32_t max_rule_bits,
                           int min_m, int max_m, uint32_t max_ticks,
                           uint32_t spawn, uint32_t spawn_mode, double density,
                           uint32_t seed, int n_threads, uint64_t counts[3], int legacy_death);

#ifdef __cplusplus
}
#endif
#endif
