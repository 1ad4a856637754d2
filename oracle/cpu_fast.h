/*
 * cpu_fast.h -- the fast CPU baseline of the search-mode step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (cave_amd/, the C-ABI
 * library, the CLI) includes, links or calls this code.  bench.py times it as
 * the host-CPU baseline; tests/test_cpu_fast.py holds it bit-exact against the
 * scalar restatement (cave_oracle.c), which stays the parity reference.
 *
 * Same semantics as orc_step_stats / orc_run_rule (no wrap, raw rule
 * bitmasks; shaders/slang/search/searchSimulationCube.slang:78-247 and
 * searchManager.slang:98-149), computed 64 cells at a time:
 *   - cells are bit-sliced: a row of X cells is ceil(X/64) words per state
 *     bit, plus one "live" word row (state == M) with a zero guard word at
 *     each end and zero guard rows around every plane;
 *   - neighbour counts are separable: per row, the live bits, their two x
 *     neighbours and the 3-cell row sum; per plane, the in-plane part and
 *     the part the plane hands to z-1 and z+1; per cell, a 5-bit sum;
 *   - the rule is a decode of the 5-bit count into minterms, OR-ed over the
 *     counts the birth / survival masks allow;
 *   - one grid is split into z slabs over threads (one barrier per tick);
 *     a list of simulations is spread over threads, one simulation each.
 */
#ifndef CAVE_CPU_FAST_H
#define CAVE_CPU_FAST_H

#include <stddef.h>
#include <stdint.h>

#include "cave_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ofs_grid ofs_grid;

/* A grid from cells in the packed layout of orc_get_cell (8 cells per
 * 32-bit word, rows padded to whole words).  NULL on bad sizes or no memory. */
ofs_grid *ofs_create(uint32_t X, uint32_t Y, uint32_t Z, const uint32_t *cells);
void ofs_destroy(ofs_grid *g);
/* The current cells, back in the packed layout (padding nibbles 0). */
void ofs_read(const ofs_grid *g, uint32_t *cells);
/* `ticks` steps of rule (M, flags, birth, survival), M in 1..15, flags in
 * 1..7, on n_threads threads.  per_tick (nullable) receives `ticks`
 * entries, equal to what orc_step_stats reports.  0 on success, -1 on bad
 * arguments, -2 when a thread cannot be started. */
int ofs_run(ofs_grid *g, uint32_t M, uint32_t flags, uint64_t birth, uint64_t survival, uint32_t ticks,
            int n_threads, orc_tick_stats *per_tick);

/* One simulation of a list: grid, rule, tick budget and initial cells
 * (the arguments of orc_init_cells). */
typedef struct {
    uint32_t X, Y, Z, M, flags, max_ticks;
    uint64_t birth, survival;
    uint32_t sx, sy, sz, spawn_mode, seed, reserved;
    double density;
} ofs_sim;

/* Runs simulations of the list as orc_run_rule does (fresh cells, manager
 * after every tick, stop at status > 0), on n_threads threads that take
 * them in list order.  No simulation is started once budget_s seconds have
 * passed (budget_s <= 0: no limit).  ticks_out[i] receives the ticks run,
 * or UINT32_MAX for a simulation not run; info_out (nullable) the final
 * grid info.  Returns the number of simulations run; *cell_updates (nullable)
 * the cells x ticks they executed. */
size_t ofs_run_sims(const ofs_sim *sims, size_t n, int n_threads, double budget_s, uint32_t *ticks_out,
                    orc_grid_info *info_out, uint64_t *cell_updates);
/* ofs_run_sims, and hash_out (nullable) receives each simulation's final
 * cells hashed as orc_canonical_hash does. */
size_t ofs_run_sims_hash(const ofs_sim *sims, size_t n, int n_threads, double budget_s, uint32_t *ticks_out,
                         orc_grid_info *info_out, uint64_t *cell_updates, uint64_t *hash_out);

#ifdef __cplusplus
}
#endif
#endif
