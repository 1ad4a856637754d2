/*
 * csv_provenance.c -- candidate histories of the reference's benchmark CSV.
 *
 * TEST INFRASTRUCTURE ONLY (driven by tests/golden/csv_provenance.py; nothing
 * in the product includes, links or calls it).
 *
 * benchmarks/reference/debug/bench_results.csv:2-9 holds the classification
 * counts of one sweep (cube, FE, --max-rule-bits 2, --min-cs 2
 * --max-cs-range 3, --ticks 20, 11^3 filled spawn) at grids 25 and 49.  The
 * oracle (the current shaders) reproduces the 49^3 unviable count under the
 * superseded death test but not the 25^3 row.  This file simulates the same
 * sweep under named alternative semantics ("variants"), each a documented or
 * plausible earlier state of the reference, so that each can be tested
 * against all six CSV numbers:
 *
 *   0  current semantics: padded rows everywhere (searchSimulationCube.slang:
 *      78-94, 159-216; searchSystem.cpp:270-300).
 *   1  the neighbour-read defect described at searchSimulationCube.slang:
 *      83-87: own cell at the padded index, neighbour cells at the linear
 *      index (cellIndex / 8) into the same padded buffer.
 *   2  initial cells written linearly (searchSystem.cpp:291-296, the ERD
 *      branch) while the shader addresses padded rows
 *      (searchSystem.cpp:133-139 "Mixing the two corrupts neighbor reads").
 *   3  initial cells written padded while the shader is the prior 1D form
 *      (one thread per linear uint, cellIndex = 8 * uint + nibble,
 *      linear neighbours; searchSimulationCube.slang:7-12).
 *   4  the prior 1D form throughout (initial cells, buffer and shader all
 *      linear), neighbours bounds-checked on the linear cell index only
 *      (0 <= index < total): x and y neighbours past a face read the
 *      adjacent row / plane.
 *   5  current semantics, manager checking the tick budget before the death
 *      test (searchManager.slang:128-135 in the other order); this one only
 *      reclassifies variant 0's runs, so it needs no simulation here.
 *
 * Storage follows Vulkan robust buffer access: reads past a buffer's end
 * return 0 and writes past it are dropped.  Each run reports, per rule, the
 * tick of the first "no survivor and no birth" (current death test,
 * searchManager.slang:128) and of the first all-dead tick (superseded test,
 * shaders/source/search/manager.comp:115), so one simulation classifies the
 * rule under both tests (searchSystem.cpp:695-711).
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    uint32_t M, flags;
    uint64_t birth, survival;
} cp_job;

static const int kFace[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
static const int kEdge[12][3] = {{1, 1, 0}, {1, -1, 0}, {-1, 1, 0}, {-1, -1, 0}, {1, 0, 1}, {1, 0, -1},
                                 {-1, 0, 1}, {-1, 0, -1}, {0, 1, 1}, {0, 1, -1}, {0, -1, 1}, {0, -1, -1}};

typedef struct {
    int variant;
    uint32_t G;        /* grid edge */
    uint32_t upr;      /* padded words per row */
    size_t nw;         /* words of the buffer */
    int linear_buf;    /* buffer sized (and addressed) linearly */
} geo;

static uint32_t rd(const uint32_t *b, size_t nw, size_t cell_slot)
{   /* nibble `cell_slot` of a buffer of nw words, 0 past the end */
    size_t w = cell_slot / 8u;
    if (w >= nw) return 0u;
    return (b[w] >> ((cell_slot % 8u) * 4u)) & 0xFu;
}

static void wr(uint32_t *b, size_t nw, size_t cell_slot, uint32_t v)
{
    size_t w = cell_slot / 8u;
    if (w >= nw) return;
    b[w] |= (v & 0xFu) << ((cell_slot % 8u) * 4u);
}

static size_t padded_slot(const geo *g, int x, int y, int z)
{
    return (size_t)8u * ((size_t)g->upr * ((size_t)y + (size_t)g->G * (size_t)z)) + (size_t)x;
}

static size_t linear_slot(const geo *g, int x, int y, int z)
{
    return (size_t)x + (size_t)g->G * ((size_t)y + (size_t)g->G * (size_t)z);
}

/* the shader's own-cell slot and a neighbour's slot (or SIZE_MAX: none) */
static size_t own_slot(const geo *g, int x, int y, int z)
{
    return (g->variant == 3 || g->variant == 4) ? linear_slot(g, x, y, z) : padded_slot(g, x, y, z);
}

static size_t nb_slot(const geo *g, int x, int y, int z, int dx, int dy, int dz)
{
    const int G = (int)g->G;
    int nx = x + dx, ny = y + dy, nz = z + dz;
    if (g->variant == 4) {
        long long li = (long long)x + G * ((long long)y + G * (long long)z) + dx + (long long)G * (dy + (long long)G * dz);
        if (li < 0 || li >= (long long)G * G * G) return SIZE_MAX;
        return (size_t)li;
    }
    if (nx < 0 || nx >= G || ny < 0 || ny >= G || nz < 0 || nz >= G) return SIZE_MAX;
    if (g->variant == 1 || g->variant == 3) return linear_slot(g, nx, ny, nz);
    return padded_slot(g, nx, ny, nz);
}

static void init_cells(const geo *g, uint32_t *b, uint32_t M, uint32_t spawn)
{   /* searchSystem.cpp:268-303, filled spawn */
    memset(b, 0, g->nw * sizeof(uint32_t));
    const uint32_t G = g->G, s = spawn < G ? spawn : G;
    for (uint32_t x = (G - s) / 2; x < (G + s) / 2; x++)
        for (uint32_t y = (G - s) / 2; y < (G + s) / 2; y++)
            for (uint32_t z = (G - s) / 2; z < (G + s) / 2; z++) {
                /* variant 2 writes linearly into the padded buffer; variant 4
                   is the 1D form throughout, its initial cells linear too */
                size_t slot = (g->variant == 2 || g->variant == 4) ? linear_slot(g, (int)x, (int)y, (int)z)
                                                                   : padded_slot(g, (int)x, (int)y, (int)z);
                wr(b, g->nw, slot, M);
            }
}

/* one tick; returns the snapshot {dead, dying, born, surviving} */
static void step(const geo *g, const uint32_t *in, uint32_t *out, uint32_t M, uint32_t flags, uint64_t B,
                 uint64_t S, uint32_t snap[4])
{
    memset(out, 0, g->nw * sizeof(uint32_t));
    snap[0] = snap[1] = snap[2] = snap[3] = 0;
    const int G = (int)g->G;
    for (int z = 0; z < G; z++)
        for (int y = 0; y < G; y++)
            for (int x = 0; x < G; x++) {
                const size_t os = own_slot(g, x, y, z);
                const uint32_t cur = rd(in, g->nw, os);
                int n = 0;
                if (flags & 1u)
                    for (int i = 0; i < 6; i++) {
                        size_t s = nb_slot(g, x, y, z, kFace[i][0], kFace[i][1], kFace[i][2]);
                        n += s != SIZE_MAX && rd(in, g->nw, s) == M;
                    }
                if (flags & 2u)
                    for (int i = 0; i < 12; i++) {
                        size_t s = nb_slot(g, x, y, z, kEdge[i][0], kEdge[i][1], kEdge[i][2]);
                        n += s != SIZE_MAX && rd(in, g->nw, s) == M;
                    }
                uint32_t ns;
                if (cur == M) {
                    if (n > 0 && ((S >> (n - 1)) & 1u)) { ns = M; snap[3]++; }
                    else { ns = cur - 1u; snap[1]++; }
                } else if (cur == 0u) {
                    if (n > 0 && ((B >> (n - 1)) & 1u)) { ns = M; snap[2]++; }
                    else { ns = 0u; snap[0]++; }
                } else {
                    ns = cur - 1u;
                    snap[1]++;
                }
                wr(out, g->nw, os, ns);
            }
}

/* Runs one rule for up to max_ticks ticks.  res[0]: the tick (1-based) at
 * which the current death test first fires (0: never), res[1]: the same for
 * the superseded all-dead test. */
static void run_rule(const geo *g, const cp_job *j, uint32_t max_ticks, uint32_t spawn, uint32_t *a, uint32_t *b,
                     uint32_t res[2])
{
    init_cells(g, a, j->M, spawn);
    res[0] = res[1] = 0;
    const uint64_t total = (uint64_t)g->G * g->G * g->G;
    for (uint32_t t = 1; t <= max_ticks; t++) {
        uint32_t snap[4];
        step(g, a, b, j->M, j->flags, j->birth, j->survival, snap);
        uint32_t *tmp = a; a = b; b = tmp;
        if (!res[0] && snap[3] == 0u && snap[2] == 0u) res[0] = t;
        if (!res[1] && (uint64_t)snap[0] == total) res[1] = t;
        if (res[0] && res[1]) break;
    }
}

typedef struct {
    const cp_job *jobs;
    size_t n, next;
    pthread_mutex_t mu;
    geo g;
    uint32_t max_ticks, spawn;
    uint32_t *res; /* 2 per job */
} cp_ctx;

static void *worker(void *arg)
{
    cp_ctx *c = (cp_ctx *)arg;
    uint32_t *a = (uint32_t *)malloc(c->g.nw * sizeof(uint32_t));
    uint32_t *b = (uint32_t *)malloc(c->g.nw * sizeof(uint32_t));
    for (;;) {
        pthread_mutex_lock(&c->mu);
        size_t i = c->next;
        c->next = i + 16;
        pthread_mutex_unlock(&c->mu);
        if (i >= c->n) break;
        size_t e = i + 16 < c->n ? i + 16 : c->n;
        for (; i < e; i++) run_rule(&c->g, &c->jobs[i], c->max_ticks, c->spawn, a, b, c->res + 2 * i);
    }
    free(a);
    free(b);
    return NULL;
}

/* Simulates every job under `variant` on a G^3 grid; res gets 2 ticks per job. */
int cp_run(int variant, uint32_t G, const cp_job *jobs, size_t n, uint32_t max_ticks, uint32_t spawn, int threads,
           uint32_t *res)
{
    cp_ctx c;
    memset(&c, 0, sizeof c);
    c.jobs = jobs;
    c.n = n;
    c.g.variant = variant;
    c.g.G = G;
    c.g.upr = (G + 7u) / 8u;
    /* variant 3: the host sized and filled the buffer for padded rows; only
       the prior 1D shader addresses it linearly */
    c.g.linear_buf = variant == 4;
    c.g.nw = c.g.linear_buf ? ((size_t)G * G * G + 7u) / 8u : (size_t)c.g.upr * G * G;
    c.max_ticks = max_ticks;
    c.spawn = spawn;
    c.res = res;
    pthread_mutex_init(&c.mu, NULL);
    if (threads < 1) threads = 1;
    if (threads > 64) threads = 64;
    pthread_t th[64];
    for (int i = 0; i < threads; i++) pthread_create(&th[i], NULL, worker, &c);
    for (int i = 0; i < threads; i++) pthread_join(th[i], NULL);
    pthread_mutex_destroy(&c.mu);
    return 0;
}
