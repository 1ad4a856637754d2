/*
 * cpu_fast.c -- the fast CPU baseline of the search-mode step (cpu_fast.h).
 *
 * TEST INFRASTRUCTURE ONLY.  Bit-sliced, 64 cells per word, threaded with
 * pthreads.  It restates the same step as cave_oracle.c (step_impl, which
 * follows shaders/slang/search/searchSimulationCube.slang:78-247) and the
 * same manager (searchManager.slang:98-149, searchSystem.cpp:496-526 for the
 * per-slot fresh start), and is held bit-exact against them by tests.
 */
#define _POSIX_C_SOURCE 200809L
#include "cpu_fast.h"

#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef uint64_t u64;

/* the two hot loops get an AVX2 clone, picked at load time where the CPU has it */
#if defined(__x86_64__) && defined(__GNUC__) && !defined(__clang__)
#define HOT __attribute__((target_clones("avx2", "default")))
#else
#define HOT
#endif

struct ofs_grid {
    uint32_t X, Y, Z;
    uint32_t nw;   /* 64-cell words per row and state bit */
    uint32_t rs;   /* live row stride: nw data words between two zero guard words */
    u64 vmask;     /* valid cells of a row's last word */
    u64 *st[2];    /* state: row (z, y) is 4 bit planes of nw words */
    u64 *live[2];  /* state == M: Z x (Y + 2) rows, guard rows and words zero */
    int cur;
};

/* rule constants of one run */
typedef struct {
    u64 mbit[4];                      /* all ones where bit b of M is set */
    u64 selL[3], selR[3], selF[3];    /* the row value of counting kind 0, 1, 2 */
    int nb, ns;
    uint8_t bcount[32], scount[32];   /* neighbour counts that give birth / survival */
} rule_t;

/* per-tick totals of one thread */
typedef struct {
    u64 dead, dying, born, surv;
    int32_t bmin[3], bmax[3];
} acc_t;

static size_t state_row(const ofs_grid *g, uint32_t z, uint32_t y) { return ((size_t)z * g->Y + y) * 4u * g->nw; }

static size_t live_row(const ofs_grid *g, uint32_t z, int64_t y)
{
    return ((size_t)z * (g->Y + 2u) + (size_t)(y + 1)) * g->rs + 1u;
}

static size_t plane_words(const ofs_grid *g) { return (size_t)g->Y * 4u * g->nw; }

/* ---- rule set-up (searchSimulationCube.slang:139-157, 96-115) ----
 * A neighbour offset with k non-zero coordinates belongs to the faces (k = 1,
 * flag bit 0), edges (k = 2, bit 1) or corners (k = 3, bit 2).  A row of
 * offsets (dy, dz) with m non-zero coordinates contributes the row's own cell
 * when class m is on, and its two x neighbours when class m + 1 is on:
 * kind m is "nothing", "the cell", "the two x neighbours" or "all three". */
static int rule_setup(rule_t *r, uint32_t M, uint32_t flags, uint64_t birth, uint64_t survival)
{
    if (M < 1u || M > 15u || flags < 1u || flags > 7u) return -1;
    memset(r, 0, sizeof *r);
    for (int b = 0; b < 4; b++) r->mbit[b] = ((M >> b) & 1u) ? ~(u64)0 : 0;
    int on[4] = {0, (flags & 1u) != 0, (flags & 2u) != 0, (flags & 4u) != 0};
    for (int m = 0; m < 3; m++) {
        int cell = on[m], side = on[m + 1];
        r->selL[m] = (cell && !side) ? ~(u64)0 : 0;
        r->selR[m] = (!cell && side) ? ~(u64)0 : 0;
        r->selF[m] = (cell && side) ? ~(u64)0 : 0;
    }
    uint32_t nmax = orc_neighbor_count(flags);
    for (uint32_t n = 1; n <= nmax; n++) {
        if ((birth >> (n - 1u)) & 1u) r->bcount[r->nb++] = (uint8_t)n;
        if ((survival >> (n - 1u)) & 1u) r->scount[r->ns++] = (uint8_t)n;
    }
    return 0;
}

/* ---- bit-sliced arithmetic ---- */
typedef struct { u64 c, lo, hi, flo, fhi; } rowv;

/* the live word j of a row, its x neighbours' sum (lo, hi) and the 3-cell sum */
static inline rowv row_at(const u64 *l, uint32_t j)
{
    rowv v;
    const u64 *w = l + j; /* w[-1] and w[1] exist: the row has a guard word at each end */
    u64 c = w[0];
    u64 left = (c << 1) | (w[-1] >> 63);  /* bit x = cell x - 1 */
    u64 right = (c >> 1) | (w[1] << 63);  /* bit x = cell x + 1 */
    v.c = c;
    v.lo = left ^ right;
    v.hi = left & right;
    v.flo = v.lo ^ c;
    v.fhi = v.hi | (v.lo & c);
    return v;
}

static inline void kind_val(const rowv *v, const rule_t *r, int k, u64 *lo, u64 *hi)
{
    *lo = (v->c & r->selL[k]) | (v->lo & r->selR[k]) | (v->flo & r->selF[k]);
    *hi = (v->hi & r->selR[k]) | (v->fhi & r->selF[k]);
}

/* a + b + c of three 2-bit values (<= 9) into 4 bits */
static inline void add3_2(u64 a0, u64 a1, u64 b0, u64 b1, u64 c0, u64 c1, u64 *s)
{
    u64 x = a0 ^ b0, k0 = (a0 & b0) | (c0 & x);
    u64 y = a1 ^ b1, t = y ^ c1, k1 = (a1 & b1) | (c1 & y);
    s[0] = x ^ c0;
    s[1] = t ^ k0;
    u64 k2 = t & k0;
    s[2] = k1 ^ k2;
    s[3] = k1 & k2;
}

/* u + a + b of three 4-bit values (<= 26 in total) into 5 bits */
static inline void add3_4(const u64 *u, const u64 *a, const u64 *b, u64 *s)
{
    u64 x0 = u[0] ^ a[0], c0 = (u[0] & a[0]) | (b[0] & x0);
    s[0] = x0 ^ b[0];
    u64 x1 = u[1] ^ a[1], y1 = x1 ^ b[1], d1 = (u[1] & a[1]) | (b[1] & x1);
    s[1] = y1 ^ c0;
    u64 e1 = y1 & c0;
    u64 x2 = u[2] ^ a[2], y2 = x2 ^ b[2], d2 = (u[2] & a[2]) | (b[2] & x2);
    u64 z2 = d1 ^ e1;
    s[2] = y2 ^ z2;
    u64 f2 = (d1 & e1) | (y2 & z2);
    u64 x3 = u[3] ^ a[3], y3 = x3 ^ b[3], d3 = (u[3] & a[3]) | (b[3] & x3);
    u64 z3 = d2 ^ f2;
    s[3] = y3 ^ z3;
    u64 f3 = (d2 & f2) | (y3 & z3);
    s[4] = d3 ^ f3;
}

/* ---- one plane's counting parts ----
 * T(y, z): what plane z adds to the counts of planes z - 1 and z + 1
 *          (kind 1 of row y, kind 2 of rows y - 1 and y + 1);
 * U(y, z): plane z's own part (kind 0 of row y, kind 1 of rows y +- 1). */
HOT static void prepare(const ofs_grid *g, const rule_t *r, const u64 *L, uint32_t z, u64 *T, u64 *U)
{
    const uint32_t nw = g->nw;
    for (uint32_t y = 0; y < g->Y; y++) {
        const u64 *lm = L + live_row(g, z, (int64_t)y - 1);
        const u64 *l0 = L + live_row(g, z, y);
        const u64 *lp = L + live_row(g, z, (int64_t)y + 1);
        u64 *t = T + (size_t)y * 4u * nw, *u = U ? U + (size_t)y * 4u * nw : NULL;
        for (uint32_t j = 0; j < nw; j++) {
            rowv vm = row_at(lm, j), v0 = row_at(l0, j), vp = row_at(lp, j);
            u64 a0l, a0h, cml, cmh, cpl, cph, s[4];
            kind_val(&v0, r, 1, &a0l, &a0h);
            kind_val(&vm, r, 2, &cml, &cmh);
            kind_val(&vp, r, 2, &cpl, &cph);
            add3_2(a0l, a0h, cml, cmh, cpl, cph, s);
            for (int b = 0; b < 4; b++) t[b * nw + j] = s[b];
            if (u) {
                u64 k0l, k0h, aml, amh, apl, aph;
                kind_val(&v0, r, 0, &k0l, &k0h);
                kind_val(&vm, r, 1, &aml, &amh);
                kind_val(&vp, r, 1, &apl, &aph);
                add3_2(k0l, k0h, aml, amh, apl, aph, s);
                for (int b = 0; b < 4; b++) u[b * nw + j] = s[b];
            }
        }
    }
}

static void acc_init(acc_t *a)
{
    memset(a, 0, sizeof *a);
    for (int k = 0; k < 3; k++) { a->bmin[k] = 0x7fffffff; a->bmax[k] = -1; }
}

static void acc_box(acc_t *a, int32_t x0, int32_t x1, int32_t y, int32_t z)
{
    if (x0 < a->bmin[0]) a->bmin[0] = x0;
    if (x1 > a->bmax[0]) a->bmax[0] = x1;
    if (y < a->bmin[1]) a->bmin[1] = y;
    if (y > a->bmax[1]) a->bmax[1] = y;
    if (z < a->bmin[2]) a->bmin[2] = z;
    if (z > a->bmax[2]) a->bmax[2] = z;
}

/* ---- the transition of plane z (searchSimulationCube.slang:178-216) ---- */
HOT static void update(const ofs_grid *g, const rule_t *r, int cur, uint32_t z, const u64 *U, const u64 *Tm,
                   const u64 *Tp, acc_t *acc)
{
    const uint32_t nw = g->nw;
    const u64 *Sin = g->st[cur], *Lin = g->live[cur];
    u64 *Sout = g->st[cur ^ 1], *Lout = g->live[cur ^ 1];
    for (uint32_t y = 0; y < g->Y; y++) {
        const u64 *si = Sin + state_row(g, z, y);
        u64 *so = Sout + state_row(g, z, y);
        const u64 *li = Lin + live_row(g, z, y);
        u64 *lo = Lout + live_row(g, z, y);
        const size_t ro = (size_t)y * 4u * nw;
        u64 any = 0;
        for (uint32_t j = 0; j < nw; j++) {
            const u64 valid = (j + 1u == nw) ? g->vmask : ~(u64)0;
            u64 uu[4], ta[4], tb[4], n[5];
            for (int b = 0; b < 4; b++) {
                uu[b] = U[ro + b * nw + j];
                ta[b] = Tm[ro + b * nw + j];
                tb[b] = Tp[ro + b * nw + j];
            }
            add3_4(uu, ta, tb, n);
            /* count == k: low 3 bits select one of 8, high 2 bits one of 4 */
            u64 n0 = ~n[0], n1 = ~n[1], n2 = ~n[2];
            u64 q[4] = {n1 & n2, n[1] & n2, n1 & n[2], n[1] & n[2]};
            u64 m[8];
            for (int k = 0; k < 4; k++) { m[2 * k] = n0 & q[k]; m[2 * k + 1] = n[0] & q[k]; }
            u64 h[4] = {~n[3] & ~n[4], n[3] & ~n[4], ~n[3] & n[4], n[3] & n[4]};
            u64 fb = 0, fs = 0;
            for (int i = 0; i < r->nb; i++) fb |= m[r->bcount[i] & 7] & h[r->bcount[i] >> 3];
            for (int i = 0; i < r->ns; i++) fs |= m[r->scount[i] & 7] & h[r->scount[i] >> 3];
            u64 p[4];
            for (int b = 0; b < 4; b++) p[b] = si[b * nw + j];
            const u64 zero = ~(p[0] | p[1] | p[2] | p[3]);
            const u64 live = li[j];
            const u64 born = fb & zero & valid;
            const u64 surv = fs & live;
            const u64 tom = born | surv;
            /* every other non-zero cell loses one state */
            const u64 dec = ~tom & ~zero;
            u64 bw = ~p[0];
            u64 d[4];
            d[0] = ~p[0];
            d[1] = p[1] ^ bw;
            bw &= ~p[1];
            d[2] = p[2] ^ bw;
            bw &= ~p[2];
            d[3] = p[3] ^ bw;
            for (int b = 0; b < 4; b++) so[b * nw + j] = (tom & r->mbit[b]) | (dec & d[b]);
            lo[j] = tom;
            any |= tom;
            acc->born += (u64)__builtin_popcountll(born);
            acc->surv += (u64)__builtin_popcountll(surv);
            acc->dead += (u64)__builtin_popcountll(zero & ~born & valid);
            acc->dying += (u64)__builtin_popcountll(dec);
        }
        if (any) {
            uint32_t j0 = 0, j1 = nw - 1u;
            while (!lo[j0]) j0++;
            while (!lo[j1]) j1--;
            acc_box(acc, (int32_t)(64u * j0 + (uint32_t)__builtin_ctzll(lo[j0])),
                    (int32_t)(64u * j1 + 63u - (uint32_t)__builtin_clzll(lo[j1])), (int32_t)y, (int32_t)z);
        }
    }
}

typedef struct {
    u64 *T[3], *U[3], *zero;
} scratch_t;

static int scratch_alloc(scratch_t *s, const ofs_grid *g)
{
    size_t w = plane_words(g);
    memset(s, 0, sizeof *s);
    u64 *blk = (u64 *)calloc(7u * w, sizeof(u64));
    if (!blk) return -1;
    for (int i = 0; i < 3; i++) { s->T[i] = blk + (size_t)i * w; s->U[i] = blk + (size_t)(3 + i) * w; }
    s->zero = blk + 6u * w;
    return 0;
}

static void scratch_free(scratch_t *s) { free(s->T[0]); }

/* planes [z0, z1) of one tick; the planes beside the slab are read, not written */
static void run_slab(const ofs_grid *g, const rule_t *r, int cur, uint32_t z0, uint32_t z1, scratch_t *s,
                     acc_t *acc)
{
    const u64 *Lin = g->live[cur];
    acc_init(acc);
    if (z0 >= z1) return;
    if (z0 > 0) prepare(g, r, Lin, z0 - 1u, s->T[(z0 - 1u) % 3u], NULL);
    prepare(g, r, Lin, z0, s->T[z0 % 3u], s->U[z0 % 3u]);
    for (uint32_t z = z0; z < z1; z++) {
        if (z + 1u < g->Z) prepare(g, r, Lin, z + 1u, s->T[(z + 1u) % 3u], z + 1u < z1 ? s->U[(z + 1u) % 3u] : NULL);
        const u64 *Tm = z > 0 ? s->T[(z - 1u) % 3u] : s->zero;
        const u64 *Tp = z + 1u < g->Z ? s->T[(z + 1u) % 3u] : s->zero;
        update(g, r, cur, z, s->U[z % 3u], Tm, Tp, acc);
    }
}

/* live = (state == M) for planes [z0, z1) of the current buffer */
static void live_from_state(ofs_grid *g, uint32_t M, uint32_t z0, uint32_t z1)
{
    const uint32_t nw = g->nw;
    const u64 *S = g->st[g->cur];
    u64 *L = g->live[g->cur];
    for (uint32_t z = z0; z < z1; z++)
        for (uint32_t y = 0; y < g->Y; y++) {
            const u64 *s = S + state_row(g, z, y);
            u64 *l = L + live_row(g, z, y);
            for (uint32_t j = 0; j < nw; j++) {
                u64 eq = ~(u64)0;
                for (int b = 0; b < 4; b++) eq &= ((M >> b) & 1u) ? s[b * nw + j] : ~s[b * nw + j];
                l[j] = eq & ((j + 1u == nw) ? g->vmask : ~(u64)0);
            }
        }
}

static void acc_merge(acc_t *into, const acc_t *a)
{
    into->dead += a->dead;
    into->dying += a->dying;
    into->born += a->born;
    into->surv += a->surv;
    for (int k = 0; k < 3; k++) {
        if (a->bmin[k] < into->bmin[k]) into->bmin[k] = a->bmin[k];
        if (a->bmax[k] > into->bmax[k]) into->bmax[k] = a->bmax[k];
    }
}

static void acc_to_stats(const acc_t *a, orc_tick_stats *st)
{
    memset(st, 0, sizeof *st);
    st->snap.numberOfDeadCells = (uint32_t)a->dead;
    st->snap.numberOfDyingCells = (uint32_t)a->dying;
    st->snap.numberOfCellsBorn = (uint32_t)a->born;
    st->snap.numberOfSurvivingCells = (uint32_t)a->surv;
    st->live = (uint32_t)(a->born + a->surv);
    st->changed = (uint32_t)(a->born + a->dying);
    for (int k = 0; k < 3; k++) { st->bbox_min[k] = a->bmin[k]; st->bbox_max[k] = a->bmax[k]; }
}

/* ---- grid life cycle ---- */
ofs_grid *ofs_create(uint32_t X, uint32_t Y, uint32_t Z, const uint32_t *cells)
{
    if (X == 0 || Y == 0 || Z == 0) return NULL;
    ofs_grid *g = (ofs_grid *)calloc(1, sizeof *g);
    if (!g) return NULL;
    g->X = X; g->Y = Y; g->Z = Z;
    g->nw = (X + 63u) / 64u;
    g->rs = g->nw + 2u;
    g->vmask = (X % 64u) ? (((u64)1 << (X % 64u)) - 1u) : ~(u64)0;
    size_t sw = (size_t)Z * plane_words(g);
    size_t lw = (size_t)Z * (Y + 2u) * g->rs;
    for (int i = 0; i < 2; i++) {
        g->st[i] = (u64 *)calloc(sw, sizeof(u64));
        g->live[i] = (u64 *)calloc(lw, sizeof(u64));
        if (!g->st[i] || !g->live[i]) { ofs_destroy(g); return NULL; }
    }
    const uint32_t upr = orc_words_per_row(X);
    for (uint32_t z = 0; z < Z; z++)
        for (uint32_t y = 0; y < Y; y++) {
            const uint32_t *row = cells + ((size_t)z * Y + y) * upr;
            u64 *s = g->st[0] + state_row(g, z, y);
            for (uint32_t x = 0; x < X; x++) {
                uint32_t v = (row[x / 8u] >> ((x % 8u) * 4u)) & 0xFu;
                for (int b = 0; b < 4; b++)
                    if ((v >> b) & 1u) s[b * g->nw + x / 64u] |= (u64)1 << (x % 64u);
            }
        }
    return g;
}

void ofs_destroy(ofs_grid *g)
{
    if (!g) return;
    for (int i = 0; i < 2; i++) { free(g->st[i]); free(g->live[i]); }
    free(g);
}

void ofs_read(const ofs_grid *g, uint32_t *cells)
{
    const uint32_t upr = orc_words_per_row(g->X);
    memset(cells, 0, (size_t)upr * g->Y * g->Z * sizeof(uint32_t));
    for (uint32_t z = 0; z < g->Z; z++)
        for (uint32_t y = 0; y < g->Y; y++) {
            uint32_t *row = cells + ((size_t)z * g->Y + y) * upr;
            const u64 *s = g->st[g->cur] + state_row(g, z, y);
            for (uint32_t x = 0; x < g->X; x++) {
                uint32_t v = 0;
                for (int b = 0; b < 4; b++) v |= (uint32_t)((s[b * g->nw + x / 64u] >> (x % 64u)) & 1u) << b;
                row[x / 8u] |= v << ((x % 8u) * 4u);
            }
        }
}

/* ---- one grid over threads ---- */
typedef struct {
    ofs_grid *g;
    const rule_t *r;
    uint32_t M, ticks;
    int n;
    orc_tick_stats *per_tick;
    pthread_barrier_t bar;
    pthread_mutex_t mu;  /* the gate: workers start once the team size is known */
    pthread_cond_t cv;
    int go;
    acc_t *slot[2];      /* per thread, by tick parity */
    atomic_int failed;
} grid_run;

typedef struct {
    grid_run *run;
    int tid;
} grid_worker_arg;

static void reduce_tick(grid_run *c, uint32_t t)
{
    acc_t a;
    acc_init(&a);
    for (int i = 0; i < c->n; i++) acc_merge(&a, &c->slot[t & 1u][i]);
    if (c->per_tick) acc_to_stats(&a, &c->per_tick[t]);
}

static void *grid_worker(void *p)
{
    grid_worker_arg *w = (grid_worker_arg *)p;
    grid_run *c = w->run;
    ofs_grid *g = c->g;
    pthread_mutex_lock(&c->mu);
    while (!c->go) pthread_cond_wait(&c->cv, &c->mu);
    pthread_mutex_unlock(&c->mu);
    const uint32_t z0 = (uint32_t)((uint64_t)g->Z * (uint64_t)w->tid / (uint64_t)c->n);
    const uint32_t z1 = (uint32_t)((uint64_t)g->Z * (uint64_t)(w->tid + 1) / (uint64_t)c->n);
    scratch_t s;
    int ok = scratch_alloc(&s, g) == 0;
    if (!ok) atomic_store(&c->failed, 1);
    live_from_state(g, c->M, z0, z1);
    pthread_barrier_wait(&c->bar);
    int cur = g->cur;
    for (uint32_t t = 0; t < c->ticks; t++) {
        if (w->tid == 0 && t > 0) reduce_tick(c, t - 1u);
        if (ok && !atomic_load(&c->failed)) run_slab(g, c->r, cur, z0, z1, &s, &c->slot[t & 1u][w->tid]);
        pthread_barrier_wait(&c->bar);
        cur ^= 1;
    }
    if (ok) scratch_free(&s);
    return NULL;
}

int ofs_run(ofs_grid *g, uint32_t M, uint32_t flags, uint64_t birth, uint64_t survival, uint32_t ticks,
            int n_threads, orc_tick_stats *per_tick)
{
    rule_t r;
    if (!g || rule_setup(&r, M, flags, birth, survival)) return -1;
    if (n_threads < 1) n_threads = 1;
    if ((uint32_t)n_threads > g->Z) n_threads = (int)g->Z;
    if (ticks == 0) return 0;
    grid_run c;
    memset(&c, 0, sizeof c);
    c.g = g; c.r = &r; c.M = M; c.ticks = ticks; c.n = n_threads; c.per_tick = per_tick;
    c.slot[0] = (acc_t *)calloc(2u * (size_t)n_threads, sizeof(acc_t));
    pthread_t *th = (pthread_t *)calloc((size_t)n_threads, sizeof(pthread_t));
    grid_worker_arg *args = (grid_worker_arg *)calloc((size_t)n_threads, sizeof(grid_worker_arg));
    if (!c.slot[0] || !th || !args) { free(c.slot[0]); free(th); free(args); return -2; }
    c.slot[1] = c.slot[0] + n_threads;
    atomic_init(&c.failed, 0);
    pthread_mutex_init(&c.mu, NULL);
    pthread_cond_init(&c.cv, NULL);
    int started = 1;
    for (int i = 0; i < n_threads; i++) {
        args[i].run = &c;
        args[i].tid = i;
    }
    for (int i = 1; i < n_threads; i++) {
        if (pthread_create(&th[i], NULL, grid_worker, &args[i]) != 0) break;
        started++;
    }
    /* the threads that started are the team */
    c.n = started;
    pthread_barrier_init(&c.bar, NULL, (unsigned)started);
    pthread_mutex_lock(&c.mu);
    c.go = 1;
    pthread_cond_broadcast(&c.cv);
    pthread_mutex_unlock(&c.mu);
    grid_worker(&args[0]);
    for (int i = 1; i < started; i++) pthread_join(th[i], NULL);
    reduce_tick(&c, ticks - 1u);
    g->cur = (int)((uint32_t)g->cur ^ (ticks & 1u));
    int rc = atomic_load(&c.failed) ? -2 : 0;
    pthread_barrier_destroy(&c.bar);
    pthread_mutex_destroy(&c.mu);
    pthread_cond_destroy(&c.cv);
    free(c.slot[0]);
    free(th);
    free(args);
    return rc;
}

/* ---- a list of simulations ---- */

/* searchManager.slang:98-149 with the per-slot rule of orc_run_rule: the
 * tick's snapshot folded into the grid info; status 1 when nothing survived
 * and nothing was born, 2 at the tick budget. */
static uint32_t manager(orc_grid_info *info, const orc_tick_stats *st, uint32_t max_ticks)
{
    info->numberOfSimulationTicksSurvived++;
    info->totalNumberOfCellsThatDied += st->snap.numberOfDeadCells;
    info->totalNumberOfCellsThatWereBorn += st->snap.numberOfCellsBorn;
    info->finalAliveCount = st->snap.numberOfSurvivingCells;
    if (st->snap.numberOfSurvivingCells > info->maxAliveCount) info->maxAliveCount = st->snap.numberOfSurvivingCells;
    if (st->snap.numberOfSurvivingCells == 0u && st->snap.numberOfCellsBorn == 0u) info->status = 1;
    else if (info->numberOfSimulationTicksSurvived >= max_ticks) info->status = 2;
    else info->status = 0;
    return info->status;
}

static uint32_t run_sim(const ofs_sim *sp, orc_grid_info *info, uint64_t *hash)
{
    memset(info, 0, sizeof *info);
    rule_t r;
    if (rule_setup(&r, sp->M, sp->flags, sp->birth, sp->survival)) return UINT32_MAX;
    uint32_t *cells = (uint32_t *)calloc(orc_cell_words(sp->X, sp->Y, sp->Z), sizeof(uint32_t));
    if (!cells) return UINT32_MAX;
    orc_init_cells(cells, sp->X, sp->Y, sp->Z, sp->M, sp->sx, sp->sy, sp->sz, sp->spawn_mode, sp->density, sp->seed);
    ofs_grid *g = ofs_create(sp->X, sp->Y, sp->Z, cells);
    free(cells);
    if (!g) return UINT32_MAX;
    scratch_t s;
    if (scratch_alloc(&s, g)) { ofs_destroy(g); return UINT32_MAX; }
    live_from_state(g, sp->M, 0, g->Z);
    uint32_t t = 0;
    for (;;) {
        acc_t a;
        orc_tick_stats st;
        run_slab(g, &r, g->cur, 0, g->Z, &s, &a);
        g->cur ^= 1;
        acc_to_stats(&a, &st);
        t++;
        if (manager(info, &st, sp->max_ticks) > 0 || t >= 0x7fffffffu) break;
    }
    scratch_free(&s);
    if (hash) {  /* FNV-1a-64 of the final cells, as orc_canonical_hash */
        uint32_t *fin = (uint32_t *)calloc(orc_cell_words(sp->X, sp->Y, sp->Z), sizeof(uint32_t));
        if (!fin) { ofs_destroy(g); return UINT32_MAX; }
        ofs_read(g, fin);
        *hash = orc_canonical_hash(fin, sp->X, sp->Y, sp->Z);
        free(fin);
    }
    ofs_destroy(g);
    return t;
}

typedef struct {
    const ofs_sim *sims;
    size_t n;
    atomic_size_t next;
    atomic_size_t done;
    atomic_uint_fast64_t updates;
    struct timespec t0;
    double budget_s;
    uint32_t *ticks_out;
    orc_grid_info *info_out;
    uint64_t *hash_out;
} sims_run;

static double since(const struct timespec *t0)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)(t.tv_sec - t0->tv_sec) + 1e-9 * (double)(t.tv_nsec - t0->tv_nsec);
}

static void *sims_worker(void *p)
{
    sims_run *c = (sims_run *)p;
    for (;;) {
        if (c->budget_s > 0 && since(&c->t0) >= c->budget_s) break;
        size_t i = atomic_fetch_add(&c->next, 1);
        if (i >= c->n) break;
        orc_grid_info info;
        uint32_t t = run_sim(&c->sims[i], &info, c->hash_out ? &c->hash_out[i] : NULL);
        c->ticks_out[i] = t;
        if (c->info_out) c->info_out[i] = info;
        if (t != UINT32_MAX) {
            atomic_fetch_add(&c->done, 1);
            atomic_fetch_add(&c->updates, (uint64_t)t * c->sims[i].X * c->sims[i].Y * c->sims[i].Z);
        }
    }
    return NULL;
}

size_t ofs_run_sims(const ofs_sim *sims, size_t n, int n_threads, double budget_s, uint32_t *ticks_out,
                    orc_grid_info *info_out, uint64_t *cell_updates)
{
    return ofs_run_sims_hash(sims, n, n_threads, budget_s, ticks_out, info_out, cell_updates, NULL);
}

size_t ofs_run_sims_hash(const ofs_sim *sims, size_t n, int n_threads, double budget_s, uint32_t *ticks_out,
                         orc_grid_info *info_out, uint64_t *cell_updates, uint64_t *hash_out)
{
    sims_run c;
    memset(&c, 0, sizeof c);
    c.sims = sims;
    c.n = n;
    c.budget_s = budget_s;
    c.ticks_out = ticks_out;
    c.info_out = info_out;
    c.hash_out = hash_out;
    atomic_init(&c.next, 0);
    atomic_init(&c.done, 0);
    atomic_init(&c.updates, 0);
    for (size_t i = 0; i < n; i++) ticks_out[i] = UINT32_MAX;
    clock_gettime(CLOCK_MONOTONIC, &c.t0);
    if (n_threads < 1) n_threads = 1;
    pthread_t *th = (pthread_t *)calloc((size_t)n_threads, sizeof(pthread_t));
    int started = 1;
    if (th)
        for (int i = 1; i < n_threads; i++) {
            if (pthread_create(&th[i], NULL, sims_worker, &c) != 0) break;
            started++;
        }
    sims_worker(&c);
    for (int i = 1; i < started; i++) pthread_join(th[i], NULL);
    free(th);
    if (cell_updates) *cell_updates = (uint64_t)atomic_load(&c.updates);
    return (size_t)atomic_load(&c.done);
}
