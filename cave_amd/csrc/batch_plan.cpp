// batch_plan.cpp -- layout of a simulation list on the persistent batched
// kernel, the autotuner behind it, and the launch itself.
//
// Autotuner.  Throughput of the batched kernel depends on the grid size and
// neighbourhood of the simulations (bucketed by grid edge and F/E/C flags), on
// the block size (which also decides which grids fit the LDS: 1024-thread
// blocks get a whole CU's LDS, 512 half of it, 256 a quarter), and on how
// finely a tick is cut into work items (items per thread).  For each bucket
// present in a batch and each candidate (block size, items per thread) the
// tuner times one short calibration launch in which every resident block runs
// a grid of that bucket, and keeps the device-wide rate (cell-updates/s) in a
// process-wide table.  The planner then picks the block size with the lowest
// predicted time for the batch's bucket mix, the items per thread per bucket,
// predicts every simulation's single-block time from its bucket's rate, splits
// the simulations that would outlast the batch's average block load over
// several blocks, and orders the queue (split simulations first, so that all
// of their parts start together, then longest first).
#include "batch_plan.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>

#include "host_util.h"

namespace cave {

uint64_t fnv_ref_layout(const uint32_t *cells, uint32_t X, uint32_t Y, uint32_t Z);

namespace {

constexpr uint32_t kLdsBytesPerCU = 160u * 1024u;
constexpr uint32_t kStaticLds = 1024u;    // s_sum / s_box / scalars of k_batch
constexpr uint64_t kPartWaitSeconds = 4;  // a split simulation's parts give up waiting after this
constexpr uint32_t kThreadChoices[] = {256u, 512u, 1024u};
constexpr uint32_t kIptChoices[] = {1u, 2u};
constexpr uint32_t kDefaultMaxParts = 32u;  // the uneven 2048-config search: 16 parts 20.5-20.8 ms, 32 parts 20.1-20.2 ms
constexpr double kTuneMargin = 1.03;  // a tuned choice must beat the default by 3%

struct TuneKey {
    uint32_t bucket, flags, threads, ipt;
    bool operator<(const TuneKey &o) const
    {
        return std::tie(bucket, flags, threads, ipt) < std::tie(o.bucket, o.flags, o.threads, o.ipt);
    }
};

std::mutex g_tune_mu;
std::map<TuneKey, double> g_rates;  // device-wide cell-updates/s, every resident block on the bucket

struct DevBuf {
    void *p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
    template <typename T>
    T *as() const { return static_cast<T *>(p); }
};

int dev_alloc(DevBuf &b, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        b.p = nullptr;
        return fail(e == hipErrorOutOfMemory ? CAVE_E_NOMEM : CAVE_E_DEVICE,
                    std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    return CAVE_OK;
}

double split_divisor()
{
    const char *e = std::getenv("CAVE_SPLIT_DIV");
    const double d = e ? std::atof(e) : 8.0;
    return d > 0.0 ? d : 8.0;
}

bool tuner_enabled()
{
    const char *e = std::getenv("CAVE_AUTOTUNE");
    return !(e && e[0] == '0');
}

// Simulations in a calibration launch on the list's own specs: eight per
// resident block of the smallest block size, so that every block size runs
// the same simulations and each block several of them -- with one each, the
// launch would time the longest-lived simulation, not the throughput.
size_t sample_copies()
{
    return (size_t)cu_count() * (1024u / kThreadChoices[0]) * 8u;
}

// Largest calibration launch on the list's own specs, in cell-updates: below
// it a bucket's rates are measured on its own simulations (49^3 x 20 ticks x
// 8192 copies is 1.9e10, ~3 ms), above it on the synthetic grid.
constexpr double kSampleBudget = 2.5e10;
constexpr size_t kSampleMax = 1024;  // spec indices kept per bucket

// One calibration launch.  With `sample` (indices into `specs` of the
// bucket's simulations), every resident block runs one of them, cycling
// through the sample: their rules, spawns, tick budgets and flags (final
// hashes off), so the rate includes what the synthetic grid does not --
// active regions, early deaths, per-tick fixed costs -- and a block size that
// runs several of them per CU is seen at its worth.  The rate counts each
// copy's budgeted ticks, as the planner's costs do.  Without a sample: every
// resident block runs a copy of a grid of the bucket (all-counts rule, no
// early exit, every cell, a few ticks).
int calibrate(const TuneKey &key, const uint32_t dims[3], hipStream_t st, double &rate,
              const CaveSimSpec *sample_specs = nullptr, const std::vector<uint32_t> *sample = nullptr)
{
    if (sample && !sample->empty()) {
        // the same simulations for every block size and items per thread:
        // the sample, repeated to sample_copies()
        const size_t copies = std::max<size_t>(sample->size(), sample_copies());
        std::vector<CaveSimSpec> specs(copies);
        double work = 0;
        for (size_t i = 0; i < copies; i++) {
            specs[i] = sample_specs[(*sample)[i % sample->size()]];
            specs[i].flags &= ~(uint32_t)CAVE_SIM_FINAL_HASH;
            work += (double)specs[i].dims[0] * specs[i].dims[1] * specs[i].dims[2] * specs[i].ticks;
        }
        BatchKnobs k;
        k.threads = key.threads;
        k.ipt = key.ipt;
        k.max_parts = 1;
        k.tune = false;
        BatchPlan plan;
        int rc = plan_batch(specs.data(), copies, k, plan);
        if (rc) return rc;
        std::vector<CaveSimResult> res(copies);
        CaveBatchOptions opt{};
        if ((rc = run_plan(plan, copies, res.data(), &opt, st))) return rc;
        rate = work / (std::max(opt.device_ms, 1e-3) * 1e-3);
        return CAVE_OK;
    }
    const uint32_t copies = cu_count() * std::max(1u, 1024u / key.threads);
    const uint64_t cells = (uint64_t)dims[0] * dims[1] * dims[2];
    const bool erd = (key.flags & kTuneErd) != 0u;
    // ERD: 12 face and 4 edge neighbours (searchSimulationERD.slang:81-97)
    const uint32_t N = erd ? ((key.flags & 1u) ? 12u : 0u) + ((key.flags & 2u) ? 4u : 0u) : neighbor_count(key.flags & 7u);
    const uint64_t all = N ? (N >= 64 ? ~0ull : ((1ull << N) - 1)) : 1ull;
    const double target = 2e9;
    const uint32_t ticks = (uint32_t)std::clamp<double>(target / ((double)copies * (double)cells), 2.0, 64.0);
    std::vector<CaveSimSpec> specs(copies);
    for (uint32_t i = 0; i < copies; i++) {
        CaveSimSpec &s = specs[i];
        std::memset(&s, 0, sizeof s);
        for (int a = 0; a < 3; a++) {
            s.dims[a] = dims[a];
            s.spawn[a] = std::max(1u, dims[a] / 2u);
        }
        s.max_cell_state = 7;
        // ERD keys: bit 2 (corners, which ERD cells do not have) marks wrap
        s.neighborhood_flags = erd ? (key.flags & 3u) | ((key.flags & 4u) ? CAVE_WRAP : 0u) : key.flags & 7u;
        s.ticks = ticks;
        s.birth = all;
        s.survival = all;
        s.spawn_mode = CAVE_SPAWN_RANDOM;
        s.density = 0.25;
        s.seed = i;
        s.flags = CAVE_SIM_NO_EARLY_EXIT | CAVE_SIM_FULL_GRID | (erd ? CAVE_SIM_SHAPE_ERD : 0u);
    }
    BatchKnobs k;
    k.threads = key.threads;
    k.ipt = key.ipt;
    k.max_parts = 1;
    k.tune = false;
    BatchPlan plan;
    int rc = plan_batch(specs.data(), copies, k, plan);
    if (rc) return rc;
    std::vector<CaveSimResult> res(copies);
    CaveBatchOptions opt{};
    if ((rc = run_plan(plan, copies, res.data(), &opt, st))) return rc;
    rate = (double)copies * (double)cells * ticks / (std::max(opt.device_ms, 1e-3) * 1e-3);
    return CAVE_OK;
}

struct BucketWork {
    uint32_t dims[3];  // a representative grid of the bucket
    double work = 0;   // budgeted cell-updates
    std::vector<uint32_t> sample;  // its first simulations (up to kSampleMax)
    double sample_work = 0;        // their budgeted cell-updates
    uint32_t rate_flags = 0;       // the rate table's flags for the bucket (kTuneSample when sampled)
};

}  // namespace

uint64_t wall_clock_khz()
{
    static int khz = 0;
    if (!khz && (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, 0) != hipSuccess || khz <= 0))
        khz = 100000;
    return (uint64_t)khz;
}

uint32_t lds_cap(uint32_t threads)
{
    const uint32_t per = std::max(1u, 1024u / threads);
    return kLdsBytesPerCU / per - kStaticLds;
}

double bucket_rate(uint32_t bucket, uint32_t flags, uint32_t threads, uint32_t ipt, const uint32_t dims[3],
                   bool measure, hipStream_t st)
{
    const TuneKey key{bucket, flags & (kTuneErd | 7u), threads, ipt};
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        auto it = g_rates.find(key);
        if (it != g_rates.end()) return it->second;
    }
    if (!measure || !tuner_enabled()) return 0.0;
    double r = 0;
    if (calibrate(key, dims, st, r)) return 0.0;
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_rates[key] = r;
    return r;
}

uint32_t cu_count()
{
    static int cus = 0;
    if (!cus) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, 0) != hipSuccess) return 256;
        cus = p.multiProcessorCount;
    }
    return (uint32_t)cus;
}

// Blocks per CU that stay resident, for the form of the batch kernel (of
// the plan's cell shape) with the lower occupancy: a plan is sized before it
// knows whether it splits a job, which picks the form launched.
uint32_t resident_per_cu(uint32_t threads, uint32_t lds_bytes, uint32_t shape)
{
    int best = 1 << 30;
    for (bool split : {false, true}) {
        int per = 0;
        const void *k = batch_kernel(split, shape);
        if (lds_bytes > 64u * 1024u) (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k, (int)threads, lds_bytes) != hipSuccess ||
            per < 1)
            per = 1;
        best = std::min(best, per);
    }
    return (uint32_t)best;
}

uint32_t spec_shape(const CaveSimSpec &s)
{
    if (!(s.flags & CAVE_SIM_SHAPE_ERD)) return kBatchShapeCube;
    return (s.neighborhood_flags & CAVE_WRAP) ? kBatchShapeErdWrap : kBatchShapeErd;
}

uint32_t tune_kind(const DevJob &j, uint32_t shape)
{
    if (shape == kBatchShapeCube) return j.flags & 7u;
    return kTuneErd | (shape == kBatchShapeErdWrap ? 4u : 0u) | (j.flags & 3u);
}

uint32_t edge_bucket(uint32_t X, uint32_t Y, uint32_t Z)
{
    const double e = std::cbrt((double)X * Y * Z);
    if (e < 32) return 0;
    if (e < 48) return 1;
    if (e < 64) return 2;
    if (e < 96) return 3;
    if (e < 128) return 4;
    return 5;
}

DevJob make_job(const CaveSimSpec &s, uint32_t slot)
{
    DevJob j{};
    j.X = s.dims[0];
    j.Y = s.dims[1];
    j.Z = s.dims[2];
    // ERD cells have no corner neighbours and the persistent kernel runs ERD
    // grids without wrap: only F|E matter there
    j.flags = s.neighborhood_flags & ((s.flags & CAVE_SIM_SHAPE_ERD) ? 3u : 7u);
    j.M = s.max_cell_state;
    const DevRule r = make_rule(s.max_cell_state, s.neighborhood_flags, s.birth, s.survival);
    j.born_lut = r.born_lut;
    j.surv_lut = r.surv_lut;
    j.ticks = s.ticks;
    j.sim_flags = s.flags;
    j.result_slot = slot;
    j.parts = 1;
    j.init = make_init(s.dims, s.spawn, s.spawn_mode, s.density, s.seed);
    return j;
}

int plan_batch(const CaveSimSpec *specs, size_t n, const BatchKnobs &k, BatchPlan &plan)
{
    plan = BatchPlan{};
    plan.shape = n > 0 ? spec_shape(specs[0]) : kBatchShapeCube;
    for (size_t i = 0; i < n; i++)
        if (spec_shape(specs[i]) != plan.shape) return fail(CAVE_E_INVALID, "a persistent launch runs one cell shape");
    plan.jobs.resize(n);
    plan.ipt.resize(n);
    std::vector<uint64_t> words(n);
    std::vector<double> cost(n);
    std::vector<uint32_t> key_bucket(n);
    std::map<std::pair<uint32_t, uint32_t>, BucketWork> buckets;  // (edge bucket, flags)
    // Lists hold long runs of one grid size and one bucket (the seeds of a
    // config): per-size values and the bucket's entry are carried over from
    // the previous spec instead of recomputed (map entries do not move).
    // the tuner's bucket kind, with jobs that compute every cell apart
    auto plan_kind = [&](const DevJob &j) {
        return tune_kind(j, plan.shape) | ((j.sim_flags & CAVE_SIM_FULL_GRID) ? kTuneFull : 0u);
    };
    uint32_t prev_dims[3] = {0u, 0u, 0u}, prev_bucket = 0u;
    uint64_t prev_words = 0u;
    std::pair<uint32_t, uint32_t> prev_key{~0u, ~0u};
    BucketWork *prev_b = nullptr;
    for (size_t i = 0; i < n; i++) {
        DevJob &j = plan.jobs[i];
        j = make_job(specs[i], (uint32_t)i);
        if (j.X != prev_dims[0] || j.Y != prev_dims[1] || j.Z != prev_dims[2]) {
            prev_dims[0] = j.X;
            prev_dims[1] = j.Y;
            prev_dims[2] = j.Z;
            prev_words = job_words(j.X, j.Y, j.Z);
            prev_bucket = edge_bucket(j.X, j.Y, j.Z);
        }
        words[i] = prev_words;
        cost[i] = (double)j.X * j.Y * j.Z * j.ticks;
        key_bucket[i] = prev_bucket;
        const std::pair<uint32_t, uint32_t> key{prev_bucket, plan_kind(j)};
        if (key != prev_key) {
            prev_key = key;
            prev_b = &buckets[key];
        }
        BucketWork &b = *prev_b;
        if (b.work == 0)
            for (int a = 0; a < 3; a++) b.dims[a] = specs[i].dims[a];
        b.work += cost[i];
        if (b.sample.size() < kSampleMax) {
            b.sample.push_back((uint32_t)i);
            b.sample_work += cost[i];
        }
        if (specs[i].flags & CAVE_SIM_FINAL_HASH) {
            j.has_final = 1u;
            j.final_addr = plan.final_words;  // word offset until run_plan places the staging
            plan.final_words += cave_cell_words(j.X, j.Y, j.Z);
        }
    }

    // ---- block size and items per thread from the rate table ----
    const bool tune = k.tune && tuner_enabled();
    std::vector<uint32_t> cand;
    if (k.threads) cand.push_back(std::max(64u, std::min(1024u, (k.threads + 63u) / 64u * 64u)));
    else if (tune) cand.assign(std::begin(kThreadChoices), std::end(kThreadChoices));
    else cand.push_back(1024u);
    std::vector<uint32_t> ipts;  // the first entry is the default
    if (k.ipt) ipts.push_back(k.ipt);
    else if (tune) ipts.assign(std::begin(kIptChoices), std::end(kIptChoices));
    else ipts.push_back(1u);

    auto rate_of = [&](uint32_t bucket, uint32_t flags, uint32_t T, uint32_t ipt) -> double {
        std::lock_guard<std::mutex> g(g_tune_mu);
        auto it = g_rates.find(TuneKey{bucket, flags, T, ipt});
        return it == g_rates.end() ? 0.0 : it->second;
    };
    // A bucket is calibrated on its own simulations when a launch of them on
    // every resident block of the smallest block size stays within
    // kSampleBudget (the small grids, where per-tick fixed costs and active
    // regions decide the block size), else on the synthetic grid.
    for (auto &[bk, bw] : buckets) {
        const size_t m = std::max<size_t>(1, bw.sample.size());
        const double launch = bw.sample_work / (double)m * (double)std::max(m, sample_copies());
        bw.rate_flags = bk.second | (launch <= kSampleBudget ? kTuneSample : 0u);
    }
    if (tune) {
        hipStream_t cst = nullptr;  // created on the first calibration this plan needs
        int rc = CAVE_OK;
        auto measure = [&](uint32_t bucket, const BucketWork &bw, uint32_t T, uint32_t ipt) {
            const TuneKey key{bucket, bw.rate_flags, T, ipt};
            if (rc || rate_of(key.bucket, key.flags, T, ipt) > 0) return;
            if (!cst && (rc = hipStreamCreateWithFlags(&cst, hipStreamNonBlocking) == hipSuccess ? CAVE_OK
                                                                                                   : CAVE_E_DEVICE))
                return;
            double r = 0;
            rc = (bw.rate_flags & kTuneSample) ? calibrate(key, bw.dims, cst, r, specs, &bw.sample)
                                               : calibrate(key, bw.dims, cst, r);
            if (!rc) {
                std::lock_guard<std::mutex> g(g_tune_mu);
                g_rates[key] = r;
            }
        };
        // every block size at the default items per thread, then the other
        // items-per-thread choices at the block size that measured best
        for (const auto &[bk, bw] : buckets) {
            uint32_t bestT = cand.back();
            double bestr = -1;
            for (uint32_t T : cand) {
                measure(bk.first, bw, T, ipts[0]);
                const double r = rate_of(bk.first, bw.rate_flags, T, ipts[0]);
                if (r > bestr) bestr = r, bestT = T;
            }
            for (size_t q = 1; q < ipts.size(); q++) measure(bk.first, bw, bestT, ipts[q]);
        }
        if (cst) {
            (void)hipStreamSynchronize(cst);
            (void)hipStreamDestroy(cst);
        }
        if (rc) return rc;
    }
    // a candidate replaces the default (the first one listed) only when it
    // measured kTuneMargin faster: calibration noise does not flip choices
    auto best_ipt = [&](uint32_t bucket, uint32_t flags, uint32_t T, double *rate) {
        uint32_t bi = ipts[0];
        double br = rate_of(bucket, flags, T, ipts[0]);
        for (uint32_t ipt : ipts) {
            const double r = rate_of(bucket, flags, T, ipt);
            if (r > br * kTuneMargin) {
                br = r;
                bi = ipt;
            }
        }
        if (rate) *rate = br;
        return bi;
    };
    uint32_t threads = cand.back();  // the largest block size is the default
    if (tune && cand.size() > 1) {
        auto predict = [&](uint32_t T) {
            double t = 0;
            for (const auto &[bk, bw] : buckets) {
                double r = 0;
                best_ipt(bk.first, bw.rate_flags, T, &r);
                t += bw.work / std::max(r, 1.0);
            }
            return t;
        };
        double best = predict(threads);
        for (uint32_t T : cand) {
            const double t = predict(T);
            if (t * kTuneMargin < best) {
                best = t;
                threads = T;
            }
        }
    }
    plan.threads = threads;

    // ---- storage tiers, residency ----
    const uint32_t cap = lds_cap(threads);
    uint64_t lds_words = 0;
    for (size_t i = 0; i < n; i++) {
        DevJob &j = plan.jobs[i];
        if (words[i] * 4u <= cap) {
            j.tier = kTierLds;
            lds_words = std::max(lds_words, words[i]);
        } else {
            j.tier = kTierBlockArena;
        }
    }
    plan.lds_bytes = (uint32_t)(lds_words * 4u);
    const uint32_t per_cu = resident_per_cu(threads, plan.lds_bytes, plan.shape);
    const uint32_t resident = cu_count() * per_cu;

    // ---- single-block time estimates, splitting, order ----
    // items per thread and rate per bucket, looked up in the (locked) rate
    // table once per bucket rather than once per simulation
    std::map<std::pair<uint32_t, uint32_t>, std::pair<uint32_t, double>> bucket_choice;
    for (const auto &[bk, bw] : buckets) {
        double r = 0;
        const uint32_t ipt = best_ipt(bk.first, bw.rate_flags, threads, &r);
        bucket_choice[bk] = {ipt, r};
    }
    std::vector<double> t(n);
    std::vector<uint32_t> ipt_of(n, ipts[0]);
    double total = 0;
    prev_key = {~0u, ~0u};
    std::pair<uint32_t, double> choice{ipts[0], 0.0};
    for (size_t i = 0; i < n; i++) {
        const DevJob &j = plan.jobs[i];
        const std::pair<uint32_t, uint32_t> key{key_bucket[i], plan_kind(j)};
        if (key != prev_key) {
            prev_key = key;
            choice = bucket_choice[key];
        }
        ipt_of[i] = choice.first;
        const double r = choice.second;
        // a bucket's device-wide rate is shared by `resident` blocks
        t[i] = r > 0 ? cost[i] * resident / r : cost[i];
        total += t[i];
    }
    const double share = total / resident / split_divisor();
    uint32_t max_parts = k.max_parts ? k.max_parts : kDefaultMaxParts;
    max_parts = std::min(max_parts, std::max(1u, resident / 4u));
    std::vector<std::tuple<int, double, size_t>> order(n);  // (unsplit, per-part time, job)
    double longest_part = 0;
    uint32_t zc_key[5] = {0u, 0u, 0u, 0u, 0u}, zc_prev = 0u;  // part_zchunks of the previous (X, Y, Z, ipt, parts)
    for (size_t i = 0; i < n; i++) {
        DevJob &j = plan.jobs[i];
        uint32_t parts = 1;
        if (j.tier != kTierLds && max_parts > 1 && t[i] > share) {
            const double want = std::ceil(t[i] / share);
            parts = (uint32_t)std::min<double>(want, (double)std::min(max_parts, std::max(1u, j.Z / 4u)));
            parts = std::max(1u, parts);
        }
        j.parts = parts;
        if (j.X != zc_key[0] || j.Y != zc_key[1] || j.Z != zc_key[2] || ipt_of[i] != zc_key[3] || parts != zc_key[4]) {
            zc_key[0] = j.X;
            zc_key[1] = j.Y;
            zc_key[2] = j.Z;
            zc_key[3] = ipt_of[i];
            zc_key[4] = parts;
            zc_prev = part_zchunks(j.X, j.Y, j.Z, threads, ipt_of[i], parts);
        }
        j.zchunks = zc_prev;
        plan.ipt[i] = ipt_of[i];
        if (parts > 1) {
            j.tier = kTierOwnRegion;
            j.region = plan.region_words;  // word offset until run_plan places the regions
            plan.region_words += (words[i] + 63u) / 64u * 64u;
            j.sync_slot = plan.nsync++;
        } else if (j.tier == kTierBlockArena) {
            plan.arena_words = std::max(plan.arena_words, (words[i] + 63u) / 64u * 64u);  // 256-byte block strides
        }
        order[i] = {parts > 1 ? 0 : 1, t[i] / parts, i};
        longest_part = std::max(longest_part, t[i] / parts);
    }
    const auto claim_order = [](const auto &x, const auto &y) {
        if (std::get<0>(x) != std::get<0>(y)) return std::get<0>(x) < std::get<0>(y);
        return std::get<1>(x) > std::get<1>(y);
    };
    // a list already in claim order (e.g. configs of one size and budget) is kept as is
    if (!std::is_sorted(order.begin(), order.end(), claim_order))
        std::stable_sort(order.begin(), order.end(), claim_order);
    plan.queue.reserve(n);
    for (const auto &o : order) {
        const DevJob &j = plan.jobs[std::get<2>(o)];
        for (uint32_t p = 0; p < j.parts; p++) {
            DevJob e = j;
            e.part = p;
            plan.queue.push_back(e);
        }
    }
    plan.blocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(resident, plan.queue.size()));
    if (tune) plan.predicted_ms = std::max(total / resident, longest_part) * 1e3;
    return CAVE_OK;
}

int run_plan(const BatchPlan &plan, size_t n, CaveSimResult *results, CaveBatchOptions *opt, hipStream_t st)
{
    const size_t nq = plan.queue.size();
    DevBuf djobs, dnext, dres, dpt, darena, dfinal, dregion, dsync;
    int rc;
    if ((rc = dev_alloc(djobs, sizeof(DevJob) * nq))) return rc;
    if ((rc = dev_alloc(dnext, sizeof(uint32_t)))) return rc;
    if ((rc = dev_alloc(dres, sizeof(DevResult) * n))) return rc;
    const uint32_t stride = opt ? opt->per_tick_stride : 0u;
    const bool want_pt = opt && opt->per_tick && stride;
    if (want_pt) {  // rows past a simulation's ticks_run come back zero
        if ((rc = dev_alloc(dpt, sizeof(CaveTickStats) * n * stride))) return rc;
        CAVE_HIP(hipMemsetAsync(dpt.p, 0, sizeof(CaveTickStats) * n * stride, st));
    }
    if (plan.arena_words && (rc = dev_alloc(darena, sizeof(uint32_t) * plan.arena_words * plan.blocks))) return rc;
    if (plan.region_words && (rc = dev_alloc(dregion, sizeof(uint32_t) * plan.region_words))) return rc;
    if (plan.nsync) {
        if ((rc = dev_alloc(dsync, sizeof(JobSync) * plan.nsync))) return rc;
        CAVE_HIP(hipMemsetAsync(dsync.p, 0, sizeof(JobSync) * plan.nsync, st));
    }
    if (plan.final_words && (rc = dev_alloc(dfinal, sizeof(uint32_t) * plan.final_words))) return rc;
    // word offsets -> device addresses, in a copy of the queue only when some
    // entry holds one
    std::vector<DevJob> queue;
    const DevJob *qsrc = plan.queue.data();
    if (plan.region_words || plan.final_words) {
        queue = plan.queue;
        for (DevJob &e : queue) {
            if (e.tier == kTierOwnRegion) e.region = (uint64_t)(dregion.as<uint32_t>() + e.region);
            if (e.has_final) e.final_addr = (uint64_t)(dfinal.as<uint32_t>() + e.final_addr);
        }
        qsrc = queue.data();
    }
    CAVE_HIP(hipMemcpyAsync(djobs.p, qsrc, sizeof(DevJob) * nq, hipMemcpyHostToDevice, st));
    CAVE_HIP(hipMemsetAsync(dnext.p, 0, sizeof(uint32_t), st));

    BatchArgs a{};
    a.jobs = djobs.as<DevJob>();
    a.njobs = (uint32_t)nq;
    a.next_job = dnext.as<uint32_t>();
    a.results = dres.as<DevResult>();
    a.per_tick = want_pt ? dpt.as<CaveTickStats>() : nullptr;
    a.per_tick_stride = stride;
    a.arena = darena.as<uint32_t>();
    a.arena_words = plan.arena_words;
    a.lds_words = plan.lds_bytes / 4u;
    a.sync = dsync.as<JobSync>();
    a.wait_limit = wall_clock_khz() * 1000u * kPartWaitSeconds;

    hipEvent_t e0, e1;
    CAVE_HIP(hipEventCreate(&e0));
    CAVE_HIP(hipEventCreate(&e1));
    CAVE_HIP(hipEventRecord(e0, st));
    rc = launch_batch(a, plan.blocks, plan.threads, plan.lds_bytes, st, plan.nsync > 0, plan.shape);
    if (rc) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return rc;
    }
    CAVE_HIP(hipEventRecord(e1, st));
    std::vector<DevResult> hres(n);
    CAVE_HIP(hipMemcpyAsync(hres.data(), dres.p, sizeof(DevResult) * n, hipMemcpyDeviceToHost, st));
    if (want_pt)
        CAVE_HIP(hipMemcpyAsync(opt->per_tick, dpt.p, sizeof(CaveTickStats) * n * stride, hipMemcpyDeviceToHost, st));
    std::vector<uint32_t> hfinal(plan.final_words);
    if (plan.final_words)
        CAVE_HIP(hipMemcpyAsync(hfinal.data(), dfinal.p, sizeof(uint32_t) * plan.final_words, hipMemcpyDeviceToHost,
                                st));
    CAVE_HIP(hipStreamSynchronize(st));
    float ms = 0.f;
    CAVE_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (opt) opt->device_ms = ms;

    for (size_t i = 0; i < n; i++)
        if (hres[i].err)
            return fail(CAVE_E_STATE, "the parts of a split simulation did not meet within the wait limit");
    for (size_t i = 0; i < n; i++) {
        CaveSimResult &r = results[i];
        std::memset(&r, 0, sizeof r);
        r.info = hres[i].info;
        r.classification = hres[i].classification;
        r.ticks_run = hres[i].ticks_run;
        r.last = hres[i].last;
    }
    if (opt && opt->timing) {
        uint64_t t0 = ~0ull;
        for (size_t i = 0; i < n; i++) t0 = std::min(t0, hres[i].t_start);
        const double us_per_tick = 1000.0 / (double)wall_clock_khz();
        for (size_t i = 0; i < n; i++) {
            CaveSimTiming &tm = opt->timing[i];
            tm.start_us = (double)(hres[i].t_start - t0) * us_per_tick;
            tm.end_us = (double)(hres[i].t_end - t0) * us_per_tick;
            tm.blocks = plan.jobs[i].parts;
            tm.items_per_thread = plan.ipt[i];
            tm.computed_cells = hres[i].computed;
        }
    }
    for (const DevJob &j : plan.jobs)
        if (j.has_final)
            results[j.result_slot].final_hash = fnv_ref_layout(hfinal.data() + j.final_addr, j.X, j.Y, j.Z);
    return CAVE_OK;
}

std::string tuning_report()
{
    std::lock_guard<std::mutex> g(g_tune_mu);
    std::string s = "[";
    bool first = true;
    for (const auto &[k, r] : g_rates) {
        if (!first) s += ",";
        first = false;
        s += "{\"edge_bucket\":" + std::to_string(k.bucket) + ",\"flags\":" + std::to_string(k.flags) +
             ",\"threads\":" + std::to_string(k.threads) + ",\"items_per_thread\":" + std::to_string(k.ipt) +
             ",\"cell_updates_per_s\":" + std::to_string(r) + "}";
    }
    return s + "]";
}

}  // namespace cave
