// api_serve.cpp -- cave_serve_*: continuous batching of rule-evaluation requests.
//
// One resident persistent kernel (k_serve) runs while requests are in flight.
// The host keeps a FIFO of submitted requests and admits them in order while
// the HBM budget, the in-flight cap and the ring allow: admission reserves the
// request's own region for its packed state (global-tier grids) from a
// budgeted HBM arena and, when a final hash is asked for, a region of pinned
// host memory that the kernel writes the final cells into, writes
// its ring entries (one per part of a split grid) into host-coherent memory
// and publishes them by raising the ring tail.  Free blocks of the resident
// kernel claim entries as soon as they are published, so requests that arrive
// while others run start on the next free block, next to the ones already
// running.  Completion is a flag per ring slot in host-coherent memory; the
// host polls it, reads the result record, releases the request's memory and
// admits what now fits.  The kernel closes itself after a stretch without
// claims; the host relaunches it when new entries are published.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <thread>
#include <vector>

#include "batch.h"
#include "batch_plan.h"
#include "common.h"
#include "host_util.h"

namespace cave {

uint64_t fnv_ref_layout(const uint32_t *cells, uint32_t X, uint32_t Y, uint32_t Z);
int validate_sim(const CaveSimSpec &s);

namespace {

constexpr uint64_t kAlign = 256;
constexpr uint64_t kChunkBytes = 1ull << 30;
constexpr uint64_t kFinalChunkBytes = 64ull << 20;  // pinned host staging of final cells
constexpr uint64_t kFinalBudget = 4ull << 30;
constexpr uint64_t kRingCap = 1ull << 16;
constexpr uint32_t kSyncPool = 1u << 17;
constexpr uint32_t kThreads = 1024;
constexpr uint64_t kDefaultBudget = 200ull << 30;
constexpr uint32_t kDefaultInflight = 16384;
constexpr uint32_t kDefaultSplitUs = 1000;
constexpr uint32_t kDefaultIdleUs = 50000;
constexpr uint32_t kDefaultMaxParts = 64;
constexpr double kFallbackBlockRate = 1e10;  // cell-updates/s of one block when a bucket was never measured

uint64_t align_up(uint64_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// Memory for resident requests: lazily allocated chunks (or one chunk of the
// request's size when larger), first-fit with coalescing free lists, and a
// byte budget over live reservations.  Addresses are device addresses.
//   * HBM (host = false): request state, 1 GiB chunks.
//   * pinned host memory (host = true): the final cells of requests that ask
//     for them.  The kernel writes them straight into host-coherent memory,
//     so reporting them needs no copy: a copy engine may fall back to a copy
//     kernel, which cannot start beside the resident kernel that fills the
//     GPU, and waiting for it would stall the request loop until that
//     kernel closed.
class Arena {
public:
    Arena(uint64_t budget, bool host, uint64_t chunk_bytes) : budget_(budget), host_(host), chunk_bytes_(chunk_bytes) {}
    ~Arena()
    {
        for (Chunk &c : chunks_) {
            if (host_) (void)hipHostFree(c.hbase);
            else (void)hipFree(c.base);
        }
    }
    uint64_t used() const { return used_; }
    uint64_t budget() const { return budget_; }

    // 0 on success; CAVE_E_BUDGET when the live reservations would exceed the budget
    int alloc(uint64_t bytes, uint64_t &addr)
    {
        bytes = align_up(bytes);
        if (used_ + bytes > budget_) return CAVE_E_BUDGET;
        for (Chunk &c : chunks_)
            if (take(c, bytes, addr)) {
                used_ += bytes;
                return CAVE_OK;
            }
        const uint64_t size = std::max(chunk_bytes_, bytes);
        Chunk c;
        if (host_) {
            if (hipHostMalloc(&c.hbase, size, hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess)
                return CAVE_E_NOMEM;
            if (hipHostGetDevicePointer(&c.base, c.hbase, 0) != hipSuccess || !c.base) {
                (void)hipHostFree(c.hbase);
                return CAVE_E_NOMEM;
            }
        } else if (hipMalloc(&c.base, size) != hipSuccess) {
            return CAVE_E_NOMEM;
        }
        c.size = size;
        c.free[0] = size;
        chunks_.push_back(std::move(c));
        take(chunks_.back(), bytes, addr);
        used_ += bytes;
        return CAVE_OK;
    }

    void release(uint64_t addr, uint64_t bytes)
    {
        bytes = align_up(bytes);
        for (Chunk &c : chunks_) {
            const uint64_t base = (uint64_t)c.base;
            if (addr < base || addr >= base + c.size) continue;
            uint64_t off = addr - base, len = bytes;
            auto next = c.free.lower_bound(off);
            if (next != c.free.end() && off + len == next->first) {
                len += next->second;
                next = c.free.erase(next);
            }
            if (next != c.free.begin()) {
                auto prev = std::prev(next);
                if (prev->first + prev->second == off) {
                    off = prev->first;
                    len += prev->second;
                    c.free.erase(prev);
                }
            }
            c.free[off] = len;
            used_ -= bytes;
            return;
        }
    }

    // host view of a device address of a pinned-host arena (nullptr if none)
    const void *host_of(uint64_t addr) const
    {
        for (const Chunk &c : chunks_) {
            const uint64_t base = (uint64_t)c.base;
            if (addr >= base && addr < base + c.size) return (const char *)c.hbase + (addr - base);
        }
        return nullptr;
    }

private:
    struct Chunk {
        void *base = nullptr;   // device address
        void *hbase = nullptr;  // host address (pinned-host arena)
        uint64_t size = 0;
        std::map<uint64_t, uint64_t> free;  // offset -> length
    };
    static bool take(Chunk &c, uint64_t bytes, uint64_t &addr)
    {
        for (auto it = c.free.begin(); it != c.free.end(); ++it)
            if (it->second >= bytes) {
                const uint64_t off = it->first, len = it->second;
                c.free.erase(it);
                if (len > bytes) c.free[off + bytes] = len - bytes;
                addr = (uint64_t)c.base + off;
                return true;
            }
        return false;
    }
    std::vector<Chunk> chunks_;
    uint64_t budget_;
    bool host_;
    uint64_t chunk_bytes_;
    uint64_t used_ = 0;
};

struct Pending {
    uint64_t id;
    CaveSimSpec spec;
    DevJob job;           // parts, tier, tiling decided at submit
    uint64_t state_bytes;
    uint64_t final_bytes;
    double submit_s;
};

struct Inflight {
    uint64_t id;
    uint32_t parts;
    uint32_t X, Y, Z;
    uint64_t region = 0, state_bytes = 0, final_addr = 0, final_bytes = 0;
    double submit_s, admit_s;
};

}  // namespace
}  // namespace cave

using namespace cave;

struct CaveServer {
    CaveServeOptions opt{};
    hipStream_t kstream = nullptr;  // the resident kernel
    ServeCtl *ctl = nullptr;        // host-coherent
    DevJob *ring = nullptr;         // host-coherent
    DevResult *results = nullptr;   // host-coherent
    uint64_t *done = nullptr;       // host-coherent
    // the same four allocations as the kernel addresses them
    ServeCtl *d_ctl = nullptr;
    DevJob *d_ring = nullptr;
    DevResult *d_results = nullptr;
    uint64_t *d_done = nullptr;
    ServeDev *sd = nullptr;         // device
    JobSync *sync = nullptr;        // device
    uint32_t sync_next = 0;
    Arena *arena = nullptr;        // HBM: request state
    Arena *final_arena = nullptr;  // pinned host memory: final cells
    uint32_t lds_bytes = 0, blocks = 0, resident = 0;
    bool kernel_live = false;
    std::deque<Pending> pending;
    std::map<uint64_t, Inflight> inflight;  // by first ring index
    std::deque<CaveServeCompletion> ready;
    uint64_t next_seq = 0;
    uint64_t admitted = 0, finished = 0, peak = 0;
    std::chrono::steady_clock::time_point t0;

    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

namespace {

void free_server(CaveServer *s)
{
    if (!s) return;
    if (s->kstream) (void)hipStreamSynchronize(s->kstream);
    delete s->arena;
    delete s->final_arena;
    if (s->sync) (void)hipFree(s->sync);
    if (s->sd) (void)hipFree(s->sd);
    for (void *p : {(void *)s->ctl, (void *)s->ring, (void *)s->results, (void *)s->done})
        if (p) (void)hipHostFree(p);
    if (s->kstream) (void)hipStreamDestroy(s->kstream);
    delete s;
}

int host_coherent(void **p, size_t bytes)
{
    if (hipHostMalloc(p, bytes, hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess) {
        *p = nullptr;
        return fail(CAVE_E_NOMEM, "hipHostMalloc (coherent) failed");
    }
    std::memset(*p, 0, bytes);
    return CAVE_OK;
}

template <typename T>
T *device_view(T *host)
{   // the kernel's address of a mapped host allocation
    void *d = nullptr;
    if (hipHostGetDevicePointer(&d, host, 0) != hipSuccess) return nullptr;
    return static_cast<T *>(d);
}

// Layout of one request on the resident kernel: storage tier, items per
// thread and the number of blocks it is split over, from the tuner's rates.
void plan_request(CaveServer *s, Pending &p)
{
    DevJob &j = p.job;
    const uint64_t words = job_words(j.X, j.Y, j.Z);
    const bool in_lds = words * 4u <= s->lds_bytes;
    j.tier = in_lds ? kTierLds : kTierOwnRegion;
    const uint32_t bucket = edge_bucket(j.X, j.Y, j.Z);
    const double r1 = bucket_rate(bucket, tune_kind(j, kBatchShapeCube), kThreads, 1, nullptr, false, nullptr);
    const double r2 = bucket_rate(bucket, tune_kind(j, kBatchShapeCube), kThreads, 2, nullptr, false, nullptr);
    const uint32_t ipt = r2 > r1 * 1.03 ? 2u : 1u;
    const double dev_rate = std::max(r1, r2);
    const double block_rate = dev_rate > 0 ? dev_rate / s->resident : kFallbackBlockRate;
    const double est_us = (double)j.X * j.Y * j.Z * j.ticks / block_rate * 1e6;
    uint32_t parts = 1;
    if (!in_lds && est_us > s->opt.split_target_us) {
        parts = (uint32_t)std::ceil(est_us / s->opt.split_target_us);
        parts = std::min({parts, s->opt.max_parts, std::max(1u, j.Z / 4u), std::max(1u, s->resident / 4u)});
    }
    j.parts = parts;
    j.zchunks = part_zchunks(j.X, j.Y, j.Z, kThreads, ipt, parts);
    p.state_bytes = in_lds ? 0 : words * 4u;
    p.final_bytes = (p.spec.flags & CAVE_SIM_FINAL_HASH) ? (uint64_t)cave_cell_words(j.X, j.Y, j.Z) * 4u : 0;
}

int ensure_kernel(CaveServer *s)
{
    if (s->kernel_live) {
        const hipError_t q = hipStreamQuery(s->kstream);
        if (q == hipErrorNotReady) return CAVE_OK;
        if (q != hipSuccess) return fail(CAVE_E_DEVICE, std::string("resident kernel: ") + hipGetErrorString(q));
        s->kernel_live = false;
    }
    if (s->inflight.empty()) return CAVE_OK;
    // relaunch: claims continue from `head`; everything else starts clean
    ServeDev d{};
    CAVE_HIP(hipMemcpyAsync(&d, s->sd, sizeof d, hipMemcpyDeviceToHost, s->kstream));
    CAVE_HIP(hipStreamSynchronize(s->kstream));
    d.tail_seen = d.head;
    d.last_claim = 0;
    d.poll_lock = 0;
    d.closing = 0;
    CAVE_HIP(hipMemcpyAsync(s->sd, &d, sizeof d, hipMemcpyHostToDevice, s->kstream));
    CAVE_HIP(hipStreamSynchronize(s->kstream));
    BatchArgs a{};
    a.jobs = s->d_ring;
    a.results = s->d_results;
    a.lds_words = s->lds_bytes / 4u;
    a.sync = s->sync;
    a.wait_limit = wall_clock_khz() * 1000u * 4u;
    a.ctl = s->d_ctl;
    a.sd = s->sd;
    a.ring_cap = kRingCap;
    a.done = s->d_done;
    a.idle_limit = wall_clock_khz() * s->opt.idle_exit_us / 1000u;
    int rc = launch_serve(a, s->blocks, kThreads, s->lds_bytes, s->kstream);
    if (rc) return rc;
    s->kernel_live = true;
    return CAVE_OK;
}

// Admits pending requests in order while memory, the in-flight cap and the
// ring allow, and publishes their entries.
int admit(CaveServer *s)
{
    bool published = false;
    while (!s->pending.empty()) {
        Pending &p = s->pending.front();
        if (s->inflight.size() >= s->opt.max_inflight) break;
        const uint64_t oldest = s->inflight.empty() ? s->next_seq : s->inflight.begin()->first;
        if (s->next_seq + p.job.parts - oldest > kRingCap) break;
        if (p.job.parts > 1 && s->sync_next + 1 > kSyncPool) {
            // the rendezvous records are used once; recycle them when idle
            if (!s->inflight.empty() || s->kernel_live) break;
            CAVE_HIP(hipMemsetAsync(s->sync, 0, sizeof(JobSync) * kSyncPool, s->kstream));
            CAVE_HIP(hipStreamSynchronize(s->kstream));
            s->sync_next = 0;
        }
        Inflight f{};
        if (p.state_bytes) {
            const int rc = s->arena->alloc(p.state_bytes, f.region);
            if (rc == CAVE_E_BUDGET) break;
            if (rc) return fail(rc, "resident state allocation failed");
            f.state_bytes = p.state_bytes;
        }
        if (p.final_bytes) {
            const int rc = s->final_arena->alloc(p.final_bytes, f.final_addr);
            if (rc) {
                if (f.state_bytes) s->arena->release(f.region, f.state_bytes);
                if (rc == CAVE_E_BUDGET) break;
                return fail(rc, "final-cell staging allocation failed");
            }
            f.final_bytes = p.final_bytes;
        }
        const uint64_t seq0 = s->next_seq;
        const uint32_t slot = (uint32_t)(seq0 % kRingCap);
        DevJob j = p.job;
        j.result_slot = slot;
        j.seq = seq0 + 1u;
        j.region = f.region;
        j.final_addr = f.final_addr;
        j.has_final = p.final_bytes ? 1u : 0u;
        if (j.parts > 1) j.sync_slot = s->sync_next++;
        __atomic_store_n(&s->done[slot], 0ull, __ATOMIC_RELAXED);
        for (uint32_t k = 0; k < j.parts; k++) {
            DevJob e = j;
            e.part = k;
            s->ring[(seq0 + k) % kRingCap] = e;
        }
        f.id = p.id;
        f.parts = j.parts;
        f.X = j.X;
        f.Y = j.Y;
        f.Z = j.Z;
        f.submit_s = p.submit_s;
        f.admit_s = s->now();
        s->inflight.emplace(seq0, f);
        s->next_seq += j.parts;
        s->admitted++;
        s->peak = std::max(s->peak, s->arena->used());
        s->pending.pop_front();
        published = true;
    }
    if (published) __atomic_store_n(&s->ctl->tail, s->next_seq, __ATOMIC_RELEASE);
    return ensure_kernel(s);
}

// Collects finished requests into `ready`.
int collect(CaveServer *s)
{
    for (auto it = s->inflight.begin(); it != s->inflight.end();) {
        const uint64_t seq0 = it->first;
        const uint32_t slot = (uint32_t)(seq0 % kRingCap);
        if (__atomic_load_n(&s->done[slot], __ATOMIC_ACQUIRE) != seq0 + 1u) {
            ++it;
            continue;
        }
        const Inflight &f = it->second;
        const DevResult r = s->results[slot];
        CaveServeCompletion c{};
        c.request_id = f.id;
        c.result.info = r.info;
        c.result.classification = r.classification;
        c.result.ticks_run = r.ticks_run;
        c.result.last = r.last;
        c.submit_s = f.submit_s;
        c.admit_s = f.admit_s;
        c.run_ms = (double)(r.t_end - r.t_start) / (double)wall_clock_khz();
        c.blocks = f.parts;
        c.status = r.err ? CAVE_E_STATE : CAVE_OK;
        if (f.final_bytes && !r.err) {
            // written by the kernel before its release fence and the completion flag
            const uint32_t *cells = static_cast<const uint32_t *>(s->final_arena->host_of(f.final_addr));
            if (!cells) return fail(CAVE_E_STATE, "final cells outside the staging arena");
            c.result.final_hash = fnv_ref_layout(cells, f.X, f.Y, f.Z);
        }
        if (f.state_bytes) s->arena->release(f.region, f.state_bytes);
        if (f.final_bytes) s->final_arena->release(f.final_addr, f.final_bytes);
        c.finish_s = s->now();
        s->ready.push_back(c);
        s->finished++;
        it = s->inflight.erase(it);
    }
    return CAVE_OK;
}

int progress(CaveServer *s)
{
    int rc = collect(s);
    if (rc) return rc;
    return admit(s);
}

}  // namespace

extern "C" {

int cave_serve_create(const CaveServeOptions *opt, CaveServer **out)
{
    if (!out) return fail(CAVE_E_INVALID, "null out");
    *out = nullptr;
    CaveServer *s = new CaveServer();
    if (opt) s->opt = *opt;
    if (!s->opt.hbm_budget_bytes) s->opt.hbm_budget_bytes = kDefaultBudget;
    if (!s->opt.max_inflight) s->opt.max_inflight = kDefaultInflight;
    if (!s->opt.split_target_us) s->opt.split_target_us = kDefaultSplitUs;
    if (!s->opt.idle_exit_us) s->opt.idle_exit_us = kDefaultIdleUs;
    if (!s->opt.max_parts) s->opt.max_parts = kDefaultMaxParts;
    s->t0 = std::chrono::steady_clock::now();
    int rc = CAVE_OK;
    auto bail = [&](int code) {
        free_server(s);
        return code;
    };
    if (hipStreamCreateWithFlags(&s->kstream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(CAVE_E_DEVICE, "stream creation failed"));
    if ((rc = host_coherent((void **)&s->ctl, sizeof(ServeCtl))) ||
        (rc = host_coherent((void **)&s->ring, sizeof(DevJob) * kRingCap)) ||
        (rc = host_coherent((void **)&s->results, sizeof(DevResult) * kRingCap)) ||
        (rc = host_coherent((void **)&s->done, sizeof(uint64_t) * kRingCap)))
        return bail(rc);
    s->d_ctl = device_view(s->ctl);
    s->d_ring = device_view(s->ring);
    s->d_results = device_view(s->results);
    s->d_done = device_view(s->done);
    if (!s->d_ctl || !s->d_ring || !s->d_results || !s->d_done)
        return bail(fail(CAVE_E_DEVICE, "hipHostGetDevicePointer failed"));
    if (hipMalloc(&s->sd, sizeof(ServeDev)) != hipSuccess ||
        hipMalloc(&s->sync, sizeof(JobSync) * kSyncPool) != hipSuccess)
        return bail(fail(CAVE_E_NOMEM, "device allocation failed"));
    if (hipMemsetAsync(s->sd, 0, sizeof(ServeDev), s->kstream) != hipSuccess ||
        hipMemsetAsync(s->sync, 0, sizeof(JobSync) * kSyncPool, s->kstream) != hipSuccess ||
        hipStreamSynchronize(s->kstream) != hipSuccess)
        return bail(fail(CAVE_E_DEVICE, "device clear failed"));
    s->arena = new Arena(s->opt.hbm_budget_bytes, false, kChunkBytes);
    s->final_arena = new Arena(kFinalBudget, true, kFinalChunkBytes);
    s->lds_bytes = lds_cap(kThreads);
    (void)hipFuncSetAttribute(serve_kernel(), hipFuncAttributeMaxDynamicSharedMemorySize, (int)s->lds_bytes);
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, serve_kernel(), (int)kThreads, s->lds_bytes) !=
            hipSuccess ||
        per < 1)
        per = 1;
    s->resident = cu_count() * (uint32_t)per;
    s->blocks = s->resident;
    // rates of the common buckets for the split decision (F, FE, FEV; every edge bucket)
    static const uint32_t kEdges[6] = {28, 40, 56, 80, 112, 144};
    for (uint32_t fl : {1u, 3u, 7u})
        for (uint32_t b = 0; b < 6; b++)
            for (uint32_t ipt : {1u, 2u}) {
                const uint32_t dims[3] = {kEdges[b], kEdges[b], kEdges[b]};
                (void)bucket_rate(b, fl, kThreads, ipt, dims, true, s->kstream);
            }
    *out = s;
    return CAVE_OK;
}

int cave_serve_submit(CaveServer *s, const CaveSimSpec *spec, uint64_t request_id)
{
    if (!s || !spec) return fail(CAVE_E_INVALID, "null argument");
    int rc = validate_sim(*spec);
    if (rc) return rc;
    if (spec->flags & CAVE_SIM_SHAPE_ERD)  // the resident kernel is the cube shape's
        return fail(CAVE_E_UNSUPPORTED, "the server runs cube-shaped requests; batch ERD grids with cave_batch_run");
    Pending p{};
    p.id = request_id;
    p.spec = *spec;
    p.job = make_job(*spec, 0);
    p.submit_s = s->now();
    plan_request(s, p);
    if (align_up(p.state_bytes) + align_up(p.final_bytes) > s->opt.hbm_budget_bytes)
        return fail(CAVE_E_BUDGET, "request state exceeds the HBM admission budget");
    s->pending.push_back(p);
    return admit(s);
}

int cave_serve_poll(CaveServer *s, CaveServeCompletion *out, size_t cap, size_t *n)
{
    if (!s || !n || (cap && !out)) return fail(CAVE_E_INVALID, "null argument");
    *n = 0;
    int rc = progress(s);
    if (rc) return rc;
    while (*n < cap && !s->ready.empty()) {
        out[(*n)++] = s->ready.front();
        s->ready.pop_front();
    }
    return CAVE_OK;
}

int cave_serve_drain(CaveServer *s)
{
    if (!s) return fail(CAVE_E_INVALID, "null server");
    while (!s->pending.empty() || !s->inflight.empty()) {
        int rc = progress(s);
        if (rc) return rc;
        if (!s->inflight.empty()) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    return CAVE_OK;
}

int cave_serve_stats(const CaveServer *s, uint64_t *admitted, uint64_t *finished, uint64_t *resident_bytes,
                     uint64_t *peak_resident_bytes)
{
    if (!s) return fail(CAVE_E_INVALID, "null server");
    if (admitted) *admitted = s->admitted;
    if (finished) *finished = s->finished;
    if (resident_bytes) *resident_bytes = s->arena->used();
    if (peak_resident_bytes) *peak_resident_bytes = s->peak;
    return CAVE_OK;
}

int cave_serve_destroy(CaveServer *s)
{
    if (!s) return CAVE_OK;
    if (s->ctl) __atomic_store_n(&s->ctl->stop, 1ull, __ATOMIC_RELEASE);
    free_server(s);
    return CAVE_OK;
}

}  // extern "C"
