// Device-side state and helpers shared by engine.hip and pagerank.hip.
// gfx950 only: 64-lane wavefronts, 160 KiB LDS per CU.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <memory>
#include <vector>

#include "og_internal.h"

#define OG_HIP(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            ::og::raise(e_ == hipErrorOutOfMemory ? OG_ENOMEM_DEVICE : OG_EHIP, "%s: %s", \
                        #x, hipGetErrorString(e_));                                        \
    } while (0)

namespace og {

constexpr int kWave = 64;
constexpr int kBlock = 256;             // threads per block of the frontier kernels
// Frontier compaction windows: each thread of a count / write workgroup
// takes FV uint4 loads of flags (16 * FV vertices), so a workgroup covers
// kBlock * 16 * FV vertices.  Graphs up to kSmallFrontierV vertices use
// FV = 1 (more, shorter-lived workgroups: their rounds are latency bound);
// larger ones FV = 2 (half the blocks for the single-workgroup scan).
constexpr uint64_t kSmallFrontierV = 1ull << 22;
constexpr int frontier_fv(uint64_t V) { return V <= kSmallFrontierV ? 1 : 2; }
constexpr uint64_t verts_per_block(int fv) { return (uint64_t)kBlock * 16 * fv; }
constexpr int kDynEdgesPerThread = 8;   // dynamic relax: edges per lane
constexpr int kDynEdgesPerBlock = kBlock * kDynEdgesPerThread;
constexpr int kMaxSlots = 16;           // staging slots of the dynamic-region pipeline
// Zero-copy loading keeps part of the dynamic region in HBM in whole chunks of
// kResChunk consecutive vertices; the link moves pinned host memory in lines
// of kLinkLine bytes, each fetched whole.
constexpr uint32_t kResChunkShift = 5;
constexpr uint32_t kResChunk = 1u << kResChunkShift;
constexpr int64_t kNotResident = INT64_MIN;
constexpr uint32_t kLinkLine = 128;
constexpr uint64_t kSlotTargetBytes = 256ull << 20;  // preferred slot (chunk) size

enum Algo { ALGO_BFS = 0, ALGO_SSSP = 1, ALGO_CC = 2, ALGO_MSBFS = 3 };

constexpr int kMsWidth = 64;            // sources per multi-source batch (bits of a u64 mask)

// Per-vertex state the relax kernels touch.
struct RelaxArgs {
    uint32_t *val;        // BFS level / SSSP distance / CC label
    uint8_t *flag;        // next-frontier flags
    uint64_t *visit;      // multi-source: sources that reached v
    const uint64_t *front;// multi-source: sources whose frontier holds v
    uint64_t *next;       // multi-source: sources offered to v this round
    uint32_t wbits;       // packed weighted edges: bits of the weight field
    const uint64_t *pull; // multi-source, optional: sources relaxed bottom-up this round
                          // (left out of the top-down relax)
    const uint64_t *pmid; // light/heavy split, optional: end of each row's light edges
};

// Edge range of u in a round: the whole row, or with the light/heavy split
// its light part [ptr[u], pmid[u]) or, in a heavy round, [pmid[u], ptr[u+1]).
__device__ __forceinline__ void row_range(const uint64_t *__restrict__ ptr, const uint64_t *__restrict__ pmid,
                                          bool heavy, uint64_t u, uint64_t &b, uint64_t &e) {
    if (!pmid) {
        b = ptr[u];
        e = ptr[u + 1];
    } else if (heavy) {
        b = pmid[u];
        e = ptr[u + 1];
    } else {
        b = ptr[u];
        e = pmid[u];
    }
}

// Degree of u's row in the round's range.  With a degree table (4 bytes per
// vertex: deg the whole-row degrees, or with the split deg the light and
// hdeg the heavy degrees) one or two loads replace the two 8-byte row
// bounds; a round over whole rows of a split graph adds the two parts.
__device__ __forceinline__ uint64_t round_deg(const uint32_t *__restrict__ deg, const uint32_t *__restrict__ hdeg,
                                              const uint64_t *__restrict__ ptr, const uint64_t *__restrict__ pmid,
                                              bool heavy, uint64_t u) {
    if (deg) {
        if (pmid) return heavy ? hdeg[u] : deg[u];
        return (uint64_t)deg[u] + (hdeg ? hdeg[u] : 0u);
    }
    uint64_t b, e;
    row_range(ptr, pmid, heavy, u, b, e);
    return e - b;
}

// Degree-bin thresholds and shapes for the static relax (autotunable).
struct Tuning {
    uint32_t t1 = 32;      // bin 0: 1 <= deg <= t1, lpv0 lanes per vertex
    uint32_t lpv0 = 8;     // 1, 2, 4, 8, 16, 32 or 64
    uint32_t t2 = 4096;    // bin 1: t1 < deg <= t2, one wave per vertex
    uint32_t tile = 2048;  // bin 2: deg > t2, tiles of `tile` edges, one workgroup each
    uint32_t delta = 0;    // SSSP near-far bucket width (0: every improved vertex every round)
    uint32_t defer = 0;    // SSSP / CC: 1 = hold dynamic vertices back while static ones are active;
                           // SSSP with delta > 0: 2 = dynamic vertices relax only after the
                           // static part of their near-far bucket has settled
    uint32_t dyn_grid = 0; // workgroups of the dynamic-region relax (0: kDynGridCap)
};

// Frontier summary produced on the device each iteration and read by the host.
struct FrontierInfo {
    uint32_t n[4];          // bin0 vertices, bin1 vertices, bin2 tiles, dynamic vertices
    uint32_t nzero;         // active vertices with no out-edge
    uint32_t nhub;          // bin2 vertices
    uint64_t static_deg;    // sum of degrees over static active vertices
    uint64_t dyn_deg;       // sum of degrees over dynamic active vertices
    uint32_t nfar;          // active vertices held back (value >= threshold)
    uint32_t min_far;       // smallest value among them
    uint32_t live;          // 1 while active or held-back vertices remain
    uint32_t dyn_held;      // 1: the dynamic list was held back this round (n[3] = 0)
    uint32_t pull;          // 1: BFS round runs bottom-up (the lists are not built)
    uint32_t heavy;         // 1: light/heavy split, this round relaxes heavy edges
};

// Device-resident driver state of one traversal.  The scan of each
// compaction advances it, so rounds can be enqueued without a host round
// trip: the near-far threshold moves on the device and the statistics the
// host used to add up per round accumulate here.
struct RoundCtl {
    uint32_t thresh;             // near-far threshold of the next compaction (0xFFFFFFFF: off)
    uint32_t delta;              // near-far bucket width (0: off)
    uint32_t defer;              // 1: hold the dynamic list back while static vertices are active;
                                 // 2: dynamic vertices are near only below fbound
    uint32_t fbound;             // defer 2: dynamic vertices with value < fbound are near
    uint32_t split;              // light/heavy split in use (rows: light edges, then heavy)
    uint32_t mode;               // split: 1 = the next compaction collects the heavy-pending vertices
    uint64_t hpend;              // split: vertices light-relaxed since the last heavy round
    uint32_t rounds;             // compactions that found active vertices
    uint32_t compactions;        // all compactions
    uint32_t live;               // last compaction's FrontierInfo::live
    uint32_t overflow;           // multi-source: a level past 255 was needed
    uint32_t dir_opt;            // BFS: choose top-down / bottom-up per round
    uint32_t pull;               // direction of the current round (1: bottom-up)
    uint32_t pull_rounds;        // rounds run bottom-up
    uint64_t unexplored;         // BFS: edges of vertices not yet in any frontier
    uint64_t nverts;             // V (the bottom-up -> top-down switch compares the frontier with V / 24)
    uint64_t node_sum;           // active vertices, summed over rounds
    uint64_t static_node_sum;    // ... of them in the static bins
    uint64_t dyn_node_sum;       // ... of them in the dynamic list
    uint64_t work_edges;         // degree sums over active vertices
    uint64_t static_work_edges;  // ... over the static bins
    uint64_t dyn_rounds;         // rounds with a dynamic list
    uint32_t halt;               // 1 (device-driven runs): once a compaction finds nothing live, the
                                 // count passes of rounds enqueued past the end skip their sweep
    uint32_t pad_;
};

constexpr int kBcntRows = 5;            // per-block rows: bins 0-3, any flag
constexpr uint32_t kSbBlocks = 64;      // compaction blocks per super-block (one lane each in the write pass)
constexpr int kSbVals = 9;              // super-block totals: bins 0|1, bins 2|3, dynamic degrees, zero, hub,
                                        // far, static degrees, lagging, min far
constexpr uint8_t kCode0 = 2;           // count pass: flag byte = kCode0 + bin of the vertex

// Per-block and per-super-block counts of the compaction passes.
struct BlockCounts {
    uint32_t *bcnt;
    uint64_t *bdeg;
    uint64_t *sbt;
    uint64_t *sbp;
    uint32_t nblk, nsb;
};

// Buffers of one frontier traversal: next-frontier flags, the compacted bin
// lists, the dynamic degree prefix and the per-block scan scratch.  A graph's
// single-source runs use DeviceState::fr; each serving engine owns its own so
// that engines can iterate concurrently on different streams.
struct Frontier {
    uint8_t *flag = nullptr;         // Vpad next-frontier flags
    uint8_t *hflag = nullptr;        // Vpad light/heavy split: heavy-pending flags
    uint32_t *list = nullptr;        // V: bin0 | bin1 | dynamic lists
    uint64_t *hub = nullptr;         // hub tiles (tile << 32 | vertex)
    uint64_t hub_cap = 0;
    uint64_t *doff = nullptr;        // V + 1 degree prefix of the dynamic list
    uint32_t *dchunk = nullptr;      // list entry holding edge c * kDynEdgesPerBlock of the dynamic edge space
    uint64_t dchunk_cap = 0;
    uint32_t *bcnt = nullptr;        // kBcntRows * nblk per-block counts
    uint64_t *bdeg = nullptr;        // nblk per-block dynamic degree sums
    uint64_t *sbt = nullptr;         // kSbVals * nsb super-block totals (added by the count pass; zero, the
                                     // min row all ones, between compactions)
    uint64_t *sbp = nullptr;         // 3 * nsb super-block exclusive prefixes (bins 0|1, 2|3, dynamic degrees)
    uint32_t nblk = 0;
    uint32_t nsb = 0;
    BlockCounts counts() const { return BlockCounts{bcnt, bdeg, sbt, sbp, nblk, nsb}; }
    uint32_t fv = 2;                 // flag vectors per thread of the compaction (frontier_fv)
    uint64_t V = 0;
    FrontierInfo *info = nullptr;
    FrontierInfo *h_info = nullptr;  // pinned
    RoundCtl *ctl = nullptr;
    RoundCtl *h_ctl = nullptr;       // pinned
    unsigned long long *counters = nullptr;    // [0] updates static, [1] updates dynamic, [2..] scratch
    unsigned long long *h_counters = nullptr;  // pinned
};

// Timing of one class of launches: pairs of events on the launching stream.
struct TimedClass {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> marks;
    uint64_t launches = 0;
    double ms = 0;  // already resolved (marks of captured graphs, read after each replay)
};

class EventPool {
  public:
    ~EventPool() {
        for (auto e : all_) (void)hipEventDestroy(e);
    }
    hipEvent_t get() {
        if (next_ == all_.size()) {
            hipEvent_t e;
            OG_HIP(hipEventCreate(&e));
            all_.push_back(e);
        }
        return all_[next_++];
    }
    void reset() { next_ = 0; }

  private:
    std::vector<hipEvent_t> all_;
    size_t next_ = 0;
};

// Owner of device and pinned host allocations; frees them all at once.
struct MemPool {
    std::vector<void *> dev, pin;
    uint64_t device_bytes = 0;
    MemPool() = default;
    MemPool(const MemPool &) = delete;
    MemPool &operator=(const MemPool &) = delete;
    ~MemPool() { release(); }
    void *dalloc(size_t bytes) {
        void *p = nullptr;
        if (bytes == 0) bytes = 256;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            raise(OG_ENOMEM_DEVICE, "hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
        }
        dev.push_back(p);
        device_bytes += bytes;
        return p;
    }
    void *palloc(size_t bytes) {
        void *p = nullptr;
        if (hipHostMalloc(&p, bytes ? bytes : 8, hipHostMallocDefault) != hipSuccess || !p) {
            (void)hipGetLastError();
            raise(OG_ENOMEM_HOST, "pinned host allocation of %zu bytes failed", bytes);
        }
        pin.push_back(p);
        return p;
    }
    void release() {
        for (void *p : dev) (void)hipFree(p);
        for (void *p : pin) (void)hipHostFree(p);
        dev.clear();
        pin.clear();
        device_bytes = 0;
    }
};

struct DeviceState {
    og_approach approach = OG_ASCETIC;
    int device = 0;
    uint64_t V = 0, E = 0, Vpad = 0;
    uint32_t eb = 4;                 // bytes per edge on the device (static region, staging) and the link
    og_plan_info plan{};
    Tuning tune;
    unsigned host_threads = 0;
    uint32_t flags = 0;

    // graph on the device
    uint64_t *d_ptr = nullptr;       // V + 1 (device copy; USM: managed)
    uint64_t *d_pmid = nullptr;      // light/heavy split: V, end of the light edges of each row
    uint32_t *d_deg = nullptr;       // degree table of the compaction: Vpad whole-row degrees, or with
                                     // the split Vpad light then Vpad heavy degrees (null: row bounds)
    bool deg_split = false;          // d_deg holds light and heavy degrees
    const uint32_t *deg_heavy() const { return d_deg && deg_split ? d_deg + Vpad : nullptr; }
    std::vector<uint64_t> h_pmid;    // host copy (the staged gather)
    uint32_t split_weight = 0;       // light: weight < split_weight (0: no split)
    const uint64_t *run_pmid = nullptr;  // d_pmid while a run uses the split (SSSP near-far), else null
    void *d_static = nullptr;        // static-region edges (USM: managed, whole graph)
    uint64_t static_used = 0;        // edges of the static vertices, = ptr[vsplit]
    void *d_staging = nullptr;       // nslots * slot_edges edges (zero-copy: the resident chunks)
    uint64_t slot_edges = 0;
    bool managed = false;
    // Weighted edges on the device and the link: one 4-byte word
    // (to << wbits | w) when packed, else the CSR's 8-byte {to, w}.
    bool packed = false;
    uint32_t wbits = 0;
    uint32_t deb = 4;                // bytes per dynamic-region edge on the link (= eb)
    uint64_t dyn_base = 0;           // first edge of the dynamic region (= ptr[vsplit])
    const void *h_dyn = nullptr;     // host source of dynamic edges (packed copy or the CSR)
    const void *hd_dyn = nullptr;    // the same, as a device address (zero-copy mode)
    uint64_t slot_dyn = 0;           // dynamic edges per staging slot in the transfer format
    // zero-copy with the light/heavy split: the host copy of the dynamic
    // region holds every row's light part in vertex order, then every heavy
    // part (light part of v at lpos[v - vsplit], heavy part at
    // split_nl + (ptr[v] - dyn_base) - lpos[v - vsplit]); lpos has
    // V - vsplit + 1 entries
    uint64_t *d_lpos = nullptr;
    uint64_t split_nl = 0;           // light edges of the dynamic region
    uint64_t dyn_resident = 0;       // zero-copy: dynamic edges held in d_staging (whole chunks)
    int64_t *d_rdelta = nullptr;     // zero-copy: per chunk of kResChunk vertices (index v >> kResChunkShift),
                                     // d_staging index - host-array position, or kNotResident; two per
                                     // chunk (light block, heavy block) with d_lpos
    uint64_t *m_ptr = nullptr;       // USM managed copies
    void *m_edges = nullptr;

    // per-vertex state
    uint32_t *d_val = nullptr;       // V
    Frontier fr;                     // frontier buffers of single-source / batch runs

    // pinned host mirrors of the dynamic list
    uint32_t *h_dlist = nullptr;     // V
    uint64_t *h_doff = nullptr;      // V + 1
    std::vector<void *> h_stage;     // pinned gather buffer per staging slot

    hipStream_t s_compute = nullptr, s_copy = nullptr;
    hipStream_t s_static = nullptr;  // static-region relax, overlapping the dynamic region
    hipEvent_t ev_fork = nullptr;    // compaction of the round finished (static relax may start)
    hipEvent_t ev_static = nullptr;  // static relax of the round finished
    uint32_t nslots = 0;             // staging slots (a pipeline of small chunks)
    std::vector<hipEvent_t> ev_h2d;    // copy into slot finished
    std::vector<hipEvent_t> ev_relax;  // relax reading slot finished
    std::vector<char> slot_h2d_pending, slot_relax_pending;
    EventPool events;
    std::unique_ptr<WorkerPool> pool;

    // multi-source BFS state (allocated on first use)
    uint64_t *d_visit = nullptr, *d_front = nullptr, *d_next = nullptr;  // V each
    uint8_t *d_mlvl = nullptr;       // V x 64 levels, vertex-major (0 = unreached)
    uint32_t *d_msrc = nullptr;      // 64 sources of the batch
    unsigned long long *d_msum = nullptr;  // 2 x 64: checksum, reached per source
    uint32_t *d_extract = nullptr;   // V scratch for one source's u32 levels
    // direction-optimising BFS: in-edge CSR (OG_PLAN_DIRECTION_OPT)
    uint64_t *d_in_ptr = nullptr;    // V + 1
    uint32_t *d_in_cols = nullptr;   // E in-neighbours

    // PageRank state (allocated on first use)
    void *d_rank[2] = {nullptr, nullptr};
    void *d_contrib = nullptr;
    uint32_t *d_outdeg = nullptr;
    double *d_partial = nullptr;     // per-block residual / dangling partials
    double *d_scalars = nullptr;     // [0] dangling, [1] l1
    double *h_scalars = nullptr;
    uint64_t pr_partial_cap = 0;     // entries of d_partial
    bool pr_fp64 = false;
    std::shared_ptr<void> pr_plan;
    std::shared_ptr<void> round_graph;  // captured batches of device-driven rounds (engine.hip)
    void *h_out[2] = {nullptr, nullptr};  // pinned staging chunks of result copy-backs (copy_to_host)
    hipEvent_t ev_out[2] = {nullptr, nullptr};
    uint32_t last_pairs[8] = {};        // relax + compaction pairs of the last device-driven run, per algorithm   // PageRank bins and chunk ranges (pagerank.hip)

    MemPool mem;                     // device + pinned allocations of this preparation

    void *dalloc(size_t bytes) { return mem.dalloc(bytes); }
    void *palloc(size_t bytes) { return mem.palloc(bytes); }
    uint64_t device_bytes() const { return mem.device_bytes; }
};

// ------------------------------------------------------------ wave helpers

__device__ __forceinline__ uint64_t shfl_up_u64(uint64_t x, int d) {
    uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    lo = __shfl_up(lo, d, kWave);
    hi = __shfl_up(hi, d, kWave);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t x, int m) {
    uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    lo = __shfl_xor(lo, m, kWave);
    hi = __shfl_xor(hi, m, kWave);
    return ((uint64_t)hi << 32) | lo;
}

// Exclusive scan over the block of x (u64); *total receives the block sum.
// `lds` must hold blockDim.x / 64 entries.  Contains __syncthreads.
__device__ __forceinline__ uint64_t block_exclusive_scan(uint64_t x, uint64_t *total, uint64_t *lds) {
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    uint64_t inc = x;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        uint64_t y = shfl_up_u64(inc, o);
        if (lane >= o) inc += y;
    }
    if (lane == kWave - 1) lds[wid] = inc;
    __syncthreads();
    uint64_t before = 0, tot = 0;
    for (int w = 0; w < nw; ++w) {
        uint64_t s = lds[w];
        before += (w < wid) ? s : 0;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return before + inc - x;
}

// N independent exclusive scans over the block in one pass (one LDS round):
// ex[i] = sum of x[i] over lower threads, tot[i] = block sum.  `lds` must
// hold N * blockDim.x / 64 entries.  Contains __syncthreads.
template <int N>
__device__ __forceinline__ void block_scan_n(const uint64_t (&x)[N], uint64_t (&ex)[N], uint64_t (&tot)[N],
                                             uint64_t *lds) {
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    uint64_t inc[N];
#pragma unroll
    for (int i = 0; i < N; ++i) inc[i] = x[i];
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const uint64_t y = shfl_up_u64(inc[i], o);
            if (lane >= o) inc[i] += y;
        }
    }
    if (lane == kWave - 1)
#pragma unroll
        for (int i = 0; i < N; ++i) lds[i * nw + wid] = inc[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < N; ++i) {
        uint64_t before = 0, t = 0;
        for (int w = 0; w < nw; ++w) {
            const uint64_t s = lds[i * nw + w];
            before += (w < wid) ? s : 0;
            t += s;
        }
        ex[i] = before + inc[i] - x[i];
        tot[i] = t;
    }
    __syncthreads();
}

// Sum over the block; result valid in every thread.  Contains __syncthreads.
__device__ __forceinline__ uint64_t block_sum(uint64_t x, uint64_t *lds) {
    uint64_t t;
    (void)block_exclusive_scan(x, &t, lds);
    return t;
}

// Adds the wave's per-lane counts to *ctr with one atomic per wave.
__device__ __forceinline__ void wave_add(unsigned long long *ctr, uint32_t v) {
    uint64_t s = v;
#pragma unroll
    for (int m = 1; m < kWave; m <<= 1) s += shfl_xor_u64(s, m);
    if ((threadIdx.x & (kWave - 1)) == 0 && s) atomicAdd(ctr, (unsigned long long)s);
}

constexpr int kRelaxBatch = 4;  // edges per lane whose reads are issued together

// Edge i of an edge array: 4-byte column (unweighted), the packed 4-byte
// weighted word (to << wbits | w), or the CSR's 8-byte AoS {to, w}.
template <bool W, bool PK>
__device__ __forceinline__ void load_edge(const void *__restrict__ edges, uint64_t i, uint32_t wbits, uint32_t &to,
                                          uint32_t &wt) {
    if (W && PK) {
        const uint32_t x = ((const uint32_t *)edges)[i];
        to = x >> wbits;
        wt = x & ((1u << wbits) - 1u);
    } else if (W) {
        const uint2 e = ((const uint2 *)edges)[i];
        to = e.x;
        wt = e.y;
    } else {
        to = ((const uint32_t *)edges)[i];
        wt = 1;
    }
}

// ------------------------------------------------------------- host timing

struct Stopwatch {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

// Records a start/stop event pair around a launch sequence on `s`.
struct EventSpan {
    DeviceState *d;
    TimedClass *cls;
    hipStream_t s;
    hipEvent_t a;
    EventSpan(DeviceState *d_, TimedClass *c, hipStream_t s_) : d(d_), cls(c), s(s_) {
        a = d->events.get();
        OG_HIP(hipEventRecord(a, s));
    }
    ~EventSpan() noexcept(false) {
        hipEvent_t b = d->events.get();
        OG_HIP(hipEventRecord(b, s));
        cls->marks.push_back({a, b});
        cls->launches++;
    }
};

inline double sum_ms(const TimedClass &c) {
    double t = c.ms;
    for (auto &m : c.marks) {
        float ms = 0;
        OG_HIP(hipEventElapsedTime(&ms, m.first, m.second));
        t += ms;
    }
    return t;
}

// Copies n bytes of device memory at src to host memory at dst (pageable or
// not) on stream s and returns when they are there.  Large copies go through
// two pinned staging chunks: chunk k + 1 crosses the link while the host
// threads copy chunk k out, instead of HIP's staged copy into pageable
// memory (~10 GB/s on the box for a 64 MB result against ~50 for pinned).
void copy_to_host(DeviceState *d, void *dst, const void *src, size_t n, hipStream_t s);

// Allocates a frontier (device buffers + pinned summaries) sized for the
// graph and the current tuning; freed with the DeviceState.
void frontier_alloc(DeviceState *d, const og_graph *g, Frontier &f, MemPool &m);

// A single-source lane over a prepared, wholly device-resident graph: it
// shares the graph's device arrays (row pointers, edges, degree table, light
// / heavy split, in-edges) and owns values, frontier buffers, streams and
// captured rounds, so several lanes run single-source traversals side by
// side on their own streams.  `delta` (0: the graph's tuning) is the lane's
// near-far width.  Freed with device_state_free, which releases only what
// the lane allocated.
DeviceState *lane_state_create(og_graph *g, uint32_t delta);
// Device bytes lane_state_create allocates for graph g (values, flags, bin
// lists, degree prefix, block counts, hub tiles).
uint64_t lane_state_bytes(const og_graph *g);
// og_sssp from src on a lane; out (V values) may be null.
void lane_run_sssp(og_graph *g, DeviceState *lane, uint32_t src, uint32_t *out, og_run_stats *st);

// Resets f.ctl for a new traversal on stream s (a kernel, so it may sit in a
// captured graph): statistics zeroed, near-far width `delta` with the first
// threshold 1 + delta (the source's value is 1), or none for delta = 0;
// `defer` 1 holds the dynamic list back in rounds with static work, 2 lets
// dynamic vertices relax only once the static part of their bucket settled;
// dir_opt_edges > 0 turns on the BFS direction choice over that many edges.
// split: light/heavy split rounds (the graph's pmid rows; SSSP near-far).
void frontier_begin(Frontier &f, hipStream_t s, uint32_t delta, uint32_t defer = 0, uint64_t dir_opt_edges = 0,
                    bool split = false);

// One compaction: flags -> bin lists / hub tiles / sorted dynamic list with
// its degree prefix, flags cleared; f.info and f.ctl updated on the device.
// With `val` non-null, vertices whose value is >= f.ctl->thresh stay flagged
// and are only counted (near-far order).  With to_host, the summary is also
// copied to f.h_info (stream-ordered, not synchronised).
// pmid (light/heavy split, with frontier_begin's split): rows are ranged by
// the round's mode.
void frontier_compact(DeviceState *d, Frontier &f, hipStream_t s, const uint32_t *val = nullptr,
                      bool to_host = true, const uint64_t *pmid = nullptr);


// Launches the static-region relax of the bins for op ALGO.  The kernels
// read their counts from f.info; `hfi` (the host copy of it, or null when the
// host does not know it) only sizes the grids and skips empty bins.
template <int ALGO, bool W>
void launch_static_relax(DeviceState *d, Frontier &f, const FrontierInfo *hfi, const RelaxArgs &a, hipStream_t s);
// Served BFS: the bottom-up scan's arrays (graphs with their in-edges).
struct SrvPull {
    const uint64_t *in_ptr;
    const uint32_t *in_cols;
    uint64_t V;
    const uint64_t *visit, *front;
    uint64_t *next;
    uint8_t *flag;
    const unsigned long long *pull;  // bit j: slot j scans bottom-up this round
};
// One launch relaxing a served round: top-down slots through the static
// bins and, with `pull`, the bottom-up scan beside them.
void launch_srv_round(DeviceState *d, Frontier &f, const RelaxArgs &a, hipStream_t s, const SrvPull *pull,
                      bool weighted);

}  // namespace og
