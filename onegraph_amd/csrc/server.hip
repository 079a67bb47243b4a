// Query server: continuous batching of BFS queries and concurrent SSSP
// lanes on device-resident graphs (no reference counterpart; the reference
// runs one query per process, BFS/oneGraph_BFS/main.cpp:3-14).
//
// BFS: every graph has one batch engine with 64 query slots, its own HIP
// stream, frontier buffers and host thread.  A slot is one bit of the
// per-vertex u64 masks visit / front / next.  The host waits once per round,
// for the commit's summary (step 2), and enqueues the next round while the
// rest of this one runs:
//   1. admit pending queries into free slots: the source's bit goes into
//      `next`, so the commit below seeds it like any other offer;
//   2. commit (one sweep over the frontier flags): the offered, unseen slots
//      of a vertex become its frontier mask and each of them adds
//      mix64(v << 32 | level) to its slot's digest and the vertex and its
//      degree to the slot's reached / traversed counts, so no level array is
//      kept; then every slot picks its direction for the round;
//   3. compact the flags into degree bins (shared with og_bfs);
//   4. relax, one launch: top-down slots through the multi-source bitmask
//      bins (one 64-bit atomic per edge serves every slot) and, beside them,
//      bottom-up slots in one pass over the in-edges of the vertices they
//      have not reached (graphs prepared with OG_PLAN_DIRECTION_OPT);
//   5. clear the masks of the relaxed vertices;
//   6. slots whose frontier emptied report their results; their visit bits
//      are cleared later, by one sweep for many slots.
// Queries that arrive while a round runs join at the next round boundary, so
// their work overlaps with the queries already in flight.
//
// SSSP: every weighted graph has K lanes, each a complete single-source
// near-far engine (values, frontier, streams, captured rounds) over the
// shared device-resident graph, taking queries from the graph's SSSP queue
// one at a time.  Distances do not share across queries the way BFS masks
// do (a 64-slot distance matrix is 256 bytes per vertex and defeats the
// caches), so SSSP queries overlap as concurrent lanes instead.
#include <algorithm>
#include <condition_variable>
#include <deque>
#include <unordered_map>

#include "device.h"

struct og_server;

namespace og {

struct SeedList {
    uint32_t n;
    uint32_t slot[64];
    uint32_t src[64];
};

struct SlotLevels {  // BFS: the level a vertex committed this round takes, per slot
    uint32_t lv[64];
};

struct KeepList {
    uint32_t *p[64];
};

__device__ __forceinline__ uint64_t srv_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t wave_or(uint64_t x) {
#pragma unroll
    for (int m = 1; m < kWave; m <<= 1) x |= shfl_xor_u64(x, m);
    return x;
}

// Seeds: the slot's bit goes into next[src]; the commit admits it.  The
// slot's explored-edge count restarts.
__global__ void k_srv_seed(SeedList sl, uint64_t *next, uint8_t *flag, unsigned long long *explored) {
    const uint32_t i = threadIdx.x;
    if (i < sl.n) {
        atomicOr((unsigned long long *)(next + sl.src[i]), 1ull << sl.slot[i]);
        flag[sl.src[i]] = 1;
        explored[sl.slot[i]] = 0;
    }
}

// Commit sweep over the flags, consecutive lanes on consecutive vertices (so
// a wave's visit / next / front / ptr traffic is whole cache lines, and a
// wave whose 64 flags are clear moves on after one coalesced byte load):
// nw = next & ~visit becomes the vertex's frontier mask (when it has
// out-edges); each slot in nw adds the vertex's level to its digest in
// acc[0..63] (and writes it to a kept result array) and counts the vertex
// and its out-degree in fstat (the round's newly reached vertices and their
// edges per slot).  Flags end up 1 exactly for vertices with a non-empty
// mask; *alive gets the OR of all masks.  The per-slot sums are kept in
// kCommitCopies copies in LDS, lane l adding into copy l % kCommitCopies:
// lanes of a wave that reach the same slot would otherwise all add into one
// address, which the LDS serialises; a stride of 65 words puts the copies of
// a slot on distinct banks.  The slots' levels and kept-result pointers are
// staged in LDS first: indexed per lane, the by-value kernel arguments would
// be read from memory once per (vertex, slot).
constexpr int kCommitCopies = 16;
constexpr int kCommitStride = 65;
__global__ __launch_bounds__(kBlock) void k_srv_commit(uint8_t *__restrict__ flag, uint64_t V,
                                                       const uint64_t *__restrict__ ptr, uint64_t *__restrict__ visit,
                                                       uint64_t *__restrict__ front, uint64_t *__restrict__ next,
                                                       SlotLevels L, KeepList keep, unsigned long long *acc,
                                                       unsigned long long *fstat, unsigned long long *alive) {
    __shared__ unsigned long long sdig[kCommitCopies * kCommitStride], sfc[kCommitCopies * kCommitStride],
        sfe[kCommitCopies * kCommitStride];
    __shared__ uint32_t slv[64];
    __shared__ uint32_t *skeep[64];
    __shared__ unsigned long long skm;
    for (int i = threadIdx.x; i < kCommitCopies * kCommitStride; i += kBlock) sdig[i] = sfc[i] = sfe[i] = 0;
    if (threadIdx.x < 64) {
        slv[threadIdx.x] = L.lv[threadIdx.x];
        skeep[threadIdx.x] = keep.p[threadIdx.x];
        const unsigned long long km = __ballot(keep.p[threadIdx.x] != nullptr);
        if (threadIdx.x == 0) skm = km;
    }
    __syncthreads();
    const int cp = (int)(threadIdx.x % kCommitCopies) * kCommitStride;
    const uint64_t kmask = skm;
    uint64_t al = 0;
    for (uint64_t b = (uint64_t)blockIdx.x * kBlock; b < V; b += (uint64_t)gridDim.x * kBlock) {
        const uint64_t v = b + threadIdx.x;
        const bool on = v < V && flag[v];
        if (!__ballot(on)) continue;  // b is the same for the whole wave
        if (on) {
            const uint64_t vis = visit[v];
            const uint64_t nw = next[v] & ~vis;
            next[v] = 0;
            const uint64_t deg = ptr[v + 1] - ptr[v];
            if (nw) {
                // each slot that reaches v: digest term, one more reached
                // vertex, deg more traversed edges (= its frontier edges)
                visit[v] = vis | nw;
                uint64_t m = nw;
                while (m) {
                    const int j = __builtin_ctzll(m);
                    m &= m - 1;
                    const uint32_t lv = slv[j];
                    atomicAdd(&sdig[cp + j], (unsigned long long)srv_mix64((v << 32) | lv));
                    atomicAdd(&sfc[cp + j], 1ull);
                    if (deg) atomicAdd(&sfe[cp + j], (unsigned long long)deg);
                    if ((kmask >> j) & 1) skeep[j][v] = lv;
                }
            }
            const uint64_t fr = deg ? nw : 0;
            if (fr) {
                front[v] = fr;
                al |= fr;
            } else {
                flag[v] = 0;
            }
        }
    }
    al = wave_or(al);
    if ((threadIdx.x & (kWave - 1)) == 0 && al) atomicOr(alive, (unsigned long long)al);
    __syncthreads();
    if (threadIdx.x < 64) {
        unsigned long long dg = 0, fc = 0, fe = 0;
        for (int c = 0; c < kCommitCopies; ++c) {
            dg += sdig[c * kCommitStride + threadIdx.x];
            fc += sfc[c * kCommitStride + threadIdx.x];
            fe += sfe[c * kCommitStride + threadIdx.x];
        }
        if (dg) atomicAdd(acc + threadIdx.x, dg);
        if (fc) {
            atomicAdd(fstat + threadIdx.x, fc);
            if (fe) atomicAdd(fstat + 64 + threadIdx.x, fe);
        }
    }
}

// End of the commit: the round's newly reached vertices and edges of each
// slot join its reached / traversed totals (acc[64 + j], acc[128 + j]), and,
// with dir_opt, every slot picks its direction for the round (Beamer et al.,
// SC'12, per slot): top-down -> bottom-up once its frontier edges exceed 1/14
// of its unexplored edges, back when its frontier holds fewer than V / 24
// vertices.  Bit j of *pull = slot j relaxes bottom-up.  Resets the stats.
__global__ void k_srv_direction(uint64_t used, uint64_t V, uint64_t E, int dir_opt, unsigned long long *fstat,
                                unsigned long long *explored, unsigned long long *acc, unsigned long long *pull) {
    const uint32_t j = threadIdx.x;
    const uint64_t fc = fstat[j], fe = fstat[64 + j];
    acc[64 + j] += fc;
    acc[128 + j] += fe;
    const uint64_t ex = explored[j];
    const uint64_t unexplored = E > ex + fe ? E - ex - fe : 0;
    const bool was = (*pull >> j) & 1;
    bool now = false;
    if (dir_opt && ((used >> j) & 1)) now = was ? fc >= V / 24 : fe > unexplored / 14;
    explored[j] = ex + fe;
    fstat[j] = fstat[64 + j] = 0;
    const uint64_t b = __ballot(now);
    if (j == 0) *pull = b;
}

// After the relax: the masks of the round's listed vertices (bins 0/1 and
// hub tiles) are cleared; the next commit sets the new ones.
__global__ __launch_bounds__(kBlock) void k_srv_clear_front(const uint32_t *__restrict__ list,
                                                            const uint64_t *__restrict__ hub, uint64_t hub_cap,
                                                            const FrontierInfo *__restrict__ info, uint64_t *front) {
    const uint64_t n01 = (uint64_t)info->n[0] + info->n[1];
    const uint64_t nh = min((uint64_t)info->n[2], hub_cap);
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n01 + nh; i += (uint64_t)gridDim.x * kBlock)
        front[i < n01 ? list[i] : (uint32_t)hub[i - n01]] = 0;
}

// Clears the visit bits of finished slots (before the slots are reused).
__global__ __launch_bounds__(kBlock) void k_srv_clear_visit(uint64_t V, uint64_t dirty, uint64_t *visit) {
    for (uint64_t v = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v < V; v += (uint64_t)gridDim.x * kBlock) {
        const uint64_t x = visit[v];
        if (x & dirty) visit[v] = x & ~dirty;
    }
}

__global__ void k_srv_zero_acc(uint64_t done, unsigned long long *acc) {
    const uint32_t j = threadIdx.x;
    if (j < 64 && ((done >> j) & 1)) acc[j] = acc[64 + j] = acc[128 + j] = 0;
}

// ------------------------------------------------------------------ host

struct Query {
    uint64_t id;
    uint32_t graph, kind, source, flags;
    double t_submit;
};

struct SlotState {
    bool used = false;
    Query q{};
    uint32_t start_round = 0;
    double t_admit = 0;
};

constexpr uint32_t kSlots = 64;     // BFS query slots per graph (bits of a u64 mask)
constexpr uint32_t kMaxLanes = 4;   // SSSP lanes per weighted graph

// Device bytes of one SSSP lane: values 4 V, flags V, list 4 V, degree
// prefix 8 V (plus per-block scratch and the hub list).
static uint64_t lane_bytes(uint64_t V) { return 17 * V; }

struct BfsEngine {
    og_server *srv = nullptr;
    og_graph *g = nullptr;
    DeviceState *d = nullptr;
    uint32_t gi = 0;
    std::deque<Query> *pending = nullptr;
    MemPool mem;
    Frontier f;
    hipStream_t s = nullptr;
    hipEvent_t ev_summary = nullptr;  // the round's alive mask and slot sums are on the host
    uint64_t *visit = nullptr, *front = nullptr, *next = nullptr;
    unsigned long long *d_alive = nullptr, *h_alive = nullptr, *d_acc = nullptr, *h_acc = nullptr;
    // direction-optimising rounds (graphs prepared with their in-edges)
    unsigned long long *d_fstat = nullptr;     // [2][64] frontier vertices / edges per slot
    unsigned long long *d_explored = nullptr;  // [64] edges of the slot's past frontiers
    unsigned long long *d_pull = nullptr;      // slots relaxed bottom-up this round
    uint64_t pull_rounds = 0;
    uint32_t *keep_buf[64] = {};
    SlotState slots[64];
    uint64_t used = 0;
    uint64_t dirty = 0;  // finished slots whose visit bits are not cleared yet
    uint32_t round = 0;
    uint64_t rounds_run = 0;
    std::thread th;
    ~BfsEngine() {
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
        if (ev_summary) (void)hipEventDestroy(ev_summary);
    }
    void loop();
    void step();
    void finalize(uint64_t done);
    void fail_held(og_status st);
    KeepList keep_list(uint64_t mask) const {
        KeepList kl{};
        for (uint32_t j = 0; j < kSlots; ++j)
            if (((mask >> j) & 1) && (slots[j].q.flags & OG_QUERY_KEEP_VALUES)) kl.p[j] = keep_buf[j];
        return kl;
    }
};

struct SsspLane {
    og_server *srv = nullptr;
    og_graph *g = nullptr;
    DeviceState *lane = nullptr;
    uint32_t gi = 0;
    std::deque<Query> *pending = nullptr;
    uint64_t runs = 0, rounds_run = 0;
    Query cur{};
    std::thread th;
    ~SsspLane() {
        if (lane) device_state_free(lane);
    }
    void loop();
};

}  // namespace og

struct og_server {
    std::vector<og_graph *> graphs;
    std::vector<std::deque<og::Query>> queues;          // [graph * 2 + kind]
    std::vector<std::unique_ptr<og::BfsEngine>> bfs;    // [graph]
    std::vector<std::vector<std::unique_ptr<og::SsspLane>>> lanes;  // [graph], empty when unweighted
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    uint64_t next_id = 1;
    std::deque<og_query_result> done;
    std::unordered_map<uint64_t, std::vector<uint32_t>> kept;
    og_server_params params{};
    uint64_t submitted = 0, completed = 0, rejected = 0, pending = 0, in_flight = 0;
    uint64_t state_bytes = 0;
    og::Stopwatch clock;
};

namespace og {

static double now_ms(og_server *srv) { return srv->clock.ms(); }

static unsigned sweep_grid(uint64_t n, unsigned cap) {
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(cap, (n + kBlock - 1) / kBlock));
}

static void push_failure(og_server *srv, const Query &q, uint32_t gi, og_status st, double t) {
    og_query_result r{};
    r.id = q.id;
    r.graph = gi;
    r.kind = q.kind;
    r.source = q.source;
    r.status = st;
    r.latency_ms = t - q.t_submit;
    srv->done.push_back(r);
    srv->completed++;
}

// Results of the slots in `done` come from the accumulators copied at the
// end of the round; their visit bits stay set (the slots are `dirty`) until
// a clearing sweep, which runs for many slots at once when a pending query
// needs a slot or half of them are dirty.
void BfsEngine::finalize(uint64_t done) {
    const uint64_t V = d->V;
    const KeepList kl = keep_list(done);
    std::vector<std::pair<uint64_t, std::vector<uint32_t>>> vals;
    for (uint32_t j = 0; j < kSlots; ++j) {
        if (kl.p[j]) {
            vals.push_back({slots[j].q.id, std::vector<uint32_t>(V)});
            OG_HIP(hipMemcpyAsync(vals.back().second.data(), kl.p[j], V * 4, hipMemcpyDeviceToHost, s));
        }
    }
    hipLaunchKernelGGL(k_srv_zero_acc, dim3(1), dim3(64), 0, s, done, d_acc);
    OG_HIP(hipGetLastError());
    if (!vals.empty()) OG_HIP(hipStreamSynchronize(s));
    dirty |= done;
    const double t = now_ms(srv);
    std::lock_guard<std::mutex> lk(srv->mu);
    for (uint32_t j = 0; j < kSlots; ++j) {
        if (!((done >> j) & 1)) continue;
        og_query_result r{};
        r.id = slots[j].q.id;
        r.graph = gi;
        r.kind = OG_QUERY_BFS;
        r.source = slots[j].q.source;
        r.status = OG_OK;
        r.checksum = h_acc[j];
        r.reached = h_acc[64 + j];
        r.traversed_edges = h_acc[128 + j];
        r.rounds = round - slots[j].start_round;
        r.queue_ms = slots[j].t_admit - slots[j].q.t_submit;
        r.latency_ms = t - slots[j].q.t_submit;
        srv->done.push_back(r);
        slots[j].used = false;
        srv->completed++;
        srv->in_flight--;
    }
    for (auto &kv : vals) srv->kept[kv.first] = std::move(kv.second);
    used &= ~done;
    srv->cv_done.notify_all();
}

void BfsEngine::step() {
    SeedList sl{};
    {
        std::unique_lock<std::mutex> lk(srv->mu);
        srv->cv_work.wait(lk, [&] { return srv->stop || used || !pending->empty(); });
        if (srv->stop) return;
        const double t = now_ms(srv);
        const uint32_t clean = kSlots - (uint32_t)__builtin_popcountll(used | dirty);
        if (dirty && (pending->size() > clean || __builtin_popcountll(dirty) >= (int)kSlots / 2)) {
            hipLaunchKernelGGL(k_srv_clear_visit, dim3(sweep_grid(d->V, 16384)), dim3(kBlock), 0, s, d->V, dirty,
                               visit);
            OG_HIP(hipGetLastError());
            dirty = 0;  // stream order: the sweep runs before this round's seeds and commit
        }
        while (!pending->empty() && (uint32_t)__builtin_popcountll(used | dirty) < kSlots) {
            uint32_t j = 0;
            while (((used | dirty) >> j) & 1) ++j;
            slots[j].used = true;
            slots[j].q = pending->front();
            slots[j].start_round = round;
            slots[j].t_admit = t;
            pending->pop_front();
            used |= 1ull << j;
            sl.slot[sl.n] = j;
            sl.src[sl.n] = slots[j].q.source;
            sl.n++;
            srv->pending--;
            srv->in_flight++;
        }
    }
    const uint64_t V = d->V;
    for (uint32_t i = 0; i < sl.n; ++i) {  // kept levels are written by the commits
        const uint32_t j = sl.slot[i];
        if (slots[j].q.flags & OG_QUERY_KEEP_VALUES) {
            if (!keep_buf[j]) keep_buf[j] = (uint32_t *)mem.dalloc(V * 4);
            OG_HIP(hipMemsetAsync(keep_buf[j], 0xFF, V * 4, s));
        }
    }
    if (sl.n) {
        hipLaunchKernelGGL(k_srv_seed, dim3(1), dim3(64), 0, s, sl, next, f.flag, d_explored);
        OG_HIP(hipGetLastError());
    }
    const bool dir_opt = d->d_in_ptr != nullptr;
    OG_HIP(hipMemsetAsync(d_alive, 0, 8, s));
    SlotLevels L{};
    for (uint32_t j = 0; j < kSlots; ++j) L.lv[j] = round - slots[j].start_round + 1;
    hipLaunchKernelGGL(k_srv_commit, dim3(sweep_grid(V, 4096)), dim3(kBlock), 0, s, f.flag, V,
                       d->d_ptr, visit, front, next, L, keep_list(used), d_acc, d_fstat, d_alive);
    OG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_srv_direction, dim3(1), dim3(64), 0, s, used, V, d->E, dir_opt ? 1 : 0, d_fstat,
                       d_explored, d_acc, d_pull);
    OG_HIP(hipMemcpyAsync(h_alive, d_alive, 16, hipMemcpyDeviceToHost, s));
    // results of the slots that finish this round are complete after the commit
    OG_HIP(hipMemcpyAsync(h_acc, d_acc, 3 * 64 * 8, hipMemcpyDeviceToHost, s));
    OG_HIP(hipEventRecord(ev_summary, s));
    // Counts stay on the device: the compaction, relax and clear kernels read
    // them from f.info, so the round needs no host round trip before its end.
    frontier_compact(d, f, s, nullptr, false);
    RelaxArgs a{nullptr, f.flag, visit, front, next, d->wbits, dir_opt ? (const uint64_t *)d_pull : nullptr};
    const SrvPull pl{d->d_in_ptr, d->d_in_cols, V, visit, front, next, f.flag, d_pull};
    launch_srv_round(d, f, a, s, dir_opt ? &pl : nullptr, g->kind == OG_KIND_WEIGHTED);
    hipLaunchKernelGGL(k_srv_clear_front, dim3(1024), dim3(kBlock), 0, s, f.list, f.hub, f.hub_cap, f.info, front);
    OG_HIP(hipGetLastError());
    // The host needs only the commit's summary: while it finishes slots and
    // admits queries, the round's compaction and relax go on, and the next
    // round's launches queue behind them in stream order (its seeds and
    // commit see everything this round offered).
    OG_HIP(hipEventSynchronize(ev_summary));
    ++round;
    ++rounds_run;
    pull_rounds += h_alive[1] ? 1 : 0;
    // a slot is done when the round's commit left it no frontier vertex
    const uint64_t doneset = used & ~(uint64_t)*h_alive;
    if (doneset) finalize(doneset);
}

void BfsEngine::fail_held(og_status st) {
    std::lock_guard<std::mutex> lk(srv->mu);
    const double t = now_ms(srv);
    for (uint32_t j = 0; j < kSlots; ++j)
        if (slots[j].used) {
            push_failure(srv, slots[j].q, gi, st, t);
            slots[j].used = false;
            srv->in_flight--;
        }
    for (auto &q : *pending) {
        push_failure(srv, q, gi, st, t);
        srv->pending--;
    }
    pending->clear();
    used = 0;
    srv->cv_done.notify_all();
}

void BfsEngine::loop() {
    try {
        OG_HIP(hipSetDevice(d->device));
        for (;;) {
            {
                std::lock_guard<std::mutex> lk(srv->mu);
                if (srv->stop) return;
            }
            step();
        }
    } catch (const Error &e) {
        // fail everything this engine holds; the server keeps serving others
        fail_held(e.st);
    }
}

void SsspLane::loop() {
    try {
        OG_HIP(hipSetDevice(lane->device));
    } catch (const Error &) {
        return;
    }
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            srv->cv_work.wait(lk, [&] { return srv->stop || !pending->empty(); });
            if (srv->stop) return;
            cur = pending->front();
            pending->pop_front();
            srv->pending--;
            srv->in_flight++;
        }
        const double t_admit = now_ms(srv);
        std::vector<uint32_t> vals;
        og_run_stats st{};
        og_status status = OG_OK;
        try {
            if (cur.flags & OG_QUERY_KEEP_VALUES) vals.resize(lane->V);
            lane_run(g, lane, ALGO_SSSP, cur.source, vals.empty() ? nullptr : vals.data(), &st);
        } catch (const Error &e) {
            status = e.st;
        } catch (...) {
            status = OG_ENOMEM_HOST;
        }
        const double t = now_ms(srv);
        std::lock_guard<std::mutex> lk(srv->mu);
        og_query_result r{};
        r.id = cur.id;
        r.graph = gi;
        r.kind = OG_QUERY_SSSP;
        r.source = cur.source;
        r.status = status;
        if (status == OG_OK) {
            r.checksum = st.checksum;
            r.reached = st.reached;
            r.traversed_edges = st.traversed_edges;
            r.rounds = st.iterations;
            if (!vals.empty()) srv->kept[cur.id] = std::move(vals);
        }
        r.queue_ms = t_admit - cur.t_submit;
        r.latency_ms = t - cur.t_submit;
        srv->done.push_back(r);
        srv->completed++;
        srv->in_flight--;
        ++runs;
        rounds_run += st.iterations;
        srv->cv_done.notify_all();
    }
}

static void server_free(og_server *srv) {
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
    }
    srv->cv_work.notify_all();
    for (auto &e : srv->bfs)
        if (e && e->th.joinable()) e->th.join();
    for (auto &ls : srv->lanes)
        for (auto &l : ls)
            if (l && l->th.joinable()) l->th.join();
    srv->bfs.clear();
    srv->lanes.clear();
    for (og_graph *g : srv->graphs) g->busy.store(0);
    delete srv;
}

}  // namespace og

using namespace og;

extern "C" og_status og_server_create(og_graph *const *graphs, uint32_t num_graphs, const og_server_params *params,
                                      og_server **out) {
    clear_error();
    if (!graphs || !out || num_graphs == 0 || num_graphs > OG_SERVER_MAX_GRAPHS)
        return fail(OG_EINVAL, "need 1..%d graphs and an output pointer", OG_SERVER_MAX_GRAPHS);
    *out = nullptr;
    for (uint32_t i = 0; i < num_graphs; ++i) {
        og_graph *g = graphs[i];
        if (!g) return fail(OG_EINVAL, "graph %u is NULL", i);
        if (!g->dev) return fail(OG_ESTATE, "graph %u has not been prepared", i);
        if (g->kind == OG_KIND_PULL) return fail(OG_EUNSUPPORTED, "graph %u is a pull-layout graph", i);
        if (g->dev->plan.vsplit < g->V)
            return fail(OG_EUNSUPPORTED, "graph %u is not device resident (prepare it with the whole graph static)", i);
        for (uint32_t k = 0; k < i; ++k)
            if (graphs[k] == g) return fail(OG_EINVAL, "graph %u is listed twice", i);
    }
    og_server *srv = new og_server;
    if (params) srv->params = *params;
    uint32_t locked = 0;
    og_status st = guard([&]() -> og_status {
        for (; locked < num_graphs; ++locked) {
            int expect = 0;
            if (!graphs[locked]->busy.compare_exchange_strong(expect, 1))
                raise(OG_EBUSY, "graph %u is running another call", locked);
            srv->graphs.push_back(graphs[locked]);
        }
        // Admission: the BFS engines' frontier masks (24 V bytes per graph:
        // visit / front / next, 64 slots) and, per weighted graph, K SSSP
        // lanes of about 17 V bytes each (values, flags, list, degree
        // prefix): K is the largest of 4, 2, 1 that fits the budget.
        uint64_t fixed = 0, per_lane = 0;
        for (og_graph *g : srv->graphs) {
            fixed += 24 * g->V;
            if (g->kind == OG_KIND_WEIGHTED) per_lane += lane_bytes(g->V);
        }
        uint64_t budget = srv->params.state_budget_bytes;
        if (!budget) {
            size_t fr = 0, tot = 0;
            OG_HIP(hipMemGetInfo(&fr, &tot));
            budget = fr / 4;
        }
        uint32_t K = kMaxLanes;
        while (K > 0 && fixed + (uint64_t)K * per_lane > budget) K >>= 1;
        if (fixed > budget || (per_lane && K == 0))
            raise(OG_EBUDGET,
                  "state budget %llu bytes holds no query state (%llu bytes of BFS masks + %llu per SSSP lane set)",
                  (unsigned long long)budget, (unsigned long long)fixed, (unsigned long long)per_lane);
        srv->queues.resize(2 * num_graphs);
        srv->bfs.resize(num_graphs);
        srv->lanes.resize(num_graphs);
        for (uint32_t gi = 0; gi < num_graphs; ++gi) {
            og_graph *g = srv->graphs[gi];
            const uint64_t V = g->V;
            OG_HIP(hipSetDevice(g->dev->device));
            std::unique_ptr<BfsEngine> e(new BfsEngine);
            e->srv = srv;
            e->g = g;
            e->d = g->dev;
            e->gi = gi;
            e->pending = &srv->queues[2 * gi];
            OG_HIP(hipStreamCreateWithFlags(&e->s, hipStreamNonBlocking));
            OG_HIP(hipEventCreateWithFlags(&e->ev_summary, hipEventDisableTiming));
            frontier_alloc(g->dev, g, e->f, e->mem);
            e->visit = (uint64_t *)e->mem.dalloc(V * 8);
            e->front = (uint64_t *)e->mem.dalloc(V * 8);
            e->next = (uint64_t *)e->mem.dalloc(V * 8);
            OG_HIP(hipMemset(e->visit, 0, V * 8));
            OG_HIP(hipMemset(e->front, 0, V * 8));
            OG_HIP(hipMemset(e->next, 0, V * 8));
            e->d_alive = (unsigned long long *)e->mem.dalloc(16);  // [0] alive slots, [1] pulling slots
            e->h_alive = (unsigned long long *)e->mem.palloc(16);
            e->d_pull = e->d_alive + 1;
            e->d_fstat = (unsigned long long *)e->mem.dalloc(2 * 64 * 8);
            e->d_explored = (unsigned long long *)e->mem.dalloc(64 * 8);
            OG_HIP(hipMemset(e->d_alive, 0, 16));
            OG_HIP(hipMemset(e->d_fstat, 0, 2 * 64 * 8));
            OG_HIP(hipMemset(e->d_explored, 0, 64 * 8));
            e->d_acc = (unsigned long long *)e->mem.dalloc(3 * 64 * 8);
            e->h_acc = (unsigned long long *)e->mem.palloc(3 * 64 * 8);
            OG_HIP(hipMemset(e->d_acc, 0, 3 * 64 * 8));
            srv->state_bytes += 24 * V;
            srv->bfs[gi] = std::move(e);
            if (g->kind != OG_KIND_WEIGHTED) continue;
            for (uint32_t k = 0; k < K; ++k) {
                std::unique_ptr<SsspLane> l(new SsspLane);
                l->srv = srv;
                l->g = g;
                l->gi = gi;
                l->pending = &srv->queues[2 * gi + 1];
                l->lane = lane_state_create(g);
                srv->state_bytes += l->lane->device_bytes();
                srv->lanes[gi].push_back(std::move(l));
            }
        }
        OG_HIP(hipDeviceSynchronize());
        for (auto &e : srv->bfs) e->th = std::thread(&BfsEngine::loop, e.get());
        for (auto &ls : srv->lanes)
            for (auto &l : ls) l->th = std::thread(&SsspLane::loop, l.get());
        return OG_OK;
    });
    if (st != OG_OK) {
        const std::string msg = og_last_error();
        for (uint32_t i = 0; i < locked && i < srv->graphs.size(); ++i) srv->graphs[i]->busy.store(0);
        srv->graphs.clear();
        server_free(srv);
        return fail(st, "%s", msg.c_str());
    }
    *out = srv;
    return OG_OK;
}

extern "C" og_status og_server_submit(og_server *srv, uint32_t graph, uint32_t kind, uint32_t source, uint32_t flags,
                                      uint64_t *id) {
    clear_error();
    if (!srv) return fail(OG_EINVAL, "server is NULL");
    if (graph >= srv->graphs.size()) return fail(OG_EINVAL, "graph index %u out of range", graph);
    if (kind > 1) return fail(OG_EINVAL, "kind must be OG_QUERY_BFS or OG_QUERY_SSSP");
    if (kind == OG_QUERY_SSSP && srv->lanes[graph].empty())
        return fail(OG_EUNSUPPORTED, "graph %u has no weights: SSSP queries are not served", graph);
    if (source >= srv->graphs[graph]->V) return fail(OG_EINVAL, "source %u >= V", source);
    std::lock_guard<std::mutex> lk(srv->mu);
    if (srv->stop) return fail(OG_ESTATE, "server is stopping");
    const uint32_t cap = srv->params.max_pending ? srv->params.max_pending : 1u << 20;
    if (srv->pending >= cap) {
        srv->rejected++;
        return fail(OG_EBUSY, "pending queue full (%u queries)", cap);
    }
    Query q{srv->next_id++, graph, kind, source, flags, now_ms(srv)};
    srv->queues[2 * graph + kind].push_back(q);
    srv->pending++;
    srv->submitted++;
    if (id) *id = q.id;
    srv->cv_work.notify_all();
    return OG_OK;
}

extern "C" og_status og_server_poll(og_server *srv, og_query_result *results, uint32_t max_results, uint32_t *count,
                                    int32_t timeout_ms) {
    clear_error();
    if (!srv || !count || (max_results && !results)) return fail(OG_EINVAL, "null argument");
    std::unique_lock<std::mutex> lk(srv->mu);
    auto ready = [&] { return !srv->done.empty(); };
    if (timeout_ms < 0)
        srv->cv_done.wait(lk, ready);
    else if (timeout_ms > 0)
        srv->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
    uint32_t n = 0;
    while (n < max_results && !srv->done.empty()) {
        results[n++] = srv->done.front();
        srv->done.pop_front();
    }
    *count = n;
    return OG_OK;
}

extern "C" og_status og_server_take_values(og_server *srv, uint64_t id, uint32_t *values) {
    clear_error();
    if (!srv || !values) return fail(OG_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->kept.find(id);
    if (it == srv->kept.end()) return fail(OG_EINVAL, "no kept values for query %llu", (unsigned long long)id);
    std::memcpy(values, it->second.data(), it->second.size() * 4);
    srv->kept.erase(it);
    return OG_OK;
}

extern "C" og_status og_server_get_info(og_server *srv, og_server_info *info) {
    clear_error();
    if (!srv || !info) return fail(OG_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    std::memset(info, 0, sizeof *info);
    info->num_graphs = (uint32_t)srv->graphs.size();
    for (uint32_t gi = 0; gi < srv->graphs.size(); ++gi) {
        info->bfs_slots[gi] = kSlots;
        info->sssp_slots[gi] = (uint32_t)srv->lanes[gi].size();
        info->rounds += srv->bfs[gi]->rounds_run;
        for (auto &l : srv->lanes[gi]) info->rounds += l->rounds_run;
    }
    info->state_bytes = srv->state_bytes;
    info->submitted = srv->submitted;
    info->completed = srv->completed;
    info->rejected = srv->rejected;
    info->pending = srv->pending;
    info->in_flight = srv->in_flight;
    return OG_OK;
}

extern "C" og_status og_server_destroy(og_server *srv) {
    clear_error();
    if (!srv) return fail(OG_EINVAL, "server is NULL");
    server_free(srv);
    return OG_OK;
}
