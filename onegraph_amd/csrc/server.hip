// Query server: continuous batching of BFS and SSSP queries on device-
// resident graphs (no reference counterpart; the reference runs one query per
// process, BFS/oneGraph_BFS/main.cpp:3-14).
//
// Every (graph, kind) pair has one batch engine with S <= 64 query slots, its
// own HIP stream, frontier buffers and host thread.  A slot is one bit of the
// per-vertex u64 masks (BFS: visit / front / next; SSSP: front / next) and,
// for SSSP, one column of the vertex-major distance rows dist[V][S].  Each
// engine round ends with one host synchronisation:
//   1. admit pending queries into free slots: the source's bit goes into
//      `next`, so the commit below seeds it like any other offer;
//   2. commit (one sweep over the frontier flags): BFS turns the offered,
//      unseen slots of a vertex into its frontier mask and adds
//      mix64(v << 32 | level) to each such slot's digest, so no level array
//      is kept; SSSP takes the offered and held-back slots as the mask;
//   3. compact the flags into degree bins (shared with og_bfs);
//   4. relax: BFS through the multi-source bitmask kernels; SSSP per slot in
//      near-far order (every slot has its own threshold), one lane group per
//      (vertex, near slot) item, a hub tile staged in LDS once for all of its
//      near slots;
//   5. clear the masks of the relaxed vertices; SSSP moves the threshold of
//      every slot whose round had held-back work only;
//   6. slots whose frontier emptied are finalised by a sweep that counts
//      their reached vertices and traversed edges and clears their bits
//      (BFS) or resets their distance column (SSSP, which also computes the
//      digest there).
// Queries that arrive while a round runs join at the next round boundary, so
// their work overlaps with the queries already in flight.  Slot counts come
// from the HBM budget for per-query state given at creation (admission).
#include <algorithm>
#include <condition_variable>
#include <deque>
#include <unordered_map>

#include "device.h"

struct og_server;

namespace og {

constexpr uint32_t kUnr = 0xFFFFFFFFu;
constexpr uint32_t kServeDelta = 32;  // SSSP near-far width when the graph's tuning has none

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

// Seeds: slot bit into next[src] (the commit admits it); SSSP also sets the
// source's distance and the slot's first threshold.
__global__ void k_srv_seed(SeedList sl, int kind, uint32_t S, uint64_t *next, uint8_t *flag, uint32_t *dist,
                           uint32_t *thresh, uint32_t delta) {
    const uint32_t i = threadIdx.x;
    if (i < sl.n) {
        const uint32_t j = sl.slot[i], s = sl.src[i];
        atomicOr((unsigned long long *)(next + s), 1ull << j);
        flag[s] = 1;
        if (kind == 1) {
            dist[(uint64_t)s * S + j] = 1;
            thresh[j] = (uint32_t)min((uint64_t)kUnr, 1ull + delta);
        }
    }
}

// Commit sweep over the flags, 16 vertices per thread step (one uint4 load).
// BFS: nw = next & ~visit becomes the vertex's frontier mask (when it has
// out-edges) and each slot in nw adds the vertex's level to its digest in
// acc[0..63] (and writes it to a kept result array).  SSSP: next (offered
// or held-back slots) becomes the mask.  Flags end up 1 exactly for vertices
// with a non-empty mask; *alive gets the OR of all masks.
template <int KIND>
__global__ __launch_bounds__(kBlock) void k_srv_commit(uint8_t *__restrict__ flag, uint64_t V, uint64_t nchunks,
                                                       const uint64_t *__restrict__ ptr, uint64_t *__restrict__ visit,
                                                       uint64_t *__restrict__ front, uint64_t *__restrict__ next,
                                                       SlotLevels L, KeepList keep, unsigned long long *acc,
                                                       unsigned long long *alive) {
    __shared__ unsigned long long sdig[64];
    if (KIND == 0) {
        if (threadIdx.x < 64) sdig[threadIdx.x] = 0;
        __syncthreads();
    }
    uint64_t al = 0;
    for (uint64_t c = (uint64_t)blockIdx.x * kBlock + threadIdx.x; c < nchunks; c += (uint64_t)gridDim.x * kBlock) {
        const uint4 w = ((const uint4 *)flag)[c];
        if (!(w.x | w.y | w.z | w.w)) continue;
        const uint32_t q[4] = {w.x, w.y, w.z, w.w};
        for (int k = 0; k < 16; ++k) {
            if (!((q[k >> 2] >> (8 * (k & 3))) & 0xFFu)) continue;
            const uint64_t v = c * 16 + k;
            if (v >= V) break;
            const bool out = ptr[v + 1] != ptr[v];
            uint64_t nw = next[v];
            next[v] = 0;
            if (KIND == 0) {
                const uint64_t vis = visit[v];
                nw &= ~vis;
                if (nw) {
                    visit[v] = vis | nw;
                    uint64_t m = nw;
                    while (m) {
                        const int j = __builtin_ctzll(m);
                        m &= m - 1;
                        atomicAdd(&sdig[j], (unsigned long long)srv_mix64((v << 32) | L.lv[j]));
                        if (keep.p[j]) keep.p[j][v] = L.lv[j];
                    }
                }
            }
            const uint64_t fr = out ? nw : 0;
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
    if (KIND == 0) {
        __syncthreads();
        if (threadIdx.x < 64 && sdig[threadIdx.x]) atomicAdd(acc + threadIdx.x, sdig[threadIdx.x]);
    }
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

// SSSP relax step shared by the kernels below: lanes of a group relax edges
// of u for slot j with u's distance dj; an improvement ORs the slot's bit
// into next[to] and flags it.
template <bool PK>
__device__ __forceinline__ void srv_relax_edge(uint32_t to, uint32_t wt, uint32_t seen, uint32_t dj, uint32_t j,
                                               uint32_t S, uint32_t *dist, uint64_t *next, uint8_t *flag) {
    uint32_t nd = dj + wt;
    if (nd < dj) nd = 0xFFFFFFFEu;
    if (nd >= seen) return;
    if (nd < atomicMin(dist + (uint64_t)to * S + j, nd)) {
        atomicOr((unsigned long long *)(next + to), 1ull << j);
        flag[to] = 1;
    }
}

// Near / held-back split of u's frontier slots (lane = slot): a slot is near
// when u's distance is below the slot's threshold.  Held-back slots go back
// into next[u] (u stays flagged) and offer their distance to the slot's
// held minimum.
__device__ __forceinline__ uint64_t srv_split(uint32_t u, uint64_t F, uint32_t lane, uint32_t T, uint32_t S,
                                              const uint32_t *dist, uint64_t *next, uint8_t *flag, uint32_t &du,
                                              uint32_t *s_near, uint32_t *s_heldmin) {
    const bool act = (F >> lane) & 1;
    du = act ? dist[(uint64_t)u * S + lane] : kUnr;
    const bool nr = act && du < T;
    const bool hd = act && !nr;
    const uint64_t N = __ballot(nr), H = __ballot(hd);
    if (nr) s_near[lane] = 1;
    if (hd) atomicMin(&s_heldmin[lane], du);
    if (H && lane == 0) {
        atomicOr((unsigned long long *)(next + u), (unsigned long long)H);
        flag[u] = 1;
    }
    return N;
}

__device__ __forceinline__ void srv_flush(const uint32_t *s_near, const uint32_t *s_heldmin, uint32_t *g_near,
                                          uint32_t *g_heldmin) {
    __syncthreads();
    if (threadIdx.x < 64) {
        if (s_near[threadIdx.x]) g_near[threadIdx.x] = 1;
        if (s_heldmin[threadIdx.x] != kUnr) atomicMin(g_heldmin + threadIdx.x, s_heldmin[threadIdx.x]);
    }
}

// Bins 0 / 1: one wave per listed vertex u.  Its near slots are ranked
// (ds_permute puts slot of rank k into lane k) and dealt to 64 / LPV lane
// groups; a group relaxes u's edges for its slot, LPV lanes striding.
template <int LPV, bool PK>
__global__ __launch_bounds__(kBlock) void k_srv_sssp_relax(const uint32_t *__restrict__ list, int bin,
                                                           const FrontierInfo *__restrict__ info,
                                                           const uint64_t *__restrict__ ptr,
                                                           const void *__restrict__ edges, uint32_t wbits, uint32_t S,
                                                           const uint64_t *__restrict__ front, uint32_t *dist,
                                                           uint64_t *next, uint8_t *flag,
                                                           const uint32_t *__restrict__ thresh, uint32_t *g_near,
                                                           uint32_t *g_heldmin) {
    __shared__ uint32_t s_near[64], s_heldmin[64];
    if (threadIdx.x < 64) {
        s_near[threadIdx.x] = 0;
        s_heldmin[threadIdx.x] = kUnr;
    }
    __syncthreads();
    constexpr int G = kWave / LPV;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t grp = lane / LPV, sub = lane % LPV;
    const uint32_t T = lane < S ? thresh[lane] : 0u;
    const uint32_t n = info->n[bin];
    const uint32_t *lst = list + (bin ? info->n[0] : 0u);
    const uint64_t nwaves = (uint64_t)gridDim.x * (kBlock / kWave);
    for (uint64_t item = ((uint64_t)blockIdx.x * kBlock + threadIdx.x) / kWave; item < n; item += nwaves) {
        const uint32_t u = lst[item];
        uint32_t du;
        const uint64_t N = srv_split(u, front[u], lane, T, S, dist, next, flag, du, s_near, s_heldmin);
        if (!N) continue;
        const uint64_t b = ptr[u], e = ptr[u + 1];
        const uint32_t np = (uint32_t)__popcll(N);
        const uint64_t below = N & ((1ull << lane) - 1ull);
        const uint32_t nb = (uint32_t)__popcll(below);
        const bool nr = (N >> lane) & 1;
        const uint32_t rank = nr ? nb : np + (lane - nb);  // a permutation of the lanes
        const uint32_t slot_at = (uint32_t)__builtin_amdgcn_ds_permute((int)(rank * 4), (int)lane);
        for (uint32_t r0 = 0; r0 < np; r0 += G) {
            const uint32_t want = r0 + grp;
            const bool has = want < np;
            const uint32_t j = (uint32_t)__shfl((int)slot_at, (int)(want & (kWave - 1)), kWave);
            const uint32_t dj = (uint32_t)__shfl((int)du, (int)j, kWave);
            if (!has) continue;
            for (uint64_t i = b + sub; i < e; i += (uint64_t)LPV * kRelaxBatch) {
                uint32_t to[kRelaxBatch], wt[kRelaxBatch], seen[kRelaxBatch];
#pragma unroll
                for (int k = 0; k < kRelaxBatch; ++k) {
                    to[k] = 0;
                    wt[k] = 0;
                    if (i + (uint64_t)k * LPV < e) load_edge<true, PK>(edges, i + (uint64_t)k * LPV, wbits, to[k], wt[k]);
                }
#pragma unroll
                for (int k = 0; k < kRelaxBatch; ++k)
                    seen[k] = (i + (uint64_t)k * LPV < e) ? dist[(uint64_t)to[k] * S + j] : 0u;
#pragma unroll
                for (int k = 0; k < kRelaxBatch; ++k)
                    if (i + (uint64_t)k * LPV < e) srv_relax_edge<PK>(to[k], wt[k], seen[k], dj, j, S, dist, next, flag);
            }
        }
    }
    srv_flush(s_near, s_heldmin, g_near, g_heldmin);
}

// Hub tiles: one workgroup per (vertex, tile); wave 0 splits the slots, the
// tile is staged in LDS once and every near slot is relaxed over it by the
// whole workgroup.
template <bool PK>
__global__ __launch_bounds__(kBlock) void k_srv_sssp_hub(const uint64_t *__restrict__ hub, uint64_t hub_cap,
                                                         const FrontierInfo *__restrict__ info, uint32_t tile,
                                                         const uint64_t *__restrict__ ptr,
                                                         const void *__restrict__ edges, uint32_t wbits, uint32_t S,
                                                         const uint64_t *__restrict__ front, uint32_t *dist,
                                                         uint64_t *next, uint8_t *flag,
                                                         const uint32_t *__restrict__ thresh, uint32_t *g_near,
                                                         uint32_t *g_heldmin) {
    extern __shared__ uint32_t s_edges[];
    __shared__ uint32_t s_near[64], s_heldmin[64], s_du[64];
    __shared__ unsigned long long s_N;
    if (threadIdx.x < 64) {
        s_near[threadIdx.x] = 0;
        s_heldmin[threadIdx.x] = kUnr;
    }
    __syncthreads();
    const uint64_t ntiles = min((uint64_t)info->n[2], hub_cap);
    const uint32_t T = threadIdx.x < S ? thresh[threadIdx.x] : 0u;
    for (uint64_t ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
        const uint64_t it = hub[ti];
        const uint32_t u = (uint32_t)it;
        const uint64_t b = ptr[u] + (it >> 32) * tile;
        const uint64_t e = min(ptr[u + 1], b + tile);
        const uint32_t cnt = (uint32_t)(e - b);
        if (threadIdx.x < kWave) {
            uint32_t du;
            const uint64_t N = srv_split(u, front[u], threadIdx.x, T, S, dist, next, flag, du, s_near, s_heldmin);
            s_du[threadIdx.x] = du;
            if (threadIdx.x == 0) s_N = N;
        }
        for (uint32_t k = threadIdx.x; k < cnt; k += kBlock) {
            if (PK) s_edges[k] = ((const uint32_t *)edges)[b + k];
            else ((uint2 *)s_edges)[k] = ((const uint2 *)edges)[b + k];
        }
        __syncthreads();
        uint64_t N = s_N;
        while (N) {
            const uint32_t j = (uint32_t)__builtin_ctzll(N);
            N &= N - 1;
            const uint32_t dj = s_du[j];
            for (uint32_t k0 = threadIdx.x; k0 < cnt; k0 += kRelaxBatch * kBlock) {
                uint32_t to[kRelaxBatch], wt[kRelaxBatch], seen[kRelaxBatch];
#pragma unroll
                for (int q = 0; q < kRelaxBatch; ++q) {
                    to[q] = 0;
                    wt[q] = 0;
                    if (k0 + q * kBlock < cnt) load_edge<true, PK>(s_edges, k0 + q * kBlock, wbits, to[q], wt[q]);
                }
#pragma unroll
                for (int q = 0; q < kRelaxBatch; ++q)
                    seen[q] = (k0 + q * kBlock < cnt) ? dist[(uint64_t)to[q] * S + j] : 0u;
#pragma unroll
                for (int q = 0; q < kRelaxBatch; ++q)
                    if (k0 + q * kBlock < cnt) srv_relax_edge<PK>(to[q], wt[q], seen[q], dj, j, S, dist, next, flag);
            }
        }
        __syncthreads();  // tile buffer, s_du and s_N are reused by the next tile
    }
    srv_flush(s_near, s_heldmin, g_near, g_heldmin);
}

// Per-slot near-far step: a used slot whose round had held-back work only
// moves its threshold to (smallest held-back distance + delta).
__global__ void k_srv_sssp_thresh(uint32_t S, uint64_t used, uint32_t delta, uint32_t *thresh, uint32_t *g_near,
                                  uint32_t *g_heldmin) {
    const uint32_t j = threadIdx.x;
    if (j >= 64) return;
    if (j < S && ((used >> j) & 1) && !g_near[j] && g_heldmin[j] != kUnr)
        thresh[j] = (uint32_t)min((uint64_t)kUnr, (uint64_t)g_heldmin[j] + delta);
    g_near[j] = 0;
    g_heldmin[j] = kUnr;
}

// BFS slots in `done`: reached vertices and traversed edges into
// acc[64 + j] / acc[128 + j], and their visit bits cleared for reuse.
__global__ __launch_bounds__(kBlock) void k_srv_bfs_finalize(uint64_t V, uint64_t done,
                                                             const uint64_t *__restrict__ ptr, uint64_t *visit,
                                                             unsigned long long *acc) {
    __shared__ unsigned long long s[128];
    if (threadIdx.x < 128) s[threadIdx.x] = 0;
    __syncthreads();
    for (uint64_t v = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v < V; v += (uint64_t)gridDim.x * kBlock) {
        const uint64_t x = visit[v];
        uint64_t dm = x & done;
        if (!dm) continue;
        visit[v] = x & ~dm;
        const uint64_t deg = ptr[v + 1] - ptr[v];
        while (dm) {
            const int j = __builtin_ctzll(dm);
            dm &= dm - 1;
            atomicAdd(&s[j], 1ull);
            if (deg) atomicAdd(&s[64 + j], (unsigned long long)deg);
        }
    }
    __syncthreads();
    if (threadIdx.x < 128 && s[threadIdx.x]) atomicAdd(acc + 64 + threadIdx.x, s[threadIdx.x]);
}

// SSSP slots in `done`: digest / reached / traversed into acc, optional copy
// of the column, and the column reset to unreached.  Lane = (row group,
// slot), kRows rows per wave in flight.
constexpr int kFinRows = 4;
__global__ __launch_bounds__(kBlock) void k_srv_sssp_finalize(uint64_t V, uint32_t S, uint64_t done,
                                                              const uint64_t *__restrict__ ptr, uint32_t *dist,
                                                              KeepList keep, unsigned long long *acc) {
    __shared__ unsigned long long s[3 * 64];
    for (int t = threadIdx.x; t < 3 * 64; t += kBlock) s[t] = 0;
    __syncthreads();
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t G = kWave / S, grp = lane / S, j = lane % S;
    const bool mine = (done >> j) & 1;
    const uint64_t wave = ((uint64_t)blockIdx.x * kBlock + threadIdx.x) / kWave;
    const uint64_t nwaves = (uint64_t)gridDim.x * kBlock / kWave;
    uint64_t h = 0, r = 0, t = 0;
    if (mine) {
        uint32_t *kp = keep.p[j];
        for (uint64_t v0 = wave * G + grp; v0 < V; v0 += (uint64_t)kFinRows * nwaves * G) {
            uint32_t x[kFinRows];
#pragma unroll
            for (int q = 0; q < kFinRows; ++q) {
                const uint64_t v = v0 + (uint64_t)q * nwaves * G;
                x[q] = v < V ? dist[v * S + j] : kUnr;
            }
#pragma unroll
            for (int q = 0; q < kFinRows; ++q) {
                const uint64_t v = v0 + (uint64_t)q * nwaves * G;
                if (v >= V) break;
                if (kp) kp[v] = x[q];
                if (x[q] != kUnr) {
                    ++r;
                    t += ptr[v + 1] - ptr[v];
                    h += srv_mix64((v << 32) | x[q]);
                    dist[v * S + j] = kUnr;
                }
            }
        }
        atomicAdd(&s[j], (unsigned long long)h);
        atomicAdd(&s[64 + j], (unsigned long long)r);
        atomicAdd(&s[128 + j], (unsigned long long)t);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 3 * 64; k += kBlock)
        if (s[k]) atomicAdd(acc + k, s[k]);
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

struct Engine {
    og_server *srv = nullptr;
    og_graph *g = nullptr;
    DeviceState *d = nullptr;
    uint32_t gi = 0;
    int kind = 0;  // 0 BFS, 1 SSSP
    uint32_t S = 0;
    uint32_t delta = kServeDelta;
    uint64_t state_bytes = 0;
    MemPool mem;
    Frontier f;
    hipStream_t s = nullptr;
    uint64_t *visit = nullptr, *front = nullptr, *next = nullptr;
    uint32_t *dist = nullptr;                                  // SSSP [V][S]
    uint32_t *thresh = nullptr, *g_near = nullptr, *g_heldmin = nullptr;  // SSSP, 64 each
    unsigned long long *d_alive = nullptr, *h_alive = nullptr, *d_acc = nullptr, *h_acc = nullptr;
    uint32_t *keep_buf[64] = {};
    std::deque<Query> pending;
    SlotState slots[64];
    uint64_t used = 0;
    uint32_t round = 0;
    uint64_t rounds_run = 0;
    std::thread th;
    ~Engine() {
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    }
    void loop();
    void step();
    void finalize(uint64_t done);
    KeepList keep_list(uint64_t mask) const {
        KeepList kl{};
        for (uint32_t j = 0; j < S; ++j)
            if (((mask >> j) & 1) && (slots[j].q.flags & OG_QUERY_KEEP_VALUES)) kl.p[j] = keep_buf[j];
        return kl;
    }
};

}  // namespace og

struct og_server {
    std::vector<og_graph *> graphs;
    std::vector<std::unique_ptr<og::Engine>> engines;  // [graph * 2 + kind], null when absent
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

void Engine::finalize(uint64_t done) {
    const uint64_t V = d->V;
    const KeepList kl = keep_list(done);
    if (kind == 0) {
        hipLaunchKernelGGL(k_srv_bfs_finalize, dim3(sweep_grid(V, 16384)), dim3(kBlock), 0, s, V, done, d->d_ptr,
                           visit, d_acc);
    } else {
        hipLaunchKernelGGL(k_srv_sssp_finalize, dim3(sweep_grid(V * S / kFinRows, 16384)), dim3(kBlock), 0, s, V, S,
                           done, d->d_ptr, dist, kl, d_acc);
    }
    OG_HIP(hipGetLastError());
    OG_HIP(hipMemcpyAsync(h_acc, d_acc, 3 * 64 * 8, hipMemcpyDeviceToHost, s));
    std::vector<std::pair<uint64_t, std::vector<uint32_t>>> vals;
    for (uint32_t j = 0; j < S; ++j) {
        if (kl.p[j]) {
            vals.push_back({slots[j].q.id, std::vector<uint32_t>(V)});
            OG_HIP(hipMemcpyAsync(vals.back().second.data(), kl.p[j], V * 4, hipMemcpyDeviceToHost, s));
        }
    }
    hipLaunchKernelGGL(k_srv_zero_acc, dim3(1), dim3(64), 0, s, done, d_acc);
    OG_HIP(hipGetLastError());
    OG_HIP(hipStreamSynchronize(s));
    const double t = now_ms(srv);
    std::lock_guard<std::mutex> lk(srv->mu);
    for (uint32_t j = 0; j < S; ++j) {
        if (!((done >> j) & 1)) continue;
        og_query_result r{};
        r.id = slots[j].q.id;
        r.graph = gi;
        r.kind = (uint32_t)kind;
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

template <bool PK>
static void launch_srv_sssp(Engine &e) {
    DeviceState *d = e.d;
    Frontier &f = e.f;
    const uint32_t lpv0 = std::min<uint32_t>(std::max<uint32_t>(d->tune.lpv0, 4), 16);
    // bin 0 (short rows): lane groups of lpv0 lanes; bin 1: one group per wave
    switch (lpv0) {
        case 4:
            hipLaunchKernelGGL((k_srv_sssp_relax<4, PK>), dim3(2048), dim3(kBlock), 0, e.s, f.list, 0, f.info,
                               d->d_ptr, d->d_static, d->wbits, e.S, e.front, e.dist, e.next, f.flag, e.thresh,
                               e.g_near, e.g_heldmin);
            break;
        case 8:
            hipLaunchKernelGGL((k_srv_sssp_relax<8, PK>), dim3(2048), dim3(kBlock), 0, e.s, f.list, 0, f.info,
                               d->d_ptr, d->d_static, d->wbits, e.S, e.front, e.dist, e.next, f.flag, e.thresh,
                               e.g_near, e.g_heldmin);
            break;
        default:
            hipLaunchKernelGGL((k_srv_sssp_relax<16, PK>), dim3(2048), dim3(kBlock), 0, e.s, f.list, 0, f.info,
                               d->d_ptr, d->d_static, d->wbits, e.S, e.front, e.dist, e.next, f.flag, e.thresh,
                               e.g_near, e.g_heldmin);
            break;
    }
    hipLaunchKernelGGL((k_srv_sssp_relax<64, PK>), dim3(2048), dim3(kBlock), 0, e.s, f.list, 1, f.info, d->d_ptr,
                       d->d_static, d->wbits, e.S, e.front, e.dist, e.next, f.flag, e.thresh, e.g_near, e.g_heldmin);
    hipLaunchKernelGGL((k_srv_sssp_hub<PK>), dim3(1024), dim3(kBlock), d->tune.tile * d->eb, e.s, f.hub, f.hub_cap,
                       f.info, d->tune.tile, d->d_ptr, d->d_static, d->wbits, e.S, e.front, e.dist, e.next, f.flag,
                       e.thresh, e.g_near, e.g_heldmin);
    OG_HIP(hipGetLastError());
}

void Engine::step() {
    SeedList sl{};
    {
        std::unique_lock<std::mutex> lk(srv->mu);
        srv->cv_work.wait(lk, [&] { return srv->stop || used || !pending.empty(); });
        if (srv->stop) return;
        const double t = now_ms(srv);
        while (!pending.empty() && (uint32_t)__builtin_popcountll(used) < S) {
            uint32_t j = 0;
            while ((used >> j) & 1) ++j;
            slots[j].used = true;
            slots[j].q = pending.front();
            slots[j].start_round = round;
            slots[j].t_admit = t;
            pending.pop_front();
            used |= 1ull << j;
            sl.slot[sl.n] = j;
            sl.src[sl.n] = slots[j].q.source;
            sl.n++;
            srv->pending--;
            srv->in_flight++;
        }
    }
    const uint64_t V = d->V;
    for (uint32_t i = 0; i < sl.n; ++i) {  // kept BFS levels are written by the commits
        const uint32_t j = sl.slot[i];
        if (kind == 0 && (slots[j].q.flags & OG_QUERY_KEEP_VALUES)) {
            if (!keep_buf[j]) keep_buf[j] = (uint32_t *)mem.dalloc(V * 4);
            OG_HIP(hipMemsetAsync(keep_buf[j], 0xFF, V * 4, s));
        } else if (kind == 1 && (slots[j].q.flags & OG_QUERY_KEEP_VALUES) && !keep_buf[j]) {
            keep_buf[j] = (uint32_t *)mem.dalloc(V * 4);
        }
    }
    if (sl.n) {
        hipLaunchKernelGGL(k_srv_seed, dim3(1), dim3(64), 0, s, sl, kind, S, next, f.flag, dist, thresh, delta);
        OG_HIP(hipGetLastError());
    }
    OG_HIP(hipMemsetAsync(d_alive, 0, 8, s));
    const uint64_t nchunks = d->Vpad / 16;
    SlotLevels L{};
    for (uint32_t j = 0; j < S; ++j) L.lv[j] = round - slots[j].start_round + 1;
    const KeepList kl = kind == 0 ? keep_list(used) : KeepList{};
    if (kind == 0)
        hipLaunchKernelGGL(k_srv_commit<0>, dim3(sweep_grid(nchunks, 4096)), dim3(kBlock), 0, s, f.flag, V, nchunks,
                           d->d_ptr, visit, front, next, L, kl, d_acc, d_alive);
    else
        hipLaunchKernelGGL(k_srv_commit<1>, dim3(sweep_grid(nchunks, 4096)), dim3(kBlock), 0, s, f.flag, V, nchunks,
                           d->d_ptr, visit, front, next, L, kl, d_acc, d_alive);
    OG_HIP(hipGetLastError());
    OG_HIP(hipMemcpyAsync(h_alive, d_alive, 8, hipMemcpyDeviceToHost, s));
    // Counts stay on the device: the compaction, relax and clear kernels read
    // them from f.info, so the round needs no host round trip before its end.
    frontier_compact(d, f, s, nullptr, false);
    if (kind == 0) {
        RelaxArgs a{nullptr, f.flag, visit, front, next, d->wbits};
        if (g->kind == OG_KIND_WEIGHTED) launch_static_relax<ALGO_MSBFS, true>(d, f, nullptr, a, s);
        else launch_static_relax<ALGO_MSBFS, false>(d, f, nullptr, a, s);
    } else if (d->packed) {
        launch_srv_sssp<true>(*this);
    } else {
        launch_srv_sssp<false>(*this);
    }
    hipLaunchKernelGGL(k_srv_clear_front, dim3(1024), dim3(kBlock), 0, s, f.list, f.hub, f.hub_cap, f.info, front);
    if (kind == 1) hipLaunchKernelGGL(k_srv_sssp_thresh, dim3(1), dim3(64), 0, s, S, used, delta, thresh, g_near,
                                      g_heldmin);
    OG_HIP(hipGetLastError());
    OG_HIP(hipStreamSynchronize(s));
    ++round;
    ++rounds_run;
    // a slot is done when the round's commit left it no frontier vertex
    const uint64_t doneset = used & ~(uint64_t)*h_alive;
    if (doneset) finalize(doneset);
}

void Engine::loop() {
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
        std::lock_guard<std::mutex> lk(srv->mu);
        const double t = now_ms(srv);
        auto fail_q = [&](const Query &q) {
            og_query_result r{};
            r.id = q.id;
            r.graph = gi;
            r.kind = (uint32_t)kind;
            r.source = q.source;
            r.status = e.st;
            r.latency_ms = t - q.t_submit;
            srv->done.push_back(r);
            srv->completed++;
        };
        for (uint32_t j = 0; j < 64; ++j)
            if (slots[j].used) {
                fail_q(slots[j].q);
                srv->in_flight--;
            }
        for (auto &q : pending) {
            fail_q(q);
            srv->pending--;
        }
        pending.clear();
        used = 0;
        srv->cv_done.notify_all();
    }
}

static void server_free(og_server *srv) {
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
    }
    srv->cv_work.notify_all();
    for (auto &e : srv->engines)
        if (e && e->th.joinable()) e->th.join();
    srv->engines.clear();
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
        // Admission: per-query state is the slot bits of the frontier masks
        // (24 V bytes per BFS engine: visit / front / next; 16 V per SSSP
        // engine: front / next) plus one 4 V distance column per SSSP slot.
        // Every engine gets the same power-of-two slot count S <= 64, the
        // largest whose state fits the budget.
        uint64_t fixed = 0, per_slot = 0;
        for (og_graph *g : srv->graphs) {
            fixed += 24 * g->V;
            if (g->kind == OG_KIND_WEIGHTED) {
                fixed += 16 * g->V;
                per_slot += 4 * g->V;
            }
        }
        uint64_t budget = srv->params.state_budget_bytes;
        if (!budget) {
            size_t fr = 0, tot = 0;
            OG_HIP(hipMemGetInfo(&fr, &tot));
            budget = fr / 4;
        }
        uint32_t S = 64;
        while (S > 0 && fixed + (uint64_t)S * per_slot > budget) S >>= 1;
        if (S == 0)
            raise(OG_EBUDGET,
                  "state budget %llu bytes holds no query slot (%llu bytes of masks + %llu per slot across engines)",
                  (unsigned long long)budget, (unsigned long long)fixed, (unsigned long long)per_slot);
        srv->engines.resize(2 * num_graphs);
        for (uint32_t gi = 0; gi < num_graphs; ++gi) {
            og_graph *g = srv->graphs[gi];
            for (int kind = 0; kind < 2; ++kind) {
                if (kind == 1 && g->kind != OG_KIND_WEIGHTED) continue;
                std::unique_ptr<Engine> e(new Engine);
                e->srv = srv;
                e->g = g;
                e->d = g->dev;
                e->gi = gi;
                e->kind = kind;
                e->S = S;
                e->delta = g->dev->tune.delta ? g->dev->tune.delta : kServeDelta;
                const uint64_t V = g->V;
                OG_HIP(hipSetDevice(g->dev->device));
                OG_HIP(hipStreamCreateWithFlags(&e->s, hipStreamNonBlocking));
                frontier_alloc(g->dev, g, e->f, e->mem);
                e->front = (uint64_t *)e->mem.dalloc(V * 8);
                e->next = (uint64_t *)e->mem.dalloc(V * 8);
                OG_HIP(hipMemset(e->front, 0, V * 8));
                OG_HIP(hipMemset(e->next, 0, V * 8));
                if (kind == 0) {
                    e->visit = (uint64_t *)e->mem.dalloc(V * 8);
                    OG_HIP(hipMemset(e->visit, 0, V * 8));
                    e->state_bytes = 24 * V;
                } else {
                    e->dist = (uint32_t *)e->mem.dalloc(V * 4 * S);
                    OG_HIP(hipMemset(e->dist, 0xFF, V * 4 * S));
                    e->thresh = (uint32_t *)e->mem.dalloc(64 * 4);
                    e->g_near = (uint32_t *)e->mem.dalloc(64 * 4);
                    e->g_heldmin = (uint32_t *)e->mem.dalloc(64 * 4);
                    OG_HIP(hipMemset(e->thresh, 0xFF, 64 * 4));
                    OG_HIP(hipMemset(e->g_near, 0, 64 * 4));
                    OG_HIP(hipMemset(e->g_heldmin, 0xFF, 64 * 4));
                    e->state_bytes = 16 * V + 4 * V * S;
                }
                e->d_alive = (unsigned long long *)e->mem.dalloc(8);
                e->h_alive = (unsigned long long *)e->mem.palloc(8);
                e->d_acc = (unsigned long long *)e->mem.dalloc(3 * 64 * 8);
                e->h_acc = (unsigned long long *)e->mem.palloc(3 * 64 * 8);
                OG_HIP(hipMemset(e->d_acc, 0, 3 * 64 * 8));
                srv->state_bytes += e->state_bytes;
                srv->engines[2 * gi + kind] = std::move(e);
            }
        }
        OG_HIP(hipDeviceSynchronize());
        for (auto &e : srv->engines)
            if (e) e->th = std::thread(&Engine::loop, e.get());
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
    Engine *e = srv->engines[2 * graph + kind].get();
    if (!e) return fail(OG_EUNSUPPORTED, "graph %u has no weights: SSSP queries are not served", graph);
    if (source >= srv->graphs[graph]->V) return fail(OG_EINVAL, "source %u >= V", source);
    std::lock_guard<std::mutex> lk(srv->mu);
    if (srv->stop) return fail(OG_ESTATE, "server is stopping");
    const uint32_t cap = srv->params.max_pending ? srv->params.max_pending : 1u << 20;
    if (srv->pending >= cap) {
        srv->rejected++;
        return fail(OG_EBUSY, "pending queue full (%u queries)", cap);
    }
    Query q{srv->next_id++, graph, kind, source, flags, now_ms(srv)};
    e->pending.push_back(q);
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
        if (srv->engines[2 * gi]) info->bfs_slots[gi] = srv->engines[2 * gi]->S;
        if (srv->engines[2 * gi + 1]) info->sssp_slots[gi] = srv->engines[2 * gi + 1]->S;
        for (int k = 0; k < 2; ++k)
            if (srv->engines[2 * gi + k]) info->rounds += srv->engines[2 * gi + k]->rounds_run;
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
