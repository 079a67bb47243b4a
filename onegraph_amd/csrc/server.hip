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
// SSSP: every weighted graph has single-source lanes (kSsLanes unless
// og_server_params.sssp_lanes says otherwise), each running og_sssp's
// captured device-driven rounds on state of its own, one query at a time.
// On request (OG_SERVER_SSSP_BATCH) a batched engine with kSs = 16 query
// slots joins them (see "batched SSSP" below): each vertex keeps the 16
// slots' distances in one 64-byte line, every slot runs near-far rounds with
// its own threshold, the slots near at a vertex share the reads of its
// out-edges, and queries join at round boundaries like BFS queries.  It
// admits queries only while every lane is busy and it already runs queries
// or kSsBatchMin of them wait.  Lanes are the default because they measured
// better in the mixed workload: a lone query takes ~25 single-source rounds
// on a lane against ~30 rounds of full-width sweeps in the batched engine,
// and a busy batched engine takes the GPU from the BFS engines.  (A 64-slot
// distance matrix, 256 bytes per vertex, was tried first and lost to the
// caches.)
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
// slot's explored-edge count and its sums (digest, reached, traversed in
// acc[j], acc[64 + j], acc[128 + j]) restart: a finished slot keeps its last
// sums until it takes another query, and nothing adds to an unused slot's.
__global__ void k_srv_seed(SeedList sl, uint64_t *next, uint8_t *flag, unsigned long long *explored,
                           unsigned long long *acc) {
    const uint32_t i = threadIdx.x;
    if (i < sl.n) {
        const uint32_t j = sl.slot[i];
        atomicOr((unsigned long long *)(next + sl.src[i]), 1ull << j);
        flag[sl.src[i]] = 1;
        explored[j] = 0;
        acc[j] = acc[64 + j] = acc[128 + j] = 0;
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
// be read from memory once per (vertex, slot).  A slot's reached-vertex
// count and their out-degrees share one word (count << kFcShift | degrees),
// so a (vertex, slot) costs two LDS atomics, not three: a workgroup commits
// fewer than 2^24 vertices (at most 4096 workgroups sweep V < 2^36) whose
// degrees sum below 2^40 (the host refuses graphs with E >= 2^40).
constexpr int kCommitCopies = 16;
constexpr int kCommitStride = 65;
constexpr int kFcShift = 40;
__global__ __launch_bounds__(kBlock) void k_srv_commit(uint8_t *__restrict__ flag, uint64_t V,
                                                       const uint64_t *__restrict__ ptr,
                                                       const uint32_t *__restrict__ deg,
                                                       const uint32_t *__restrict__ hdeg, uint64_t *__restrict__ visit,
                                                       uint64_t *__restrict__ front, uint64_t *__restrict__ next,
                                                       SlotLevels L, KeepList keep, unsigned long long *acc,
                                                       unsigned long long *fstat, unsigned long long *alive) {
    __shared__ unsigned long long sdig[kCommitCopies * kCommitStride], sfe[kCommitCopies * kCommitStride];
    __shared__ uint32_t slv[64];
    __shared__ uint32_t *skeep[64];
    __shared__ unsigned long long skm;
    for (int i = threadIdx.x; i < kCommitCopies * kCommitStride; i += kBlock) sdig[i] = sfe[i] = 0;
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
            const uint64_t dv = round_deg(deg, hdeg, ptr, nullptr, false, v);
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
                    atomicAdd(&sfe[cp + j], (1ull << kFcShift) | (unsigned long long)dv);
                    if ((kmask >> j) & 1) skeep[j][v] = lv;
                }
            }
            const uint64_t fr = dv ? nw : 0;
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
        unsigned long long dg = 0, x = 0;
        for (int c = 0; c < kCommitCopies; ++c) {
            dg += sdig[c * kCommitStride + threadIdx.x];
            x += sfe[c * kCommitStride + threadIdx.x];
        }
        const unsigned long long fc = x >> kFcShift, fe = x & ((1ull << kFcShift) - 1);
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
// SC'12, per slot): top-down -> bottom-up once its frontier edges exceed
// 1/kSrvAlpha of its unexplored edges, back when its frontier holds fewer
// than V / 24 vertices.  kSrvAlpha is twice the single-source BFS's 14: the
// bottom-up pass sweeps V once for every pulling slot of the round, so in a
// batch it pays off earlier (R-MAT 26, 333 queries/s: rounds 3.7 -> 3.3 ms,
// p99 59 -> 45 ms; 56 gave the same as 28).  Bit j of *pull = slot j relaxes
// bottom-up.  Resets the stats, and moves the commit's alive mask from
// sum[0] (which the next commit ORs into) to sum[2], next to the pull mask
// sum[1] and the slot sums acc = sum + 3, so one copy brings the round's
// summary to the host and no memset clears sum[0] before the next commit.
constexpr uint64_t kSrvAlpha = 28;
__global__ void k_srv_direction(uint64_t used, uint64_t V, uint64_t E, int dir_opt, unsigned long long *fstat,
                                unsigned long long *explored, unsigned long long *sum) {
    unsigned long long *const pull = sum + 1, *const acc = sum + 3;
    const uint32_t j = threadIdx.x;
    const uint64_t fc = fstat[j], fe = fstat[64 + j];
    acc[64 + j] += fc;
    acc[128 + j] += fe;
    const uint64_t ex = explored[j];
    const uint64_t unexplored = E > ex + fe ? E - ex - fe : 0;
    const bool was = (*pull >> j) & 1;
    bool now = false;
    if (dir_opt && ((used >> j) & 1)) now = was ? fc >= V / 24 : fe > unexplored / kSrvAlpha;
    explored[j] = ex + fe;
    fstat[j] = fstat[64 + j] = 0;
    const uint64_t b = __ballot(now);
    if (j == 0) {
        *pull = b;
        sum[2] = sum[0];
        sum[0] = 0;
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

// Clears the visit bits of finished slots (before the slots are reused).
__global__ __launch_bounds__(kBlock) void k_srv_clear_visit(uint64_t V, uint64_t dirty, uint64_t *visit) {
    for (uint64_t v = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v < V; v += (uint64_t)gridDim.x * kBlock) {
        const uint64_t x = visit[v];
        if (x & dirty) visit[v] = x & ~dirty;
    }
}


// ---------------------------------------------------- batched SSSP (server)
//
// Every weighted graph has one SSSP engine with kSs query slots.  Vertex v
// keeps the kSs slots' distances side by side in one 64-byte line
// (dist[v * kSs + j], distance + 1 as og_sssp stores it), and a pending mask
// pend[v] (bit j: slot j improved v and has not relaxed it since).  Each
// slot runs near-far SSSP with its own threshold thresh[j]: a round relaxes
// the pending (v, j) with dist < thresh[j]; when a slot has nothing near
// left but far entries, its threshold moves to (smallest far value + delta)
// for the next round.  On a graph stored with the light/heavy split (rows
// list edges of weight < split_weight first, pmid[v] = start of the heavy
// ones; delta-stepping, Meyer & Sanders) a near (v, j) relaxes only its light
// edges and turns heavy-pending (hpend[v] bit j); once slot j has nothing
// near in its bucket, one heavy round relaxes the heavy edges of all its
// heavy-pending vertices, then its light rounds go on.  The slots of a round
// share the edge reads: an active vertex carries the masks of its light and
// heavy slots, and one pass over its edges offers every such slot.  Rounds:
//   select  one sweep over the pending masks: per pending vertex, the near
//           mask from its distance line (light-mode slots) and the
//           heavy-pending slots of heavy-mode slots; per-block counts,
//           masks and per-slot far minima;
//   scan    one workgroup: list offsets, each slot's next mode (light,
//           heavy, threshold move, done), the summary the host reads;
//   write   the active list (vertex, light | heavy << 16 slot masks), the
//           prefix of its work items and the chunk table of the relax;
//   relax   one lane per (edge, slot) work item, atomicMin per improved
//           slot, one atomicOr of it into the target's pending mask.
// Results need no extra pass: every successful atomicMin of (v, j) from old
// to new adds mix64(v << 32 | new) - mix64(v << 32 | old) (old = unreached
// adds nothing and counts v as reached, with its out-degree as traversed) to
// slot j's sums, so the sums telescope to og_levels_checksum of the final
// distances.  A finished slot's column is reset to "unreached" by a later
// sweep shared by many slots, before the slot takes another query.

constexpr uint32_t kSs = 16;                  // SSSP slots per graph (distances per 64-byte vertex line)
constexpr uint32_t kSsWin = 8192;             // vertices per workgroup of the select / write sweeps
constexpr uint32_t kSsPer = kSsWin / kBlock;  // per thread: vertices base + t + kBlock * k, k < kSsPer
constexpr uint32_t kSsRun = 4;                // work items per lane per chunk (consecutive)
constexpr uint32_t kSsChunk = kBlock * kSsRun;  // work items per relax workgroup iteration
constexpr unsigned kSsRelaxGridCap = 4096;    // relax workgroups at most (one chunk each below it); each
                                              // workgroup adds its per-slot sums with 48 global atomics

struct SsCtl {
    uint32_t thresh[kSs];  // near-far threshold of each slot (0xFFFFFFFF: every pending entry is near)
    uint32_t delta;        // bucket width (0: no near-far order)
    uint32_t nactive;      // vertices in this round's active list
    uint64_t dsum;         // their work items (the relax's item space)
    uint32_t near, far;    // slots with near / far light-pending entries after the select sweep
    uint32_t hp;           // slots with heavy-pending entries after the select sweep
    uint32_t heavy;        // slots in a heavy round (the select sweep reads it; the scan sets the next)
    uint32_t done;         // slots in use with nothing pending: finished
    uint32_t rounds;       // select sweeps so far
};

struct SsSeeds {
    uint32_t n;
    uint32_t slot[kSs];
    uint32_t src[kSs];
};

// New queries: slot j starts at its source (value 1, pending, light mode),
// its sums at the source's own terms, its threshold at 1 + delta.
__global__ void k_ss_seed(SsSeeds sl, uint32_t *dist, uint32_t *pend, const uint64_t *__restrict__ ptr,
                          unsigned long long *acc, SsCtl *ctl) {
    const uint32_t i = threadIdx.x;
    if (i >= sl.n) return;
    const uint32_t j = sl.slot[i], v = sl.src[i];
    dist[(uint64_t)v * kSs + j] = 1;
    atomicOr(pend + v, 1u << j);
    atomicAnd(&ctl->heavy, ~(1u << j));
    acc[j] = srv_mix64(((uint64_t)v << 32) | 1u);
    acc[kSs + j] = 1;
    acc[2 * kSs + j] = ptr[v + 1] - ptr[v];
    const uint32_t dl = ctl->delta;
    ctl->thresh[j] = dl ? (uint32_t)min((uint64_t)0xFFFFFFFFull, 1ull + dl) : 0xFFFFFFFFu;
}

__device__ __forceinline__ uint32_t quad_at(const uint4 &x, int b) {
    return b == 0 ? x.x : b == 1 ? x.y : b == 2 ? x.z : x.w;
}

// Light and heavy edge counts of row v (no split: the whole row is light).
__device__ __forceinline__ void row_parts(const uint64_t *__restrict__ ptr, const uint64_t *__restrict__ pmid,
                                          uint64_t v, uint64_t &ld, uint64_t &hd) {
    const uint64_t b = ptr[v], e = ptr[v + 1];
    const uint64_t m = pmid ? pmid[v] : e;
    ld = m - b;
    hd = e - m;
}

// Select sweep.  Per pending vertex: L = light-mode slots with a pending
// entry below their threshold (their far entries stay pending; the block's
// minimum far value per slot is kept with LDS atomicMin, tried only when it
// would lower it), H = heavy-pending slots of heavy-mode slots.  Selected
// slots move to the upper halves (pend = L << 16 | far, hpend = H << 16 |
// still heavy-pending) when the vertex has edges of that kind; near (v, j)
// turn heavy-pending when v has heavy edges.  Per block: listed vertices,
// their work items, and the OR of near / far / heavy-pending masks.
__global__ __launch_bounds__(kBlock) void k_ss_select(uint32_t *__restrict__ pend, uint32_t *__restrict__ hpend,
                                                      const uint32_t *__restrict__ dist,
                                                      const uint64_t *__restrict__ ptr,
                                                      const uint64_t *__restrict__ pmid, uint64_t V,
                                                      const SsCtl *__restrict__ ctl, uint32_t *__restrict__ bnv,
                                                      uint64_t *__restrict__ bdeg, uint32_t *__restrict__ bmask,
                                                      uint32_t *__restrict__ bmin, uint32_t nblk) {
    __shared__ uint32_t s_thr[kSs], s_min[kSs];
    __shared__ uint64_t lds[2 * kBlock / kWave];
    __shared__ uint32_t s_or[3];
    if (threadIdx.x < kSs) {
        s_thr[threadIdx.x] = ctl->thresh[threadIdx.x];
        s_min[threadIdx.x] = 0xFFFFFFFFu;
    }
    if (threadIdx.x < 3) s_or[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t heavy = ctl->heavy;
    const uint64_t base = (uint64_t)blockIdx.x * kSsWin + threadIdx.x;
    uint32_t cnt = 0, nm = 0, fm = 0, hm = 0;
    uint64_t items = 0;
    for (uint32_t k = 0; k < kSsPer; ++k) {
        const uint64_t v = base + (uint64_t)k * kBlock;
        if (v >= V) break;
        const uint32_t p = pend[v] & 0xFFFFu;
        const uint32_t h = pmid ? hpend[v] & 0xFFFFu : 0u;
        if (!(p | h)) continue;
        const uint32_t pl = p & ~heavy;  // light candidates (heavy-mode slots keep theirs pending)
        uint32_t near = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t qb = (pl >> (4 * q)) & 0xFu;
            if (!qb) continue;
            const uint4 x = *(const uint4 *)(dist + v * kSs + 4 * q);
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                if (!((qb >> b) & 1u)) continue;
                const int j = 4 * q + b;
                const uint32_t d = quad_at(x, b);
                if (d < s_thr[j]) near |= 1u << j;
                else if (d < s_min[j]) atomicMin(&s_min[j], d);
            }
        }
        const uint32_t far = p & ~near;
        const uint32_t H = h & heavy;
        uint64_t ld = 0, hd = 0;
        if (near | H) row_parts(ptr, pmid, v, ld, hd);
        // slots selected with no edges of their kind have nothing to relax:
        // their entries are consumed here and do not enter the list (so
        // every listed vertex owns at least one work item)
        const uint32_t L = ld ? near : 0u, Hs = hd ? H : 0u;
        pend[v] = (L << 16) | far;
        const uint32_t hrest = (h & ~H) | (hd ? near : 0u);
        if (pmid) hpend[v] = (Hs << 16) | hrest;
        nm |= near;
        fm |= far;
        hm |= hrest;
        const uint64_t it = ld * (uint32_t)__builtin_popcount(L) + hd * (uint32_t)__builtin_popcount(Hs);
        if (it) {
            ++cnt;
            items += it;
        }
    }
    const uint64_t x[2] = {cnt, items};
    uint64_t ex[2], t[2];
    block_scan_n<2>(x, ex, t, lds);
    nm = (uint32_t)wave_or(nm);
    fm = (uint32_t)wave_or(fm);
    hm = (uint32_t)wave_or(hm);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        if (nm) atomicOr(&s_or[0], nm);
        if (fm) atomicOr(&s_or[1], fm);
        if (hm) atomicOr(&s_or[2], hm);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        bnv[blockIdx.x] = (uint32_t)t[0];
        bdeg[blockIdx.x] = t[1];
        bmask[blockIdx.x] = s_or[0];
        bmask[nblk + blockIdx.x] = s_or[1];
        bmask[2 * nblk + blockIdx.x] = s_or[2];
    }
    if (threadIdx.x < kSs) bmin[(uint64_t)threadIdx.x * nblk + blockIdx.x] = s_min[threadIdx.x];
}

// Scan: exclusive offsets of the blocks' list counts and work items (in
// place), and each slot's next round: after a heavy round, light; with
// anything near, light again; else a heavy round when it has heavy-pending
// vertices; else, with far entries only, its threshold moves to the
// smallest far value + delta; with nothing at all it is done.  Publishes
// the list's total work at doff[nactive].
__global__ __launch_bounds__(1024) void k_ss_scan(uint32_t *__restrict__ bnv, uint64_t *__restrict__ bdeg,
                                                  const uint32_t *__restrict__ bmask, const uint32_t *__restrict__ bmin,
                                                  uint32_t nblk, uint32_t used, SsCtl *__restrict__ ctl,
                                                  uint64_t *__restrict__ doff) {
    __shared__ uint64_t lds[2 * 1024 / kWave];
    __shared__ uint32_t s_or[3], s_min[kSs];
    if (threadIdx.x < 3) s_or[threadIdx.x] = 0;
    if (threadIdx.x < kSs) s_min[threadIdx.x] = 0xFFFFFFFFu;
    __syncthreads();
    const uint32_t per = (nblk + blockDim.x - 1) / blockDim.x;
    const uint32_t b0 = min(nblk, threadIdx.x * per), b1 = min(nblk, b0 + per);
    uint64_t c = 0, dg = 0;
    uint32_t nm = 0, fm = 0, hm = 0;
    for (uint32_t b = b0; b < b1; ++b) {
        c += bnv[b];
        dg += bdeg[b];
        nm |= bmask[b];
        fm |= bmask[nblk + b];
        hm |= bmask[2 * nblk + b];
    }
    const uint64_t x[2] = {c, dg};
    uint64_t ex[2], t[2];
    block_scan_n<2>(x, ex, t, lds);
    uint64_t rc = ex[0], rd = ex[1];
    for (uint32_t b = b0; b < b1; ++b) {
        const uint32_t a = bnv[b];
        const uint64_t e = bdeg[b];
        bnv[b] = (uint32_t)rc;
        bdeg[b] = rd;
        rc += a;
        rd += e;
    }
    nm = (uint32_t)wave_or(nm);
    fm = (uint32_t)wave_or(fm);
    hm = (uint32_t)wave_or(hm);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        if (nm) atomicOr(&s_or[0], nm);
        if (fm) atomicOr(&s_or[1], fm);
        if (hm) atomicOr(&s_or[2], hm);
    }
    __syncthreads();
    const uint32_t near = s_or[0], far = s_or[1], hp = s_or[2];
    const uint32_t was_heavy = ctl->heavy;
    const uint32_t light = ~was_heavy;
    const uint32_t next_heavy = light & ~near & hp;
    const uint32_t moving = light & ~near & ~hp & far;  // threshold moves for the next round
    for (uint32_t m = moving; m;) {
        const int j = __builtin_ctz(m);
        m &= m - 1;
        uint32_t mn = 0xFFFFFFFFu;
        for (uint32_t b = b0; b < b1; ++b) mn = min(mn, bmin[(uint64_t)j * nblk + b]);
        if (mn != 0xFFFFFFFFu) atomicMin(&s_min[j], mn);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        SsCtl cl = *ctl;
        cl.nactive = (uint32_t)t[0];
        cl.dsum = t[1];
        cl.near = near;
        cl.far = far;
        cl.hp = hp;
        cl.done = used & light & ~(near | far | hp);
        cl.heavy = next_heavy;
        cl.rounds++;
        for (uint32_t m = moving; m;) {
            const int j = __builtin_ctz(m);
            m &= m - 1;
            cl.thresh[j] = (uint32_t)min((uint64_t)0xFFFFFFFFull, (uint64_t)s_min[j] + cl.delta);
        }
        *ctl = cl;
        doff[t[0]] = t[1];
    }
}

// Write sweep: the active list in block order (vertex, light | heavy << 16
// slot masks), the prefix of its work items (light edges x light slots, then
// heavy edges x heavy slots), the list entry holding the first item of every
// relax chunk (cstart[c], chunks of kSsChunk items), and the pending masks
// back to their lower halves.
__global__ __launch_bounds__(kBlock) void k_ss_write(uint32_t *__restrict__ pend, uint32_t *__restrict__ hpend,
                                                     const uint64_t *__restrict__ ptr,
                                                     const uint64_t *__restrict__ pmid, uint64_t V,
                                                     const uint32_t *__restrict__ bnv,
                                                     const uint64_t *__restrict__ bdeg, uint32_t *__restrict__ alist,
                                                     uint32_t *__restrict__ amask, uint64_t *__restrict__ doff,
                                                     uint32_t *__restrict__ cstart) {
    __shared__ uint64_t lds[2 * kBlock / kWave];
    const uint64_t base = (uint64_t)blockIdx.x * kSsWin + threadIdx.x;
    auto items_of = [&](uint64_t v, uint32_t L, uint32_t H) -> uint64_t {
        uint64_t ld, hd;
        row_parts(ptr, pmid, v, ld, hd);
        return ld * (uint32_t)__builtin_popcount(L) + hd * (uint32_t)__builtin_popcount(H);
    };
    uint32_t cnt = 0;
    uint64_t items = 0;
    for (uint32_t k = 0; k < kSsPer; ++k) {
        const uint64_t v = base + (uint64_t)k * kBlock;
        if (v >= V) break;
        const uint32_t L = pend[v] >> 16, H = pmid ? hpend[v] >> 16 : 0u;
        if (L | H) {
            ++cnt;
            items += items_of(v, L, H);
        }
    }
    const uint64_t x[2] = {cnt, items};
    uint64_t ex[2], t[2];
    block_scan_n<2>(x, ex, t, lds);
    if (!t[0]) return;
    uint64_t pos = bnv[blockIdx.x] + ex[0], dpos = bdeg[blockIdx.x] + ex[1];
    for (uint32_t k = 0; k < kSsPer && cnt; ++k) {
        const uint64_t v = base + (uint64_t)k * kBlock;
        if (v >= V) break;
        const uint32_t p = pend[v], hp = pmid ? hpend[v] : 0u;
        const uint32_t L = p >> 16, H = hp >> 16;
        if (!(L | H)) continue;
        const uint64_t it = items_of(v, L, H);  // >= 1
        alist[pos] = (uint32_t)v;
        amask[pos] = L | (H << 16);
        doff[pos] = dpos;
        // chunks of the relax whose first item is one of this vertex's
        for (uint64_t c = (dpos + kSsChunk - 1) / kSsChunk; c * kSsChunk < dpos + it; ++c) cstart[c] = (uint32_t)pos;
        ++pos;
        dpos += it;
        if (L) pend[v] = p & 0xFFFFu;
        if (H) hpend[v] = hp & 0xFFFFu;
        --cnt;
    }
}

// Relax: one work item per (edge, slot) of the active list.  An entry's
// items come light part first (edge e of the light edges with the k-th of
// its light slots at item e * mL + k), then the heavy part likewise, so
// neighbouring lanes offer the slots of one edge to one 64-byte target line
// and no lane idles for a slot that is not active.  Workgroups stride over
// chunks of kSsChunk items; a chunk's entries start at cstart[c] and its
// slice of the item prefix is staged in LDS.  Each lane takes kSsRun
// consecutive items: one binary search, then it walks on.  Result terms are
// summed per wave in LDS and added to the slots' sums once per workgroup.
template <bool W, bool PK>
__global__ __launch_bounds__(kBlock) void k_ss_relax(const uint32_t *__restrict__ alist,
                                                     const uint32_t *__restrict__ amask,
                                                     const uint64_t *__restrict__ doff,
                                                     const uint32_t *__restrict__ cstart, const SsCtl *__restrict__ ctl,
                                                     const uint64_t *__restrict__ ptr,
                                                     const uint64_t *__restrict__ pmid, const void *__restrict__ edges,
                                                     uint32_t wbits, uint32_t *dist, uint32_t *pend,
                                                     unsigned long long *acc) {
    __shared__ uint64_t s_off[kSsChunk + 2];
    __shared__ unsigned long long s_acc[kBlock / kWave][3 * kSs];
    const uint32_t wv = threadIdx.x / kWave;
    for (uint32_t k = threadIdx.x; k < (kBlock / kWave) * 3 * kSs; k += kBlock) (&s_acc[0][0])[k] = 0;
    const uint32_t n = ctl->nactive;
    const uint64_t D = ctl->dsum;
    const uint64_t nch = (D + kSsChunk - 1) / kSsChunk;
    for (uint64_t c = blockIdx.x; c < nch; c += gridDim.x) {
        const uint64_t q0 = c * kSsChunk, q1 = min(q0 + kSsChunk, D);
        const uint32_t i0 = cstart[c];
        const uint32_t i1 = q1 == D ? n - 1 : cstart[c + 1];  // the entry holding q1 (or the last one)
        const uint32_t nv = i1 - i0 + 1;                       // <= kSsChunk + 1
        __syncthreads();  // the previous chunk's slice is no longer read
        for (uint32_t k = threadIdx.x; k <= nv; k += kBlock) s_off[k] = doff[i0 + k];
        __syncthreads();
        const uint64_t w0 = q0 + (uint64_t)kSsRun * threadIdx.x;
        if (w0 >= q1) continue;
        uint32_t to[kSsRun], wt[kSsRun], slot[kSsRun], du[kSsRun], dt[kSsRun];
        bool ok[kSsRun];
        uint32_t lo = 0, hi = nv;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (s_off[mid] <= w0) lo = mid; else hi = mid;
        }
        // walking state: the current part (light or heavy) of entry lo
        uint32_t u = 0, ms = 0, M = 0, mm = 0;
        uint64_t rb = 0, e = 0, pend_item = 0, lpart = 0;
        bool hv = false;
        auto enter = [&](uint32_t idx) {  // entry idx; the light part if it has items, else the heavy part
            u = alist[i0 + idx];
            ms = amask[i0 + idx];
            uint64_t ld, hd;
            row_parts(ptr, pmid, u, ld, hd);
            lpart = ld * (uint32_t)__builtin_popcount(ms & 0xFFFFu);
            hv = lpart == 0;
            M = hv ? ms >> 16 : ms & 0xFFFFu;
            rb = hv ? pmid[u] : ptr[u];
            pend_item = hv ? s_off[idx + 1] : s_off[idx] + lpart;
        };
        enter(lo);
        {
            uint64_t off = w0 - (hv ? s_off[lo] : s_off[lo]);
            if (!hv && off >= lpart) {  // w0 lies in the heavy part
                off -= lpart;
                hv = true;
                M = ms >> 16;
                rb = pmid[u];
                pend_item = s_off[lo + 1];
            }
            const uint32_t m = (uint32_t)__builtin_popcount(M);
            e = off / m;
            mm = M;
            for (uint32_t t = (uint32_t)(off - e * m); t; --t) mm &= mm - 1;
        }
#pragma unroll
        for (uint32_t r = 0; r < kSsRun; ++r) {
            const uint64_t w = w0 + r;
            ok[r] = w < q1;
            if (!ok[r]) continue;
            if (w >= pend_item) {  // the next part: this entry's heavy part, or the next entry
                if (!hv && pend_item < s_off[lo + 1]) {
                    hv = true;
                    M = ms >> 16;
                    rb = pmid[u];
                    pend_item = s_off[lo + 1];
                } else {
                    ++lo;
                    enter(lo);
                }
                mm = M;
                e = 0;
            }
            const uint32_t j = (uint32_t)__builtin_ctz(mm);
            slot[r] = j;
            load_edge<W, PK>(edges, rb + e, wbits, to[r], wt[r]);
            du[r] = dist[(uint64_t)u * kSs + j];
            mm &= mm - 1;
            if (!mm) {
                mm = M;
                ++e;
            }
        }
#pragma unroll
        for (uint32_t r = 0; r < kSsRun; ++r)
            if (ok[r]) dt[r] = dist[(uint64_t)to[r] * kSs + slot[r]];
#pragma unroll
        for (uint32_t r = 0; r < kSsRun; ++r) {
            if (!ok[r]) continue;
            const uint64_t s64 = (uint64_t)du[r] + wt[r];
            const uint32_t nd = s64 > 0xFFFFFFFEull ? 0xFFFFFFFEu : (uint32_t)s64;
            if (nd >= dt[r]) continue;
            const uint32_t j = slot[r];
            const uint32_t old = atomicMin(dist + (uint64_t)to[r] * kSs + j, nd);
            if (nd >= old) continue;
            atomicOr(pend + to[r], 1u << j);
            const uint64_t key = (uint64_t)to[r] << 32;
            unsigned long long dd = srv_mix64(key | nd);
            if (old == OG_UNREACHED) {
                atomicAdd(&s_acc[wv][kSs + j], 1ull);
                atomicAdd(&s_acc[wv][2 * kSs + j], (unsigned long long)(ptr[to[r] + 1] - ptr[to[r]]));
            } else {
                dd -= srv_mix64(key | old);
            }
            atomicAdd(&s_acc[wv][j], dd);
        }
    }
    __syncthreads();
    if (threadIdx.x < 3 * kSs) {
        unsigned long long t = 0;
        for (uint32_t k = 0; k < kBlock / kWave; ++k) t += s_acc[k][threadIdx.x];
        if (t) atomicAdd(acc + threadIdx.x, t);
    }
}

// Resets the distance columns of the slots in `dirty` to "unreached".
__global__ __launch_bounds__(kBlock) void k_ss_clear(uint32_t *dist, uint64_t V, uint32_t dirty) {
    for (uint64_t v = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v < V; v += (uint64_t)gridDim.x * kBlock) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t qb = (dirty >> (4 * q)) & 0xFu;
            if (!qb) continue;
            uint4 *p = (uint4 *)(dist + v * kSs + 4 * q);
            uint4 x = qb == 0xFu ? make_uint4(0, 0, 0, 0) : *p;
            if (qb & 1u) x.x = OG_UNREACHED;
            if (qb & 2u) x.y = OG_UNREACHED;
            if (qb & 4u) x.z = OG_UNREACHED;
            if (qb & 8u) x.w = OG_UNREACHED;
            *p = x;
        }
    }
}

// Column j of the distance lines (a finished query's result array).
__global__ void k_ss_extract(const uint32_t *__restrict__ dist, uint64_t V, uint32_t j, uint32_t *out) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x)
        out[v] = dist[v * kSs + j];
}

__global__ void k_ss_zero_acc(uint32_t done, unsigned long long *acc) {
    const uint32_t j = threadIdx.x;
    if (j < kSs && ((done >> j) & 1u)) acc[j] = acc[kSs + j] = acc[2 * kSs + j] = 0;
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
    int keep = -1;  // BFS: index of the kept-levels buffer the commits write, or -1
};

constexpr uint32_t kSlots = 64;     // BFS query slots per graph (bits of a u64 mask)
constexpr uint32_t kKeepBufs = 4;   // BFS: kept-levels buffers per graph (4 V bytes each)

// Device bytes of per-query state the admission rule charges: the BFS
// engine's masks and kept-levels buffers, and the SSSP engine's distance
// lines (4 kSs V), light and heavy pending masks, active list with masks and
// work prefix,
// the extraction buffer of kept results and the relax's chunk table.
static uint64_t bfs_state_bytes(uint64_t V) { return 24 * V + kKeepBufs * 4 * V; }
static uint64_t sssp_state_bytes(uint64_t V, uint64_t E) {
    return 4ull * kSs * V + 8 * V + 8 * V + 8 * (V + 1) + 4 * V + 4 * (kSs * E / kSsChunk + 2);
}

// What an engine thread marks when it fails: its queue stops taking queries.
struct EngineFailure {
    og_status st = OG_OK;
    std::string msg;
};

struct BfsEngine {
    og_server *srv = nullptr;
    og_graph *g = nullptr;
    DeviceState *d = nullptr;
    uint32_t gi = 0;
    std::deque<Query> *pending = nullptr;
    EngineFailure *failure = nullptr;
    MemPool mem;
    Frontier f;
    hipStream_t s = nullptr;
    hipEvent_t ev_summary = nullptr;  // the round's alive mask and slot sums are on the host
    uint64_t *visit = nullptr, *front = nullptr, *next = nullptr;
    // round summary: [0] alive slots (the commit ORs into it), [1] pulling
    // slots, [2] the round's alive slots, [3, 195) slot sums (acc); the host
    // copy h_sum receives [1, 195) each round
    unsigned long long *d_sum = nullptr, *h_sum = nullptr;
    unsigned long long *d_alive = nullptr, *d_acc = nullptr, *h_acc = nullptr;
    // direction-optimising rounds (graphs prepared with their in-edges)
    unsigned long long *d_fstat = nullptr;     // [2][64] frontier vertices / edges per slot
    unsigned long long *d_explored = nullptr;  // [64] edges of the slot's past frontiers
    unsigned long long *d_pull = nullptr;      // slots relaxed bottom-up this round
    std::atomic<uint64_t> pull_rounds{0};
    uint32_t *keep_buf[kKeepBufs] = {};        // allocated with the engine, charged to the state budget
    uint32_t keep_used = 0;                    // bit k: keep_buf[k] belongs to a slot in flight
    SlotState slots[64];
    uint64_t used = 0;
    uint64_t dirty = 0;  // finished slots whose visit bits are not cleared yet
    uint32_t round = 0;
    std::atomic<uint64_t> rounds_run{0};
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
    void fail_held(og_status st, const std::string &msg);
    KeepList keep_list(uint64_t mask) const {
        KeepList kl{};
        for (uint32_t j = 0; j < kSlots; ++j)
            if (((mask >> j) & 1) && slots[j].keep >= 0) kl.p[j] = keep_buf[slots[j].keep];
        return kl;
    }
};

// Batched SSSP: kSs query slots over the graph's device-resident edges.
struct SsspEngine {
    og_server *srv = nullptr;
    og_graph *g = nullptr;
    DeviceState *d = nullptr;
    uint32_t gi = 0;
    std::deque<Query> *pending = nullptr;
    EngineFailure *failure = nullptr;
    MemPool mem;
    hipStream_t s = nullptr;
    hipEvent_t ev_summary = nullptr;
    uint32_t *dist = nullptr, *pend = nullptr, *hpend = nullptr, *alist = nullptr, *amask = nullptr;
    uint32_t *extract = nullptr, *cstart = nullptr;
    uint64_t *doff = nullptr;
    uint32_t *bnv = nullptr, *bmask = nullptr, *bmin = nullptr;
    uint64_t *bdeg = nullptr;
    uint32_t nblk = 0;
    SsCtl *ctl = nullptr, *h_ctl = nullptr;
    unsigned long long *acc = nullptr, *h_acc = nullptr;
    SlotState slots[kSs];
    uint32_t used = 0, dirty = 0;
    uint32_t round = 0;
    std::atomic<uint64_t> rounds_run{0};
    std::thread th;
    ~SsspEngine() {
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
        if (ev_summary) (void)hipEventDestroy(ev_summary);
    }
    void setup();
    void loop();
    void step();
    void finalize(uint32_t done);
    void fail_held(og_status st, const std::string &msg);
    bool may_admit() const;
};

constexpr uint32_t kSsLanes = 2;           // default single-source SSSP lanes per weighted graph
constexpr uint32_t kSsLanesMax = 8;        // og_server_params.sssp_lanes at most
constexpr uint32_t kSsBatchMin = kSs / 2;  // waiting SSSP queries (every lane busy) that start a batch

// A single-source SSSP lane: takes one query at a time from its graph's SSSP
// queue and runs og_sssp's traversal on its own DeviceState (values,
// frontier, streams, captured rounds) over the shared device-resident graph.
struct SsspLane {
    og_server *srv = nullptr;
    og_graph *g = nullptr;
    DeviceState *lane = nullptr;
    uint32_t gi = 0;
    std::deque<Query> *pending = nullptr;
    EngineFailure *failure = nullptr;
    bool busy = false;  // under srv->mu; stays true once the lane has failed
    std::atomic<uint64_t> rounds_run{0};
    std::thread th;
    ~SsspLane() { device_state_free(lane); }
    void loop();
    void fail_lane(const Query &q, og_status st, const std::string &msg);
};

}  // namespace og

struct og_server {
    std::vector<og_graph *> graphs;
    std::vector<std::deque<og::Query>> queues;          // [graph * 2 + kind]
    std::vector<og::EngineFailure> failures;            // [graph * 2 + kind]
    std::vector<std::unique_ptr<og::BfsEngine>> bfs;    // [graph]
    std::vector<std::unique_ptr<og::SsspEngine>> sssp;  // [graph], null when unweighted
    std::vector<std::vector<std::unique_ptr<og::SsspLane>>> lanes;  // [graph], empty when unweighted
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    uint64_t next_id = 1;
    std::deque<og_query_result> done;
    std::unordered_map<uint64_t, std::vector<uint32_t>> kept;
    og_server_params params{};
    uint64_t submitted = 0, completed = 0, rejected = 0, pending = 0, in_flight = 0;
    uint32_t lanes_running = 0;  // SSSP lanes running a query (og_server_params.sssp_running_max)
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

// Queued queries whose deadline passed before a slot took them complete with
// OG_ETIMEDOUT (the queue is in submit order, so the expired ones lead it).
// Caller holds srv->mu.
static void expire_queued(og_server *srv, uint32_t gi, std::deque<Query> *pending, double t) {
    const uint32_t dl = srv->params.deadline_ms;
    if (!dl) return;
    bool any = false;
    while (!pending->empty() && t - pending->front().t_submit > (double)dl) {
        push_failure(srv, pending->front(), gi, OG_ETIMEDOUT, t);
        pending->pop_front();
        srv->pending--;
        any = true;
    }
    if (any) srv->cv_done.notify_all();
}

// Fails the queries an engine holds (in slots and queued) and marks its queue
// failed, so that later submits for it return the error instead of waiting.
// Caller holds srv->mu.
template <class Slots>
static void fail_engine(og_server *srv, uint32_t gi, std::deque<Query> *pending, EngineFailure *failure,
                        Slots &slots, uint32_t nslots, og_status st, const std::string &msg) {
    const double t = now_ms(srv);
    failure->st = st == OG_OK ? OG_EHIP : st;
    failure->msg = msg;
    for (uint32_t j = 0; j < nslots; ++j)
        if (slots[j].used) {
            push_failure(srv, slots[j].q, gi, failure->st, t);
            slots[j].used = false;
            srv->in_flight--;
        }
    for (auto &q : *pending) {
        push_failure(srv, q, gi, failure->st, t);
        srv->pending--;
    }
    pending->clear();
    srv->cv_done.notify_all();
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
        if (slots[j].keep >= 0) keep_used &= ~(1u << slots[j].keep);
        slots[j].keep = -1;
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
        expire_queued(srv, gi, pending, t);
        const uint32_t clean = kSlots - (uint32_t)__builtin_popcountll(used | dirty);
        if (dirty && (pending->size() > clean || __builtin_popcountll(dirty) >= (int)kSlots / 2)) {
            hipLaunchKernelGGL(k_srv_clear_visit, dim3(sweep_grid(d->V, 16384)), dim3(kBlock), 0, s, d->V, dirty,
                               visit);
            OG_HIP(hipGetLastError());
            dirty = 0;  // stream order: the sweep runs before this round's seeds and commit
        }
        while (!pending->empty() && (uint32_t)__builtin_popcountll(used | dirty) < kSlots) {
            // a query keeping its levels needs one of the kKeepBufs buffers;
            // admission stays in arrival order, so it waits for one
            int kb = -1;
            if (pending->front().flags & OG_QUERY_KEEP_VALUES) {
                if (keep_used == (1u << kKeepBufs) - 1) break;
                kb = __builtin_ctz(~keep_used);
                keep_used |= 1u << kb;
            }
            uint32_t j = 0;
            while (((used | dirty) >> j) & 1) ++j;
            slots[j].used = true;
            slots[j].q = pending->front();
            slots[j].start_round = round;
            slots[j].t_admit = t;
            slots[j].keep = kb;
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
        const int kb = slots[sl.slot[i]].keep;
        if (kb >= 0) OG_HIP(hipMemsetAsync(keep_buf[kb], 0xFF, V * 4, s));
    }
    if (sl.n) {
        hipLaunchKernelGGL(k_srv_seed, dim3(1), dim3(64), 0, s, sl, next, f.flag, d_explored, d_acc);
        OG_HIP(hipGetLastError());
    }
    const bool dir_opt = d->d_in_ptr != nullptr;
    SlotLevels L{};
    for (uint32_t j = 0; j < kSlots; ++j) L.lv[j] = round - slots[j].start_round + 1;
    hipLaunchKernelGGL(k_srv_commit, dim3(sweep_grid(V, 4096)), dim3(kBlock), 0, s, f.flag, V,
                       d->d_ptr, d->d_deg, d->deg_heavy(), visit, front, next, L, keep_list(used), d_acc, d_fstat,
                       d_alive);
    OG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_srv_direction, dim3(1), dim3(64), 0, s, used, V, d->E, dir_opt ? 1 : 0, d_fstat,
                       d_explored, d_sum);
    // pull mask, alive mask and the slot sums (results of the slots that
    // finish this round are complete after the commit) in one copy
    OG_HIP(hipMemcpyAsync(h_sum + 1, d_sum + 1, (2 + 3 * 64) * 8, hipMemcpyDeviceToHost, s));
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
    rounds_run.fetch_add(1, std::memory_order_relaxed);
    if (h_sum[1]) pull_rounds.fetch_add(1, std::memory_order_relaxed);
    // a slot is done when the round's commit left it no frontier vertex
    const uint64_t doneset = used & ~(uint64_t)h_sum[2];
    if (doneset) finalize(doneset);
}

void BfsEngine::fail_held(og_status st, const std::string &msg) {
    std::lock_guard<std::mutex> lk(srv->mu);
    fail_engine(srv, gi, pending, failure, slots, kSlots, st, msg);
    used = 0;
    keep_used = 0;
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
        // this graph's BFS queue fails (held and later queries get the error);
        // the other graphs' engines keep serving
        fail_held(e.st, e.msg);
    } catch (...) {
        fail_held(OG_ENOMEM_HOST, "BFS engine stopped: host allocation failed");
    }
}

void SsspEngine::setup() {
    const uint64_t V = d->V;
    nblk = (uint32_t)((V + kSsWin - 1) / kSsWin);
    OG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    OG_HIP(hipEventCreateWithFlags(&ev_summary, hipEventDisableTiming));
    dist = (uint32_t *)mem.dalloc(V * kSs * 4);
    pend = (uint32_t *)mem.dalloc(V * 4);
    hpend = (uint32_t *)mem.dalloc(V * 4);
    alist = (uint32_t *)mem.dalloc(V * 4);
    amask = (uint32_t *)mem.dalloc(V * 4);
    doff = (uint64_t *)mem.dalloc((V + 1) * 8);
    extract = (uint32_t *)mem.dalloc(V * 4);
    cstart = (uint32_t *)mem.dalloc((kSs * d->E / kSsChunk + 2) * 4);  // items <= kSs per edge
    bnv = (uint32_t *)mem.dalloc((uint64_t)nblk * 4);
    bdeg = (uint64_t *)mem.dalloc((uint64_t)nblk * 8);
    bmask = (uint32_t *)mem.dalloc(3ull * nblk * 4);  // per block: near, far, heavy-pending slot masks
    bmin = (uint32_t *)mem.dalloc((uint64_t)kSs * nblk * 4);
    ctl = (SsCtl *)mem.dalloc(sizeof(SsCtl));
    h_ctl = (SsCtl *)mem.palloc(sizeof(SsCtl));
    acc = (unsigned long long *)mem.dalloc(3 * kSs * 8);
    h_acc = (unsigned long long *)mem.palloc(3 * kSs * 8);
    OG_HIP(hipMemset(dist, 0xFF, V * kSs * 4));
    OG_HIP(hipMemset(pend, 0, V * 4));
    OG_HIP(hipMemset(hpend, 0, V * 4));
    OG_HIP(hipMemset(acc, 0, 3 * kSs * 8));
    SsCtl c{};
    c.delta = srv->params.sssp_delta ? srv->params.sssp_delta : d->tune.delta;
    for (uint32_t j = 0; j < kSs; ++j) c.thresh[j] = 0xFFFFFFFFu;
    OG_HIP(hipMemcpy(ctl, &c, sizeof c, hipMemcpyHostToDevice));
}

// Finished slots report the sums copied with the round's summary; kept
// results are read from the distance column before it is reset.  The slots
// turn dirty: their columns are reset by one sweep for many slots, when a
// queued query needs a slot or half of them are dirty.
void SsspEngine::finalize(uint32_t done) {
    const uint64_t V = d->V;
    std::vector<std::pair<uint64_t, std::vector<uint32_t>>> vals;
    for (uint32_t j = 0; j < kSs; ++j) {
        if (!((done >> j) & 1u) || !(slots[j].q.flags & OG_QUERY_KEEP_VALUES)) continue;
        hipLaunchKernelGGL(k_ss_extract, dim3(sweep_grid(V, 8192)), dim3(kBlock), 0, s, dist, V, j, extract);
        OG_HIP(hipGetLastError());
        vals.push_back({slots[j].q.id, std::vector<uint32_t>(V)});
        OG_HIP(hipMemcpyAsync(vals.back().second.data(), extract, V * 4, hipMemcpyDeviceToHost, s));
        OG_HIP(hipStreamSynchronize(s));  // the extraction buffer is reused
    }
    hipLaunchKernelGGL(k_ss_zero_acc, dim3(1), dim3(64), 0, s, done, acc);
    OG_HIP(hipGetLastError());
    dirty |= done;
    const double t = now_ms(srv);
    std::lock_guard<std::mutex> lk(srv->mu);
    for (uint32_t j = 0; j < kSs; ++j) {
        if (!((done >> j) & 1u)) continue;
        og_query_result r{};
        r.id = slots[j].q.id;
        r.graph = gi;
        r.kind = OG_QUERY_SSSP;
        r.source = slots[j].q.source;
        r.status = OG_OK;
        r.checksum = h_acc[j];
        r.reached = h_acc[kSs + j];
        r.traversed_edges = h_acc[2 * kSs + j];
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

// One round: admit queued queries into free slots (resetting dirty columns
// first when needed), then select / scan / write / relax on the engine's
// stream; the host waits for the scan's summary only (it sizes the relax),
// so the round's write and relax run while it reports the slots that
// finished.
void SsspEngine::step() {
    if (srv->params.flags & OG_SERVER_TEST_FAIL_SSSP) {
        std::unique_lock<std::mutex> lk(srv->mu);
        srv->cv_work.wait(lk, [&] { return srv->stop || !pending->empty(); });
        if (srv->stop) return;
        lk.unlock();
        raise(OG_EHIP, "forced failure (OG_SERVER_TEST_FAIL_SSSP)");
    }
    SsSeeds sl{};
    {
        std::unique_lock<std::mutex> lk(srv->mu);
        srv->cv_work.wait(lk, [&] { return srv->stop || used || may_admit(); });
        if (srv->stop) return;
        const double t = now_ms(srv);
        expire_queued(srv, gi, pending, t);
        const uint32_t clean = kSs - (uint32_t)__builtin_popcount(used | dirty);
        if (dirty && may_admit() && (pending->size() > clean || __builtin_popcount(dirty) >= (int)kSs / 2)) {
            hipLaunchKernelGGL(k_ss_clear, dim3(sweep_grid(d->V, 16384)), dim3(kBlock), 0, s, dist, d->V, dirty);
            OG_HIP(hipGetLastError());
            dirty = 0;  // stream order: the reset runs before this round's seeds
        }
        while (may_admit() && (uint32_t)__builtin_popcount(used | dirty) < kSs) {
            uint32_t j = 0;
            while (((used | dirty) >> j) & 1u) ++j;
            slots[j].used = true;
            slots[j].q = pending->front();
            slots[j].start_round = round;
            slots[j].t_admit = t;
            pending->pop_front();
            used |= 1u << j;
            sl.slot[sl.n] = j;
            sl.src[sl.n] = slots[j].q.source;
            sl.n++;
            srv->pending--;
            srv->in_flight++;
        }
    }
    if (!used) return;
    const uint64_t V = d->V;
    const uint64_t *pmid = d->d_pmid;  // rows stored light edges first: light / heavy rounds
    if (sl.n) hipLaunchKernelGGL(k_ss_seed, dim3(1), dim3(64), 0, s, sl, dist, pend, d->d_ptr, acc, ctl);
    hipLaunchKernelGGL(k_ss_select, dim3(nblk), dim3(kBlock), 0, s, pend, hpend, dist, d->d_ptr, pmid, V, ctl, bnv,
                       bdeg, bmask, bmin, nblk);
    hipLaunchKernelGGL(k_ss_scan, dim3(1), dim3(1024), 0, s, bnv, bdeg, bmask, bmin, nblk, used, ctl, doff);
    OG_HIP(hipGetLastError());
    OG_HIP(hipMemcpyAsync(h_ctl, ctl, sizeof(SsCtl), hipMemcpyDeviceToHost, s));
    OG_HIP(hipMemcpyAsync(h_acc, acc, 3 * kSs * 8, hipMemcpyDeviceToHost, s));
    OG_HIP(hipEventRecord(ev_summary, s));
    hipLaunchKernelGGL(k_ss_write, dim3(nblk), dim3(kBlock), 0, s, pend, hpend, d->d_ptr, pmid, V, bnv, bdeg, alist,
                       amask, doff, cstart);
    OG_HIP(hipGetLastError());
    // The relax is sized from the summary: one chunk per workgroup up to
    // kSsRelaxGridCap workgroups (a small round launches only what it needs,
    // so the other engines' kernels get CUs beside it), striding beyond.
    OG_HIP(hipEventSynchronize(ev_summary));
    const uint64_t nch = (h_ctl->dsum + kSsChunk - 1) / kSsChunk;
    if (nch) {
        const dim3 grid((unsigned)std::min<uint64_t>(nch, kSsRelaxGridCap));
        if (d->packed)
            hipLaunchKernelGGL((k_ss_relax<true, true>), grid, dim3(kBlock), 0, s, alist, amask, doff, cstart, ctl,
                               d->d_ptr, pmid, (const void *)d->d_static, d->wbits, dist, pend, acc);
        else
            hipLaunchKernelGGL((k_ss_relax<true, false>), grid, dim3(kBlock), 0, s, alist, amask, doff, cstart, ctl,
                               d->d_ptr, pmid, (const void *)d->d_static, d->wbits, dist, pend, acc);
        OG_HIP(hipGetLastError());
    }
    ++round;
    rounds_run.fetch_add(1, std::memory_order_relaxed);
    // a slot is done when the select sweep found nothing pending for it; its
    // sums (copied with the summary) are final: this round relaxes none of it
    const uint32_t doneset = used & h_ctl->done;
    if (doneset) finalize(doneset);
    if (!used && dirty) {  // idle: reset the finished columns now, off the next query's path
        hipLaunchKernelGGL(k_ss_clear, dim3(sweep_grid(V, 16384)), dim3(kBlock), 0, s, dist, V, dirty);
        OG_HIP(hipGetLastError());
        dirty = 0;
    }
}

// Lanes first: with lanes, queued queries are admitted only while no lane
// can take them (every lane busy, or the running cap reached) and the engine
// already runs queries or kSsBatchMin wait (a lone query is answered sooner
// by a single-source lane).  Caller holds srv->mu.
bool SsspEngine::may_admit() const {
    if (pending->empty()) return false;
    const auto &ls = srv->lanes[gi];
    const uint32_t cap = srv->params.sssp_running_max;
    if (!(cap && srv->lanes_running >= cap))
        for (const auto &l : ls)
            if (!l->busy) return false;
    return ls.empty() || used != 0 || pending->size() >= kSsBatchMin;
}

void SsspEngine::fail_held(og_status st, const std::string &msg) {
    std::lock_guard<std::mutex> lk(srv->mu);
    fail_engine(srv, gi, pending, failure, slots, kSs, st, msg);
    used = 0;
}

void SsspEngine::loop() {
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
        fail_held(e.st, e.msg);
    } catch (...) {
        fail_held(OG_ENOMEM_HOST, "SSSP engine stopped: host allocation failed");
    }
}

void SsspLane::loop() {
    Query q{};
    try {
        OG_HIP(hipSetDevice(lane->device));
        for (;;) {
            double t_admit = 0;
            {
                std::unique_lock<std::mutex> lk(srv->mu);
                const uint32_t cap = srv->params.sssp_running_max;
                srv->cv_work.wait(lk, [&] {
                    return srv->stop || (!pending->empty() && (!cap || srv->lanes_running < cap));
                });
                if (srv->stop) return;
                t_admit = now_ms(srv);
                expire_queued(srv, gi, pending, t_admit);
                if (pending->empty()) continue;
                q = pending->front();
                pending->pop_front();
                srv->pending--;
                srv->in_flight++;
                srv->lanes_running++;
                busy = true;
            }
            srv->cv_work.notify_all();  // with every lane busy the batched engine may admit the rest
            if (srv->params.flags & OG_SERVER_TEST_FAIL_SSSP)
                raise(OG_EHIP, "forced failure (OG_SERVER_TEST_FAIL_SSSP)");
            std::vector<uint32_t> vals;
            if (q.flags & OG_QUERY_KEEP_VALUES) vals.resize(lane->V);
            og_run_stats st{};
            lane_run_sssp(g, lane, q.source, vals.empty() ? nullptr : vals.data(), &st);
            rounds_run.fetch_add(st.iterations, std::memory_order_relaxed);
            const double t = now_ms(srv);
            std::lock_guard<std::mutex> lk(srv->mu);
            og_query_result r{};
            r.id = q.id;
            r.graph = gi;
            r.kind = OG_QUERY_SSSP;
            r.source = q.source;
            r.status = OG_OK;
            r.checksum = st.checksum;
            r.reached = st.reached;
            r.traversed_edges = st.traversed_edges;
            r.rounds = st.iterations;
            r.path = OG_PATH_LANE;
            r.queue_ms = t_admit - q.t_submit;
            r.latency_ms = t - q.t_submit;
            srv->done.push_back(r);
            if (!vals.empty()) srv->kept[q.id] = std::move(vals);
            srv->completed++;
            srv->in_flight--;
            srv->lanes_running--;
            busy = false;
            srv->cv_done.notify_all();
            srv->cv_work.notify_all();  // a lane of another graph may wait for the running cap
        }
    } catch (const Error &e) {
        fail_lane(q, e.st, e.msg);
    } catch (...) {
        fail_lane(q, OG_ENOMEM_HOST, "SSSP lane stopped: host allocation failed");
    }
    srv->cv_work.notify_all();
}

// The query in hand and the graph's SSSP queue fail, as when the batched
// engine stops; the lane stays busy, so it takes no more, and gives back its
// place under the running cap.
void SsspLane::fail_lane(const Query &q, og_status st, const std::string &msg) {
    std::lock_guard<std::mutex> lk(srv->mu);
    SlotState held[1];
    held[0].used = busy;
    held[0].q = q;
    if (busy) srv->lanes_running--;
    fail_engine(srv, gi, pending, failure, held, 1, st, msg);
    busy = true;
}

static void server_free(og_server *srv) {
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
    }
    srv->cv_work.notify_all();
    for (auto &e : srv->bfs)
        if (e && e->th.joinable()) e->th.join();
    for (auto &e : srv->sssp)
        if (e && e->th.joinable()) e->th.join();
    for (auto &ls : srv->lanes)
        for (auto &l : ls)
            if (l->th.joinable()) l->th.join();
    srv->lanes.clear();
    srv->bfs.clear();
    srv->sssp.clear();
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
        if (g->E >> kFcShift)  // the commit's per-workgroup degree sums
            return fail(OG_EUNSUPPORTED, "graph %u has 2^%d edges or more", i, kFcShift);
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
        // Admission: every graph's BFS engine state and, when batching is
        // asked for, every weighted graph's batched SSSP engine state must
        // fit the budget.
        const uint32_t fl = srv->params.flags;
        const bool batch = (fl & (OG_SERVER_SSSP_BATCH | OG_SERVER_NO_SSSP_LANES)) != 0;
        const bool lanes_on = !(fl & OG_SERVER_NO_SSSP_LANES);
        uint64_t need = 0;
        for (og_graph *g : srv->graphs) {
            need += bfs_state_bytes(g->V);
            if (g->kind == OG_KIND_WEIGHTED && batch) need += sssp_state_bytes(g->V, g->E);
        }
        uint64_t budget = srv->params.state_budget_bytes;
        if (!budget) {
            size_t fr = 0, tot = 0;
            OG_HIP(hipMemGetInfo(&fr, &tot));
            budget = fr / 4;
        }
        if (need > budget)
            raise(OG_EBUDGET, "state budget %llu bytes cannot hold the engines' query state (%llu bytes)",
                  (unsigned long long)budget, (unsigned long long)need);
        // single-source SSSP lanes per weighted graph, as many (up to
        // sssp_lanes) as the rest of the budget holds for every weighted graph
        uint64_t per_lane_set = 0;
        for (og_graph *g : srv->graphs)
            if (g->kind == OG_KIND_WEIGHTED) per_lane_set += lane_state_bytes(g);
        const uint32_t want = srv->params.sssp_lanes ? srv->params.sssp_lanes : kSsLanes;
        if (want > kSsLanesMax) raise(OG_EINVAL, "sssp_lanes %u > %u", want, kSsLanesMax);
        uint32_t nlanes = lanes_on && per_lane_set ? want : 0;
        while (nlanes && need + nlanes * per_lane_set > budget) --nlanes;
        if (lanes_on && per_lane_set && !nlanes && !batch)
            raise(OG_EBUDGET, "state budget %llu bytes holds no SSSP lane (%llu bytes of BFS state + %llu per lane set)",
                  (unsigned long long)budget, (unsigned long long)need, (unsigned long long)per_lane_set);
        srv->queues.resize(2 * num_graphs);
        srv->failures.resize(2 * num_graphs);
        srv->bfs.resize(num_graphs);
        srv->sssp.resize(num_graphs);
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
            e->failure = &srv->failures[2 * gi];
            OG_HIP(hipStreamCreateWithFlags(&e->s, hipStreamNonBlocking));
            OG_HIP(hipEventCreateWithFlags(&e->ev_summary, hipEventDisableTiming));
            frontier_alloc(g->dev, g, e->f, e->mem);
            e->visit = (uint64_t *)e->mem.dalloc(V * 8);
            e->front = (uint64_t *)e->mem.dalloc(V * 8);
            e->next = (uint64_t *)e->mem.dalloc(V * 8);
            for (auto &kb : e->keep_buf) kb = (uint32_t *)e->mem.dalloc(V * 4);
            OG_HIP(hipMemset(e->visit, 0, V * 8));
            OG_HIP(hipMemset(e->front, 0, V * 8));
            OG_HIP(hipMemset(e->next, 0, V * 8));
            e->d_sum = (unsigned long long *)e->mem.dalloc((3 + 3 * 64) * 8);
            e->h_sum = (unsigned long long *)e->mem.palloc((3 + 3 * 64) * 8);
            OG_HIP(hipMemset(e->d_sum, 0, (3 + 3 * 64) * 8));
            e->d_alive = e->d_sum;
            e->d_pull = e->d_sum + 1;
            e->d_acc = e->d_sum + 3;
            e->h_acc = e->h_sum + 3;
            e->d_fstat = (unsigned long long *)e->mem.dalloc(2 * 64 * 8);
            e->d_explored = (unsigned long long *)e->mem.dalloc(64 * 8);
            OG_HIP(hipMemset(e->d_fstat, 0, 2 * 64 * 8));
            OG_HIP(hipMemset(e->d_explored, 0, 64 * 8));
            srv->state_bytes += bfs_state_bytes(V);
            srv->bfs[gi] = std::move(e);
            if (g->kind != OG_KIND_WEIGHTED) continue;
            if (batch) {
                std::unique_ptr<SsspEngine> x(new SsspEngine);
                x->srv = srv;
                x->g = g;
                x->d = g->dev;
                x->gi = gi;
                x->pending = &srv->queues[2 * gi + 1];
                x->failure = &srv->failures[2 * gi + 1];
                x->setup();
                srv->state_bytes += sssp_state_bytes(V, g->E);
                srv->sssp[gi] = std::move(x);
            }
            for (uint32_t k = 0; k < nlanes; ++k) {
                std::unique_ptr<SsspLane> l(new SsspLane);
                l->srv = srv;
                l->g = g;
                l->gi = gi;
                l->pending = &srv->queues[2 * gi + 1];
                l->failure = &srv->failures[2 * gi + 1];
                l->lane = lane_state_create(g, srv->params.sssp_delta);
                srv->state_bytes += l->lane->device_bytes();
                srv->lanes[gi].push_back(std::move(l));
            }
        }
        OG_HIP(hipDeviceSynchronize());
        for (auto &e : srv->bfs) e->th = std::thread(&BfsEngine::loop, e.get());
        for (auto &x : srv->sssp)
            if (x) x->th = std::thread(&SsspEngine::loop, x.get());
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
    if (kind == OG_QUERY_SSSP && !srv->sssp[graph] && srv->lanes[graph].empty())
        return fail(OG_EUNSUPPORTED, "graph %u has no weights: SSSP queries are not served", graph);
    if (source >= srv->graphs[graph]->V) return fail(OG_EINVAL, "source %u >= V", source);
    std::lock_guard<std::mutex> lk(srv->mu);
    if (srv->stop) return fail(OG_ESTATE, "server is stopping");
    const EngineFailure &ef = srv->failures[2 * graph + kind];
    if (ef.st != OG_OK)
        return fail(ef.st, "the %s engine of graph %u has stopped: %s", kind ? "SSSP" : "BFS", graph, ef.msg.c_str());
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
        info->sssp_slots[gi] = (srv->sssp[gi] ? kSs : 0) + (uint32_t)srv->lanes[gi].size();
        info->rounds += srv->bfs[gi]->rounds_run.load(std::memory_order_relaxed);
        if (srv->sssp[gi]) info->rounds += srv->sssp[gi]->rounds_run.load(std::memory_order_relaxed);
        for (auto &l : srv->lanes[gi]) info->rounds += l->rounds_run.load(std::memory_order_relaxed);
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
