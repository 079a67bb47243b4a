// Query server: continuous batching of BFS and SSSP queries on device-
// resident graphs (no reference counterpart; the reference runs one query per
// process, BFS/oneGraph_BFS/main.cpp:3-14).
//
// Every (graph, kind) pair has one batch engine with S <= 64 query slots and
// its own HIP stream, frontier buffers and host thread.  A slot is one bit of
// the per-vertex u64 masks (front / next, plus visit for BFS) and one column
// of the per-vertex slot state (u8 levels for BFS, u32 distances for SSSP,
// vertex-major [V][S]).  Each engine round:
//   1. admit pending queries into free slots (seed: source bit, level/dist 1);
//   2. compact the frontier flags into degree bins (shared with og_bfs);
//   3. relax: BFS through the multi-source bitmask kernels; SSSP with one
//      wave per frontier vertex (or hub tile quarter), lanes = slots, so a
//      vertex's S distances are one coalesced row;
//   4. commit: next -> front for flagged vertices, stale front masks cleared,
//      and the OR of the new frontier masks gives the slots still alive;
//   5. finalise the slots that emptied: per-slot digest, reached count and
//      traversed edges (and the values when asked for), reset their column.
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

struct SeedList {
    uint32_t n;
    uint32_t slot[64];
    uint32_t src[64];
};

struct LevelList {
    uint8_t lv[64];
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

__global__ void k_srv_seed(SeedList sl, uint32_t S, int kind, uint64_t *visit, uint64_t *front, uint8_t *flag,
                           uint8_t *lvl, uint32_t *dist) {
    const uint32_t i = threadIdx.x;
    if (i < sl.n) {
        const uint32_t j = sl.slot[i], s = sl.src[i];
        const unsigned long long bit = 1ull << j;
        atomicOr((unsigned long long *)(front + s), bit);
        flag[s] = 1;
        if (kind == 0) {
            atomicOr((unsigned long long *)(visit + s), bit);
            lvl[(uint64_t)s * S + j] = 1;
        } else {
            dist[(uint64_t)s * S + j] = 1;
        }
    }
}

// BFS commit over all vertices: flagged ones take their unseen offered
// slots as the new frontier mask (and level); the others drop stale masks.
__global__ __launch_bounds__(kBlock) void k_srv_bfs_commit(uint8_t *flag, uint64_t V, uint32_t S, uint64_t *visit,
                                                           uint64_t *front, uint64_t *next, uint8_t *lvl, LevelList L,
                                                           unsigned long long *alive) {
    uint64_t al = 0;
    for (uint64_t v = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v < V; v += (uint64_t)gridDim.x * kBlock) {
        if (flag[v]) {
            const uint64_t vis = visit[v];
            const uint64_t nw = next[v] & ~vis;
            next[v] = 0;
            front[v] = nw;
            if (nw) {
                visit[v] = vis | nw;
                uint64_t m = nw;
                while (m) {
                    const int j = __builtin_ctzll(m);
                    m &= m - 1;
                    lvl[v * S + j] = L.lv[j];
                }
                al |= nw;
            } else {
                flag[v] = 0;
            }
        } else if (front[v]) {
            front[v] = 0;
        }
    }
    al = wave_or(al);
    if ((threadIdx.x & (kWave - 1)) == 0 && al) atomicOr(alive, (unsigned long long)al);
}

__global__ __launch_bounds__(kBlock) void k_srv_sssp_commit(uint8_t *flag, uint64_t V, uint64_t *front,
                                                            uint64_t *next, unsigned long long *alive) {
    uint64_t al = 0;
    for (uint64_t v = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v < V; v += (uint64_t)gridDim.x * kBlock) {
        if (flag[v]) {
            const uint64_t nw = next[v];
            next[v] = 0;
            front[v] = nw;
            al |= nw;
            if (!nw) flag[v] = 0;
        } else if (front[v]) {
            front[v] = 0;
        }
    }
    al = wave_or(al);
    if ((threadIdx.x & (kWave - 1)) == 0 && al) atomicOr(alive, (unsigned long long)al);
}

// Multi-slot SSSP relax.  Wave w < n01 takes list vertex w; the rest take a
// quarter of a hub tile.  Lane = (edge group g, slot j) with S lanes per
// group and 64/S groups per wave, so each round of the loop relaxes 64/S
// edges for all S slots with coalesced rows of dist.
template <bool PK>
__global__ __launch_bounds__(kBlock) void k_srv_sssp_relax(const uint32_t *__restrict__ list, uint32_t n01,
                                                           const uint64_t *__restrict__ hub, uint32_t nhub_items,
                                                           uint32_t tile, const uint64_t *__restrict__ ptr,
                                                           const void *__restrict__ edges, uint32_t wbits, uint32_t S,
                                                           const uint64_t *__restrict__ front, uint32_t *dist,
                                                           uint64_t *next, uint8_t *flag) {
    const uint64_t wave = ((uint64_t)blockIdx.x * kBlock + threadIdx.x) / kWave;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t G = kWave / S, grp = lane / S, j = lane % S;
    uint32_t u;
    uint64_t rb, re;
    if (wave < n01) {
        u = list[wave];
        rb = ptr[u];
        re = ptr[u + 1];
    } else {
        const uint64_t h = wave - n01;
        if (h >= 4ull * nhub_items) return;  // whole wave exits together
        const uint64_t it = hub[h / 4];
        u = (uint32_t)it;
        const uint64_t tb = ptr[u] + (it >> 32) * tile, te = min(ptr[u + 1], tb + tile);
        const uint64_t q = (te - tb + 3) / 4;
        rb = min(te, tb + (h % 4) * q);
        re = min(te, rb + q);
    }
    const uint64_t A = front[u];
    const bool on = (A >> j) & 1;
    const uint32_t du = on ? dist[(uint64_t)u * S + j] : 0xFFFFFFFFu;
    for (uint64_t base = rb; base < re; base += G) {
        const uint64_t i = base + grp;
        bool upd = false;
        uint32_t to = 0;
        if (i < re) {
            uint32_t wt;
            if (PK) {  // packed word to << wbits | w
                const uint32_t x = ((const uint32_t *)edges)[i];
                to = x >> wbits;
                wt = x & ((1u << wbits) - 1u);
            } else {   // 8-byte {to, w}
                const uint2 e = ((const uint2 *)edges)[i];
                to = e.x;
                wt = e.y;
            }
            if (on && du != 0xFFFFFFFFu) {
                uint32_t nd = du + wt;
                if (nd < du) nd = 0xFFFFFFFEu;
                uint32_t *p = dist + (uint64_t)to * S + j;
                if (nd < *p) upd = nd < atomicMin(p, nd);
            }
        }
        const uint64_t b = __ballot(upd);
        const uint64_t gm = (S == 64) ? b : ((b >> (grp * S)) & ((1ull << S) - 1));
        if (j == 0 && gm) {
            atomicOr((unsigned long long *)(next + to), (unsigned long long)gm);
            flag[to] = 1;
        }
    }
}

// Finalises the slots in `done`: digest / reached / traversed per slot,
// optional copy of the values, reset of the slot column (and visit bits).
__global__ __launch_bounds__(kBlock) void k_srv_finalize(int kind, uint64_t V, uint32_t S, uint64_t done,
                                                         const uint64_t *__restrict__ ptr, uint64_t *visit,
                                                         uint8_t *lvl, uint32_t *dist, KeepList keep,
                                                         unsigned long long *out) {
    __shared__ unsigned long long acc[3 * 64];
    for (int t = threadIdx.x; t < 3 * 64; t += kBlock) acc[t] = 0;
    __syncthreads();
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t G = kWave / S, grp = lane / S, j = lane % S;
    const bool mine = (done >> j) & 1;
    const uint64_t wave = ((uint64_t)blockIdx.x * kBlock + threadIdx.x) / kWave;
    const uint64_t nwaves = (uint64_t)gridDim.x * kBlock / kWave;
    uint64_t h = 0, r = 0, t = 0;
    for (uint64_t v = wave * G + grp; v < V; v += nwaves * G) {
        if (mine) {
            uint32_t val;
            if (kind == 0) {
                const uint32_t x = lvl[v * S + j];
                val = x ? x : OG_UNREACHED;
                if (x) lvl[v * S + j] = 0;
            } else {
                val = dist[v * S + j];
                if (val != OG_UNREACHED) dist[v * S + j] = OG_UNREACHED;
            }
            if (val != OG_UNREACHED) {
                ++r;
                t += ptr[v + 1] - ptr[v];
                h += srv_mix64((v << 32) | val);
            }
            if (keep.p[j]) keep.p[j][v] = val;
        }
        if (kind == 0 && j == 0) {
            const uint64_t vis = visit[v];
            if (vis & done) visit[v] = vis & ~done;
        }
    }
    if (mine) {
        atomicAdd(&acc[j], (unsigned long long)h);
        atomicAdd(&acc[64 + j], (unsigned long long)r);
        atomicAdd(&acc[128 + j], (unsigned long long)t);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 3 * 64; k += kBlock)
        if (acc[k]) atomicAdd(out + k, acc[k]);
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
    uint64_t per_slot_bytes = 0;
    MemPool mem;
    Frontier f;
    hipStream_t s = nullptr;
    uint64_t *visit = nullptr, *front = nullptr, *next = nullptr;
    uint8_t *lvl = nullptr;
    uint32_t *dist = nullptr;
    unsigned long long *d_alive = nullptr, *h_alive = nullptr, *d_fin = nullptr, *h_fin = nullptr;
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

void Engine::finalize(uint64_t done) {
    KeepList kl{};
    for (uint32_t j = 0; j < S; ++j) {
        if (((done >> j) & 1) && (slots[j].q.flags & OG_QUERY_KEEP_VALUES)) {
            if (!keep_buf[j]) keep_buf[j] = (uint32_t *)mem.dalloc(d->V * 4);
            kl.p[j] = keep_buf[j];
        }
    }
    OG_HIP(hipMemsetAsync(d_fin, 0, 3 * 64 * 8, s));
    const unsigned grid = (unsigned)std::min<uint64_t>(2048, (d->V * S + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_srv_finalize, dim3(std::max(grid, 1u)), dim3(kBlock), 0, s, kind, d->V, S, done, d->d_ptr,
                       visit, lvl, dist, kl, d_fin);
    OG_HIP(hipGetLastError());
    OG_HIP(hipMemcpyAsync(h_fin, d_fin, 3 * 64 * 8, hipMemcpyDeviceToHost, s));
    std::vector<std::pair<uint64_t, std::vector<uint32_t>>> vals;
    for (uint32_t j = 0; j < S; ++j) {
        if (kl.p[j]) {
            vals.push_back({slots[j].q.id, std::vector<uint32_t>(d->V)});
            OG_HIP(hipMemcpyAsync(vals.back().second.data(), kl.p[j], d->V * 4, hipMemcpyDeviceToHost, s));
        }
    }
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
        r.checksum = h_fin[j];
        r.reached = h_fin[64 + j];
        r.traversed_edges = h_fin[128 + j];
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
    if (sl.n) {
        hipLaunchKernelGGL(k_srv_seed, dim3(1), dim3(64), 0, s, sl, S, kind, visit, front, f.flag, lvl, dist);
        OG_HIP(hipGetLastError());
    }
    frontier_compact(d, f, s);
    OG_HIP(hipStreamSynchronize(s));
    const FrontierInfo fi = *f.h_info;
    const bool w = g->kind == OG_KIND_WEIGHTED;
    if (fi.n[0] | fi.n[1] | fi.n[2]) {
        if (kind == 0) {
            RelaxArgs a{nullptr, f.flag, visit, front, next, d->wbits};
            if (w) launch_static_relax<ALGO_MSBFS, true>(d, f, &fi, a, s);
            else launch_static_relax<ALGO_MSBFS, false>(d, f, &fi, a, s);
        } else {
            const uint32_t n01 = fi.n[0] + fi.n[1];
            const uint64_t waves = (uint64_t)n01 + 4ull * fi.n[2];
            const unsigned grid = (unsigned)((waves + (kBlock / kWave) - 1) / (kBlock / kWave));
            if (d->packed)
                hipLaunchKernelGGL((k_srv_sssp_relax<true>), dim3(grid), dim3(kBlock), 0, s, f.list, n01, f.hub,
                                   fi.n[2], d->tune.tile, d->d_ptr, d->d_static, d->wbits, S, front, dist, next,
                                   f.flag);
            else
                hipLaunchKernelGGL((k_srv_sssp_relax<false>), dim3(grid), dim3(kBlock), 0, s, f.list, n01, f.hub,
                                   fi.n[2], d->tune.tile, d->d_ptr, d->d_static, d->wbits, S, front, dist, next,
                                   f.flag);
            OG_HIP(hipGetLastError());
        }
    }
    OG_HIP(hipMemsetAsync(d_alive, 0, 8, s));
    const unsigned cgrid = (unsigned)std::min<uint64_t>(4096, (d->V + kBlock - 1) / kBlock);
    if (kind == 0) {
        LevelList L{};
        for (uint32_t j = 0; j < S; ++j) {
            const uint32_t lv = round - slots[j].start_round + 2;
            L.lv[j] = (uint8_t)std::min<uint32_t>(lv, 255);
        }
        hipLaunchKernelGGL(k_srv_bfs_commit, dim3(cgrid), dim3(kBlock), 0, s, f.flag, d->V, S, visit, front, next,
                           lvl, L, d_alive);
    } else {
        hipLaunchKernelGGL(k_srv_sssp_commit, dim3(cgrid), dim3(kBlock), 0, s, f.flag, d->V, front, next, d_alive);
    }
    OG_HIP(hipGetLastError());
    OG_HIP(hipMemcpyAsync(h_alive, d_alive, 8, hipMemcpyDeviceToHost, s));
    OG_HIP(hipStreamSynchronize(s));
    ++round;
    ++rounds_run;
    uint64_t doneset = used & ~(uint64_t)*h_alive;
    // BFS levels are u8: a query deeper than 254 levels is closed at 255
    if (kind == 0)
        for (uint32_t j = 0; j < S; ++j)
            if (((used >> j) & 1) && round - slots[j].start_round + 2 > 255) doneset |= 1ull << j;
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
        // Admission: every engine gets the same power-of-two slot count S <= 64,
        // the largest with S * (sum of the engines' per-slot bytes) inside the
        // per-query state budget.
        uint64_t sum_slot = 0;
        for (og_graph *g : srv->graphs) sum_slot += g->V * (g->kind == OG_KIND_WEIGHTED ? 5 : 1);
        uint64_t budget = srv->params.state_budget_bytes;
        if (!budget) {
            size_t fr = 0, tot = 0;
            OG_HIP(hipMemGetInfo(&fr, &tot));
            budget = fr / 4;
        }
        uint32_t S = 64;
        while (S > 0 && (uint64_t)S * sum_slot > budget) S >>= 1;
        if (S == 0)
            raise(OG_EBUDGET, "state budget %llu bytes holds no query slot (%llu bytes per slot across engines)",
                  (unsigned long long)budget, (unsigned long long)sum_slot);
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
                const uint64_t V = g->V;
                e->per_slot_bytes = kind == 0 ? V : 4 * V;
                e->S = S;
                OG_HIP(hipSetDevice(g->dev->device));
                OG_HIP(hipStreamCreateWithFlags(&e->s, hipStreamNonBlocking));
                frontier_alloc(g->dev, g, e->f, e->mem);
                e->front = (uint64_t *)e->mem.dalloc(V * 8);
                e->next = (uint64_t *)e->mem.dalloc(V * 8);
                OG_HIP(hipMemset(e->front, 0, V * 8));
                OG_HIP(hipMemset(e->next, 0, V * 8));
                if (kind == 0) {
                    e->visit = (uint64_t *)e->mem.dalloc(V * 8);
                    e->lvl = (uint8_t *)e->mem.dalloc(V * S);
                    OG_HIP(hipMemset(e->visit, 0, V * 8));
                    OG_HIP(hipMemset(e->lvl, 0, V * S));
                } else {
                    e->dist = (uint32_t *)e->mem.dalloc(V * 4 * S);
                    OG_HIP(hipMemset(e->dist, 0xFF, V * 4 * S));
                }
                e->d_alive = (unsigned long long *)e->mem.dalloc(8);
                e->h_alive = (unsigned long long *)e->mem.palloc(8);
                e->d_fin = (unsigned long long *)e->mem.dalloc(3 * 64 * 8);
                e->h_fin = (unsigned long long *)e->mem.palloc(3 * 64 * 8);
                srv->state_bytes += e->per_slot_bytes * S;
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
