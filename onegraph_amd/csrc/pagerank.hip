// PageRank over the pull layout (outDegree + in-edge CSR) on gfx950.
//
// Reference shape: prSumKernel_static / prSumKernel_dynamic / prKernel_Opt
// (PR/OneGraph/gpu_kernels.h:107-158, 295-318) driven by pr_opt
// (PR/OneGraph/CalculateOpt.h:13-185).  The numerics follow the normalised
// definition instead of the reference's 0.15 + 0.85*sum:
//   contrib[u] = rank[u] / outdeg[u]            (0 for dangling u)
//   dangling   = sum of rank[u] over outdeg[u] == 0
//   new[v]     = (1-d)/n + d * (sum_{u in in(v)} contrib[u] + dangling/n)
//   stop when sum_v |new[v] - rank[v]| < tol or after max_iterations.
// FP32 mode keeps ranks in FP32 and sums in-edges with Kahan compensation;
// the dangling mass and the L1 residual are reduced in FP64 in a fixed order
// (per-block partials, then one workgroup), so runs are deterministic.
//
// Every vertex is active each iteration, so the dynamic region is the
// contiguous in-edge range [ptr[vsplit], E): it is copied chunk by chunk
// straight from the pinned host CSR (no gather) on the copy stream while the
// previous chunk is being summed.
#include <algorithm>

#include "device.h"

namespace og {

constexpr int kPrLpv0 = 8;  // lanes per vertex for the low-degree bin

template <typename T>
struct Acc {  // running sum with Kahan compensation: value = s - c
    T s = 0, c = 0;
    __device__ __forceinline__ void add(T x) {
        const T y = x - c;
        const T t = s + y;
        c = (t - s) - y;
        s = t;
    }
    // Merges another accumulator using an error-free TwoSum of the heads.
    __device__ __forceinline__ void merge(T s2, T c2) {
        const T t = s + s2;
        const T bp = t - s;
        const T e = (s - (t - bp)) + (s2 - bp);
        s = t;
        c = c + c2 - e;
    }
    __device__ __forceinline__ T value() const { return s - c; }
};

template <typename T>
__device__ __forceinline__ T shfl_xor_t(T x, int m) {
    return __shfl_xor(x, m, kWave);
}

// Reduces the accumulators of LPV consecutive lanes (LPV a power of two <= 64).
template <typename T, int LPV>
__device__ __forceinline__ void group_merge(Acc<T> &a) {
#pragma unroll
    for (int m = 1; m < LPV; m <<= 1) {
        const T s2 = shfl_xor_t(a.s, m), c2 = shfl_xor_t(a.c, m);
        a.merge(s2, c2);
    }
}

__device__ __forceinline__ double block_sum_f64(double x, double *lds) {
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
    for (int m = 1; m < kWave; m <<= 1) x += __shfl_xor(x, m, kWave);
    if (lane == 0) lds[wid] = x;
    __syncthreads();
    double t = 0;
    for (int w = 0; w < (int)(blockDim.x / kWave); ++w) t += lds[w];
    __syncthreads();
    return t;
}

// contrib[v] = rank[v] / outdeg[v]; per-block FP64 partials of the dangling mass.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pr_contrib(const T *__restrict__ rank, const uint32_t *__restrict__ outdeg,
                                                       T *__restrict__ contrib, uint64_t V, double *partial) {
    __shared__ double lds[kBlock / kWave];
    double dang = 0;
    for (uint64_t v = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v < V; v += (uint64_t)gridDim.x * kBlock) {
        const T r = rank[v];
        const uint32_t od = outdeg[v];
        contrib[v] = od ? r / (T)od : (T)0;
        if (!od) dang += (double)r;
    }
    dang = block_sum_f64(dang, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = dang;
}

// Sums n partials in a fixed order into *out.
__global__ __launch_bounds__(1024) void k_reduce_partials(const double *partial, uint32_t n, double *out) {
    __shared__ double lds[1024 / kWave];
    double s = 0;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) s += partial[i];
    s = block_sum_f64(s, lds);
    if (threadIdx.x == 0) *out = s;
}

template <typename T>
__device__ __forceinline__ T pr_apply(T sum, const double *scal, double damping, double inv_n) {
    // base = (1-d)/n + d * dangling / n, formed in FP64 then rounded once.
    const double base = (1.0 - damping) * inv_n + damping * scal[0] * inv_n;
    return (T)base + (T)damping * sum;
}

// Pull over a list of static vertices with LPV lanes each.  Residual
// partials go to partial[blockIdx.x + poff].
template <typename T, int LPV>
__global__ __launch_bounds__(kBlock) void k_pr_pull_list(const uint32_t *__restrict__ list, uint32_t n,
                                                         const uint64_t *__restrict__ ptr,
                                                         const uint32_t *__restrict__ edges, uint64_t edge_base,
                                                         const T *__restrict__ contrib, const T *__restrict__ rank,
                                                         T *__restrict__ next, const double *scal, double damping,
                                                         double inv_n, double *partial) {
    __shared__ double lds[kBlock / kWave];
    const uint64_t gid = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const uint64_t item = gid / LPV;
    const uint32_t sub = (uint32_t)(gid % LPV);
    Acc<T> a;
    uint32_t v = 0;
    if (item < n) {
        v = list[item];
        const uint64_t b = ptr[v], e = ptr[v + 1];
        for (uint64_t i = b + sub; i < e; i += LPV) a.add(contrib[edges[i - edge_base]]);
    }
    group_merge<T, LPV>(a);
    double res = 0;
    if (item < n && sub == 0) {
        const T nv = pr_apply<T>(a.value(), scal, damping, inv_n);
        res = fabs((double)nv - (double)rank[v]);
        next[v] = nv;
    }
    res = block_sum_f64(res, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = res;
}

// Pull for one hub vertex per workgroup (deterministic block reduction).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pr_pull_hub(const uint32_t *__restrict__ list,
                                                        const uint64_t *__restrict__ ptr,
                                                        const uint32_t *__restrict__ edges, uint64_t edge_base,
                                                        const T *__restrict__ contrib, const T *__restrict__ rank,
                                                        T *__restrict__ next, const double *scal, double damping,
                                                        double inv_n, double *partial) {
    __shared__ T s_s[kBlock / kWave], s_c[kBlock / kWave];
    const uint32_t v = list[blockIdx.x];
    const uint64_t b = ptr[v], e = ptr[v + 1];
    Acc<T> a;
    for (uint64_t i = b + threadIdx.x; i < e; i += kBlock) a.add(contrib[edges[i - edge_base]]);
    group_merge<T, kWave>(a);
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    if (lane == 0) {
        s_s[wid] = a.s;
        s_c[wid] = a.c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        Acc<T> t;
        t.s = s_s[0];
        t.c = s_c[0];
        for (int w = 1; w < kBlock / kWave; ++w) t.merge(s_s[w], s_c[w]);
        const T nv = pr_apply<T>(t.value(), scal, damping, inv_n);
        partial[blockIdx.x] = fabs((double)nv - (double)rank[v]);
        next[v] = nv;
    }
}

// Pull for the contiguous vertex range [va, vb) whose in-edges sit in `stage`
// starting at edge ptr[va].
template <typename T, int LPV>
__global__ __launch_bounds__(kBlock) void k_pr_pull_range(uint32_t va, uint32_t vb, const uint64_t *__restrict__ ptr,
                                                          const uint32_t *__restrict__ stage,
                                                          const T *__restrict__ contrib, const T *__restrict__ rank,
                                                          T *__restrict__ next, const double *scal, double damping,
                                                          double inv_n, double *partial) {
    __shared__ double lds[kBlock / kWave];
    const uint64_t gid = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const uint64_t item = gid / LPV;
    const uint32_t sub = (uint32_t)(gid % LPV);
    const uint64_t base = ptr[va];
    Acc<T> a;
    const uint64_t v = va + item;
    if (v < vb) {
        const uint64_t b = ptr[v] - base, e = ptr[v + 1] - base;
        for (uint64_t i = b + sub; i < e; i += LPV) a.add(contrib[stage[i]]);
    }
    group_merge<T, LPV>(a);
    double res = 0;
    if (v < vb && sub == 0) {
        const T nv = pr_apply<T>(a.value(), scal, damping, inv_n);
        res = fabs((double)nv - (double)rank[v]);
        next[v] = nv;
    }
    res = block_sum_f64(res, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = res;
}

template <typename T>
__global__ void k_fill_t(T *a, uint64_t n, T v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        a[i] = v;
}

// --------------------------------------------------------------- host

namespace {

struct PrPlan {
    std::vector<uint32_t> bin0, bin1, hub;            // static vertices by in-degree
    std::vector<std::pair<uint32_t, uint32_t>> chunks;  // dynamic vertex ranges
    uint32_t *d_bins = nullptr;
    uint64_t partial_cap = 0;
};

unsigned grid_of(uint64_t threads) {
    return (unsigned)std::max<uint64_t>(1, (threads + kBlock - 1) / kBlock);
}

}  // namespace

// Per-graph PageRank plan, cached on the device state.
static PrPlan &pr_plan(DeviceState *d, const og_graph *g) {
    if (d->pr_plan) return *(PrPlan *)d->pr_plan.get();
    std::shared_ptr<PrPlan> p = std::make_shared<PrPlan>();
    const uint64_t vs = d->plan.vsplit;
    for (uint64_t v = 0; v < vs; ++v) {
        const uint64_t deg = g->ptr[v + 1] - g->ptr[v];
        if (deg <= d->tune.t1) p->bin0.push_back((uint32_t)v);
        else if (deg <= d->tune.t2) p->bin1.push_back((uint32_t)v);
        else p->hub.push_back((uint32_t)v);
    }
    if (vs < d->V) {
        const uint64_t C = d->slot_edges;
        uint64_t va = vs;
        while (va < d->V) {
            uint64_t vb = va;
            // greedy: extend while the range fits in one slot
            uint64_t lo = va + 1, hi = d->V;
            if (g->ptr[va + 1] - g->ptr[va] > C)
                raise(OG_EBUDGET, "vertex %llu has %llu in-edges, more than a staging slot (%llu)",
                      (unsigned long long)va, (unsigned long long)(g->ptr[va + 1] - g->ptr[va]),
                      (unsigned long long)C);
            while (lo < hi) {  // largest vb with ptr[vb] - ptr[va] <= C
                uint64_t mid = (lo + hi + 1) / 2;
                if (g->ptr[mid] - g->ptr[va] <= C) lo = mid; else hi = mid - 1;
            }
            vb = lo;
            p->chunks.push_back({(uint32_t)va, (uint32_t)vb});
            va = vb;
        }
    }
    const uint64_t nb = p->bin0.size() + p->bin1.size() + p->hub.size();
    p->d_bins = (uint32_t *)d->dalloc(std::max<uint64_t>(nb, 1) * 4);
    std::vector<uint32_t> all;
    all.reserve(nb);
    all.insert(all.end(), p->bin0.begin(), p->bin0.end());
    all.insert(all.end(), p->bin1.begin(), p->bin1.end());
    all.insert(all.end(), p->hub.begin(), p->hub.end());
    if (nb) OG_HIP(hipMemcpy(p->d_bins, all.data(), nb * 4, hipMemcpyHostToDevice));
    // partial slots: one per block of every launch in an iteration
    uint64_t cap = grid_of(p->bin0.size() * (uint64_t)kPrLpv0) + grid_of(p->bin1.size() * 64ull) + p->hub.size();
    for (auto &c : p->chunks) cap += grid_of((uint64_t)(c.second - c.first) * kPrLpv0);
    p->partial_cap = std::max<uint64_t>(cap, grid_of(d->V));
    d->pr_plan = p;
    return *p;
}

template <typename T>
static void run_pr(og_graph *g, const og_pr_params &prm, T *out, og_pr_stats *stats) {
    DeviceState *d = g->dev;
    const uint64_t V = d->V;
    PrPlan &pp = pr_plan(d, g);
    hipStream_t s = d->s_compute;
    const bool f64 = sizeof(T) == 8;
    if (!d->d_outdeg) {
        d->d_outdeg = (uint32_t *)d->dalloc(V * 4);
        OG_HIP(hipMemcpy(d->d_outdeg, g->out_degree, V * 4, hipMemcpyHostToDevice));
        d->d_scalars = (double *)d->dalloc(4 * sizeof(double));
        d->h_scalars = (double *)d->palloc(4 * sizeof(double));
    }
    if (pp.partial_cap > d->pr_partial_cap) {
        d->d_partial = (double *)d->dalloc(pp.partial_cap * sizeof(double));
        d->pr_partial_cap = pp.partial_cap;
    }
    if (!d->d_rank[0] || d->pr_fp64 != f64) {
        // sized for double so either precision fits
        if (!d->d_rank[0]) {
            d->d_rank[0] = d->dalloc(V * 8);
            d->d_rank[1] = d->dalloc(V * 8);
            d->d_contrib = d->dalloc(V * 8);
        }
        d->pr_fp64 = f64;
    }
    T *rank = (T *)d->d_rank[0], *next = (T *)d->d_rank[1], *contrib = (T *)d->d_contrib;
    const double inv_n = 1.0 / (double)V, damping = prm.damping;
    const uint32_t *edges = (const uint32_t *)d->d_static;
    TimedClass tc_contrib, tc_static, tc_dyn, tc_copy;
    d->events.reset();
    og_pr_stats st{};
    st.run.approach = (uint32_t)d->approach;

    OG_HIP(hipStreamSynchronize(d->s_copy));
    OG_HIP(hipStreamSynchronize(s));
    Stopwatch total;
    hipLaunchKernelGGL(k_fill_t<T>, dim3(2048), dim3(256), 0, s, rank, V, (T)(1.0 / (double)V));
    OG_HIP(hipGetLastError());
    const unsigned cgrid = (unsigned)std::min<uint64_t>(2048, grid_of(V));
    const uint32_t n0 = (uint32_t)pp.bin0.size(), n1 = (uint32_t)pp.bin1.size(), nh = (uint32_t)pp.hub.size();
    const uint64_t eb = 4, C = d->slot_edges;
    double l1 = 0;
    uint32_t it = 0;
    const uint64_t vb_bytes = f64 ? 8 : 4;
    while (it < prm.max_iterations) {
        {
            EventSpan span(d, &tc_contrib, s);
            hipLaunchKernelGGL(k_pr_contrib<T>, dim3(cgrid), dim3(kBlock), 0, s, rank, d->d_outdeg, contrib, V,
                               d->d_partial);
            hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(1024), 0, s, d->d_partial, cgrid, d->d_scalars + 0);
            OG_HIP(hipGetLastError());
        }
        st.run.frontier_bytes += V * (4 + 2 * vb_bytes);
        uint64_t poff = 0;
        {
            if (n0 | n1 | nh) {
                EventSpan span(d, &tc_static, s);
                if (n0) {
                    const unsigned gr = grid_of((uint64_t)n0 * kPrLpv0);
                    hipLaunchKernelGGL((k_pr_pull_list<T, kPrLpv0>), dim3(gr), dim3(kBlock), 0, s, pp.d_bins, n0,
                                       d->d_ptr, edges, 0ull, contrib, rank, next, d->d_scalars, damping, inv_n,
                                       d->d_partial + poff);
                    poff += gr;
                }
                if (n1) {
                    const unsigned gr = grid_of((uint64_t)n1 * 64);
                    hipLaunchKernelGGL((k_pr_pull_list<T, 64>), dim3(gr), dim3(kBlock), 0, s, pp.d_bins + n0, n1,
                                       d->d_ptr, edges, 0ull, contrib, rank, next, d->d_scalars, damping, inv_n,
                                       d->d_partial + poff);
                    poff += gr;
                }
                if (nh) {
                    hipLaunchKernelGGL((k_pr_pull_hub<T>), dim3(nh), dim3(kBlock), 0, s, pp.d_bins + n0 + n1,
                                       d->d_ptr, edges, 0ull, contrib, rank, next, d->d_scalars, damping, inv_n,
                                       d->d_partial + poff);
                    poff += nh;
                }
                OG_HIP(hipGetLastError());
            }
        }
        const uint64_t sedges = d->plan.vsplit ? g->ptr[d->plan.vsplit] : 0;
        st.run.relax_static_bytes += (16 + 2 * vb_bytes) * d->plan.vsplit + (4 + vb_bytes) * sedges;
        st.run.static_work_edges += sedges;
        for (size_t c = 0; c < pp.chunks.size(); ++c) {
            const uint32_t va = pp.chunks[c].first, vb = pp.chunks[c].second;
            const uint64_t e0 = g->ptr[va], e1 = g->ptr[vb];
            const int slot = (int)(c % d->nslots);
            uint8_t *dslot = (uint8_t *)d->d_staging + (uint64_t)slot * C * eb;
            if (d->slot_relax_pending[slot]) OG_HIP(hipStreamWaitEvent(d->s_copy, d->ev_relax[slot], 0));
            if (e1 > e0) {
                EventSpan span(d, &tc_copy, d->s_copy);
                OG_HIP(hipMemcpyAsync(dslot, g->cols + e0, (e1 - e0) * eb, hipMemcpyHostToDevice, d->s_copy));
            }
            OG_HIP(hipEventRecord(d->ev_h2d[slot], d->s_copy));
            OG_HIP(hipStreamWaitEvent(s, d->ev_h2d[slot], 0));
            {
                EventSpan span(d, &tc_dyn, s);
                const unsigned gr = grid_of((uint64_t)(vb - va) * kPrLpv0);
                hipLaunchKernelGGL((k_pr_pull_range<T, kPrLpv0>), dim3(gr), dim3(kBlock), 0, s, va, vb, d->d_ptr,
                                   (const uint32_t *)dslot, contrib, rank, next, d->d_scalars, damping, inv_n,
                                   d->d_partial + poff);
                OG_HIP(hipGetLastError());
                poff += gr;
            }
            OG_HIP(hipEventRecord(d->ev_relax[slot], s));
            d->slot_relax_pending[slot] = true;
            st.run.h2d_bytes += (e1 - e0) * eb;
            st.run.chunks++;
            st.run.relax_dynamic_bytes += (16 + 2 * vb_bytes) * (vb - va) + (4 + vb_bytes) * (e1 - e0);
        }
        hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(1024), 0, s, d->d_partial, (uint32_t)poff,
                           d->d_scalars + 1);
        OG_HIP(hipGetLastError());
        OG_HIP(hipMemcpyAsync(d->h_scalars, d->d_scalars, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
        OG_HIP(hipStreamSynchronize(s));
        l1 = d->h_scalars[1];
        std::swap(rank, next);
        ++it;
        st.run.work_edges += g->E;
        if (l1 < prm.tolerance) break;
    }
    OG_HIP(hipStreamSynchronize(s));
    st.run.total_ms = total.ms();
    if (out) copy_to_host(d, out, rank, V * sizeof(T), s);
    // keep rank buffer order stable for the next call
    if (rank != (T *)d->d_rank[0]) std::swap(d->d_rank[0], d->d_rank[1]);
    st.iterations = it;
    st.final_l1 = l1;
    st.run.iterations = it;
    st.run.reached = V;
    st.run.traversed_edges = g->E;
    st.run.node_sum = (uint64_t)it * V;
    st.run.frontier_kernel_ms = sum_ms(tc_contrib);
    st.run.frontier_launches = tc_contrib.launches;
    st.run.relax_static_kernel_ms = sum_ms(tc_static);
    st.run.relax_static_launches = tc_static.launches;
    st.run.relax_dynamic_kernel_ms = sum_ms(tc_dyn);
    st.run.relax_dynamic_launches = tc_dyn.launches;
    st.run.copy_ms = sum_ms(tc_copy);
    st.run.pre_ms = st.run.frontier_kernel_ms;
    st.run.static_ms = st.run.relax_static_kernel_ms;
    st.run.overload_ms = st.run.relax_dynamic_kernel_ms;
    st.run.overload_move_ms = st.run.copy_ms;
    if (stats) *stats = st;
}

}  // namespace og

using namespace og;

extern "C" og_status og_pagerank(og_graph *g, const og_pr_params *params, float *ranks_f32, double *ranks_f64,
                                 og_pr_stats *stats) {
    clear_error();
    if (!g || !params) return fail(OG_EINVAL, "graph/params is NULL");
    if (!g->dev) return fail(OG_ESTATE, "og_prepare has not been called for this graph");
    if (g->kind != OG_KIND_PULL) return fail(OG_EUNSUPPORTED, "PageRank needs a pull-layout graph");
    if (!(params->damping > 0 && params->damping < 1)) return fail(OG_EINVAL, "damping must be in (0, 1)");
    if (params->max_iterations == 0) return fail(OG_EINVAL, "max_iterations must be > 0");
    int expect = 0;
    if (!g->busy.compare_exchange_strong(expect, 1)) return fail(OG_EBUSY, "graph is running another call");
    og_status st = guard([&]() -> og_status {
        OG_HIP(hipSetDevice(g->dev->device));
        if (params->fp64)
            run_pr<double>(g, *params, ranks_f64, stats);
        else
            run_pr<float>(g, *params, ranks_f32, stats);
        return OG_OK;
    });
    g->busy.store(0);
    return st;
}
