// Frontier engine for BFS, SSSP and CC on gfx950: wave64 ballot/scan frontier
// compaction, degree-binned relax kernels over the HBM-resident static region,
// and the dynamic-region loader that streams active edges from pinned host
// memory through a double-buffered staging area on a copy stream.
//
// One iteration (reference loop: BFS/oneGraph_BFS/CalculateOpt.h:75-188):
//   1. k_frontier_count / k_frontier_scan (one workgroup) / k_frontier_write:
//      two passes over the next-frontier flags build the
//      static bin lists (ordered), the hub tile list, the sorted dynamic list
//      and its degree prefix, and clear the flags.  Replaces K1-K6
//      (gpu_kernels.h:243-379), the host prefix H4 (CalculateOpt.h:110-123)
//      and every per-iteration .wait(): counts stay on the device, and rounds
//      without a host-side loader are replayed from captured hipGraphs.
//   2. static relax: bin 0 (lpv0 lanes/vertex), bin 1 (wave/vertex), bin 2
//      (workgroup per `tile` edges, neighbour list staged in LDS).  Replaces
//      bfs_kernelStatic (gpu_kernels.h:211-241), sssp_kernel
//      (SSSP/oneGraph/gpu_kernels.h:177-203), cc_kernelStaticSwap
//      (CC/onegraph/gpu_kernels.h:133-157).
//   3. dynamic region: the host gathers the active edges of the sorted dynamic
//      list into a pinned slot (H5, GraphMeta.h:470-506), hipMemcpyAsync moves
//      the slot on the copy stream while the other slot is being gathered, and
//      an edge-parallel relax consumes it (K8/K10/K12).
#include <algorithm>
#include <atomic>
#include <vector>
#include <cstdio>
#include <cstring>

#include "device.h"

namespace og {

// ================================================================= kernels

__global__ void k_fill_u32(uint32_t *a, uint64_t n, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        a[i] = v;
}

__global__ void k_iota(uint32_t *a, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        a[i] = (uint32_t)i;
}

__global__ void k_fill_u8(uint8_t *a, uint64_t n, uint8_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        a[i] = v;
}

__global__ void k_set_source(uint32_t *val, uint8_t *flag, uint32_t src) {
    val[src] = 1;
    flag[src] = 1;
}

// Per-thread flag window: 16 * FV consecutive vertices.
template <int FV>
__device__ __forceinline__ void load_flags(const uint8_t *flag, uint64_t base, uint64_t Vpad, uint4 (&w)[FV]) {
#pragma unroll
    for (int k = 0; k < FV; ++k) {
        uint64_t at = base + 16ull * k;
        w[k] = at < Vpad ? *(const uint4 *)(flag + at) : make_uint4(0, 0, 0, 0);
    }
}

template <int FV>
__device__ __forceinline__ void store_flags(uint8_t *flag, uint64_t base, uint64_t Vpad,
                                            const uint32_t (&word)[4 * FV]) {
#pragma unroll
    for (int k = 0; k < FV; ++k) {
        const uint64_t at = base + 16ull * k;
        if (at < Vpad)
            *(uint4 *)(flag + at) = make_uint4(word[4 * k], word[4 * k + 1], word[4 * k + 2], word[4 * k + 3]);
    }
}

// Bin of an active vertex: 0..2 static bins, 3 dynamic, 4 no out-edges,
// 5 held back for a later round (near-far order: value >= threshold, checked
// first so that held-back vertices need no degree).
__device__ __forceinline__ int bin_of(bool far, uint64_t v, uint64_t vsplit, uint64_t deg, const Tuning &tn) {
    if (far) return 5;
    if (deg == 0) return 4;
    if (v >= vsplit) return 3;
    if (deg <= tn.t1) return 0;
    if (deg <= tn.t2) return 1;
    return 2;
}

__device__ __forceinline__ uint32_t block_min_u32(uint32_t x, uint64_t *lds) {
#pragma unroll
    for (int m = 1; m < kWave; m <<= 1) x = min(x, (uint32_t)__shfl_xor(x, m, kWave));
    __syncthreads();
    if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = x;
    __syncthreads();
    uint32_t r = 0xFFFFFFFFu;
    for (int w = 0; w < (int)(blockDim.x / kWave); ++w) r = min(r, (uint32_t)lds[w]);
    __syncthreads();
    return r;
}

constexpr int kLoadBatch = 8;  // flagged vertices whose row bounds are read together

// Bit j set when byte j of the thread's flag window is >= lo.  The passes
// then walk the set bits only (lowest first, i.e. in vertex order), which
// keeps their code small: a per-slot unrolled switch over 32 slots does not
// fit the instruction cache.
template <int FV>
__device__ __forceinline__ uint32_t window_mask(const uint4 (&w)[FV], uint32_t lo) {
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < FV; ++k) {
        const uint32_t q[4] = {w[k].x, w[k].y, w[k].z, w[k].w};
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                m |= (((q[i] >> (8 * b)) & 0xFFu) >= lo ? 1u : 0u) << (16 * k + 4 * i + b);
    }
    return m;
}

__device__ void scan_counts(const BlockCounts &bc, FrontierInfo *__restrict__ info, uint64_t *__restrict__ doff,
                            RoundCtl *__restrict__ ctl, uint64_t *__restrict__ m_doff, uint64_t *lds);

// Most lanes of the one-workgroup scan, one per super-block
constexpr uint32_t kScanThreads = 1024;

// Super-block totals between compactions: sums zero, the min row all ones.
__global__ void k_sb_init(uint64_t *sbt, uint32_t nsb) {
    const uint64_t n = (uint64_t)kSbVals * nsb;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        sbt[i] = i / nsb == kSbVals - 1 ? ~0ull : 0ull;
}

// Compaction pass 1: classifies every flagged vertex of the block's window,
// stores kCode0 + bin in its flag byte and publishes the per-block counts.
// The window's flags are loaded 16 bytes per lane and transposed through
// LDS, so that slot j of thread t is vertex j * kBlock + t of the window:
// consecutive lanes then read consecutive row bounds and values and store
// consecutive flag bytes (only the block totals depend on the mapping).  Row
// bounds (and values, for near-far) of kLoadBatch flagged vertices are read
// before any is used, so their latencies overlap.
template <int FV>
__global__ __launch_bounds__(kBlock) void k_frontier_count(uint8_t *__restrict__ flag_a, uint8_t *__restrict__ hflag,
                                                           uint64_t Vpad, uint64_t vsplit,
                                                           const uint64_t *__restrict__ ptr,
                                                           const uint64_t *__restrict__ pmid,
                                                           const uint32_t *__restrict__ deg,
                                                           const uint32_t *__restrict__ hdeg, Tuning tn,
                                                           const uint32_t *__restrict__ val_a,
                                                           const RoundCtl *__restrict__ ctl, BlockCounts bc) {
    const uint32_t nblk = bc.nblk;
    __shared__ uint64_t lds[6 * kBlock / kWave];
    __shared__ uint4 s_win[(kBlock * 16 * FV) / 16];  // the window's flag bytes
    // a round enqueued after the traversal ended: no vertex is flagged and
    // the block counts of the last pass (all zero) stand
    if (ctl->halt && !ctl->live) return;
    // light/heavy split: a heavy compaction collects the heavy-pending
    // vertices (no near-far test) and sizes them by their heavy edges
    const bool heavy = pmid && ctl->mode;
    uint8_t *flag = heavy ? hflag : flag_a;
    const uint32_t *val = heavy ? nullptr : val_a;
    const uint64_t wbase = (uint64_t)blockIdx.x * (kBlock * 16 * FV);
    uint4 w[FV];
    load_flags(flag, wbase + (uint64_t)threadIdx.x * (16 * FV), Vpad, w);
    uint32_t nz = 0;
#pragma unroll
    for (int k = 0; k < FV; ++k) nz |= w[k].x | w[k].y | w[k].z | w[k].w;
    // a window without flags (most of them in a small round) publishes zero
    // counts without the transpose and the block scans; the write pass
    // skips it by row 7
    if (!__syncthreads_or(nz != 0)) {
        if (threadIdx.x == 0) {
#pragma unroll
            for (int r = 0; r < kBcntRows; ++r) bc.bcnt[r * nblk + blockIdx.x] = 0;
            bc.bdeg[blockIdx.x] = 0;
        }
        return;
    }
#pragma unroll
    for (int k = 0; k < FV; ++k) s_win[threadIdx.x * FV + k] = w[k];
    __syncthreads();
    const uint8_t *sb = (const uint8_t *)s_win;
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < (16 * FV); ++j) m |= (sb[j * kBlock + threadIdx.x] != 0 ? 1u : 0u) << j;
    const uint64_t base = wbase + threadIdx.x;  // slot j: vertex base + j * kBlock
    uint32_t c0 = 0, c1 = 0, c2 = 0, c3 = 0, cz = 0, ch = 0, cf = 0, cl = 0, mf = 0xFFFFFFFFu;
    uint64_t sdeg = 0, ddeg = 0;
    if (m) {
        const uint32_t thresh = val ? ctl->thresh : 0xFFFFFFFFu;
        // lagged order: a dynamic vertex below the threshold but not below
        // fbound waits until the static part of its bucket has settled
        const bool lagged = val && ctl->defer == 2;
        const uint32_t fbound = lagged ? ctl->fbound : 0u;
        while (m) {
            uint32_t jj[kLoadBatch];
            uint64_t dg[kLoadBatch];
            uint32_t vv[kLoadBatch];
            bool far[kLoadBatch];
#pragma unroll
            for (int k = 0; k < kLoadBatch; ++k) {
                jj[k] = (16 * FV);  // empty slot
                dg[k] = 0;
                vv[k] = 0;
                if (m) {
                    const uint32_t j = (uint32_t)__builtin_ctz(m);
                    m &= m - 1;
                    jj[k] = j;
                    if (val) vv[k] = val[base + (uint64_t)j * kBlock];
                    else dg[k] = round_deg(deg, hdeg, ptr, pmid, heavy, base + (uint64_t)j * kBlock);
                }
            }
            // near-far: a held-back vertex needs only its value, so the
            // degrees are read for the near ones alone
#pragma unroll
            for (int k = 0; k < kLoadBatch; ++k) {
                far[k] = false;
                if (jj[k] == (16 * FV) || !val) continue;
                const bool lag = lagged && vv[k] < thresh && vv[k] >= fbound && base + (uint64_t)jj[k] * kBlock >= vsplit;
                far[k] = vv[k] >= thresh || lag;
                if (!far[k]) dg[k] = round_deg(deg, hdeg, ptr, pmid, heavy, base + (uint64_t)jj[k] * kBlock);
            }
#pragma unroll
            for (int k = 0; k < kLoadBatch; ++k) {
                if (jj[k] == (16 * FV)) continue;
                const uint64_t d = dg[k];
                const bool lag = lagged && vv[k] < thresh && vv[k] >= fbound && base + (uint64_t)jj[k] * kBlock >= vsplit;
                cl += lag ? 1u : 0u;
                const int b = bin_of(far[k], base + (uint64_t)jj[k] * kBlock, vsplit, d, tn);
                switch (b) {
                    case 0: ++c0; sdeg += d; break;
                    case 1: ++c1; sdeg += d; break;
                    case 2: c2 += (uint32_t)((d + tn.tile - 1) / tn.tile); ++ch; sdeg += d; break;
                    case 3: ++c3; ddeg += d; break;
                    case 5: ++cf; mf = min(mf, vv[k]); break;
                    default: ++cz; break;
                }
                flag[base + (uint64_t)jj[k] * kBlock] = (uint8_t)(kCode0 + b);
            }
        }
    }
    const uint64_t x[6] = {(uint64_t)c0 | ((uint64_t)c1 << 32), (uint64_t)c2 | ((uint64_t)c3 << 32),
                           (uint64_t)cz | ((uint64_t)ch << 32), sdeg, ddeg, (uint64_t)cf | ((uint64_t)cl << 16)};
    uint64_t ex[6], t[6];
    block_scan_n<6>(x, ex, t, lds);
    const uint32_t bm = block_min_u32(mf, lds);
    if (threadIdx.x == 0) {
        bc.bcnt[0 * nblk + blockIdx.x] = (uint32_t)t[0];
        bc.bcnt[1 * nblk + blockIdx.x] = (uint32_t)(t[0] >> 32);
        bc.bcnt[2 * nblk + blockIdx.x] = (uint32_t)t[1];
        bc.bcnt[3 * nblk + blockIdx.x] = (uint32_t)(t[1] >> 32);
        bc.bcnt[4 * nblk + blockIdx.x] = 1;
        bc.bdeg[blockIdx.x] = t[4];
    }
    // the block's totals into its super-block's, one lane per value (packed
    // pairs add without carries: each half stays below 2^32)
    if (threadIdx.x < (unsigned)kSbVals) {
        const uint32_t sb = blockIdx.x / kSbBlocks;
        const uint64_t v[kSbVals] = {t[0], t[1], t[4], (uint32_t)t[2], t[2] >> 32, t[5] & 0xFFFFu, t[3],
                                     t[5] >> 16, bm};
        uint64_t x = 0;
#pragma unroll
        for (int i = 0; i < kSbVals; ++i)
            if ((int)threadIdx.x == i) x = v[i];
        unsigned long long *at = (unsigned long long *)bc.sbt + threadIdx.x * (uint64_t)bc.nsb + sb;
        if (threadIdx.x == kSbVals - 1) {
            if (x != 0xFFFFFFFFull) atomicMin(at, (unsigned long long)x);
        } else if (x) {
            atomicAdd(at, (unsigned long long)x);
        }
    }
}

// Compaction pass 2: one workgroup scans the counts (scan_counts).  A
// separate launch rather than a last-workgroup-arrives tail in pass 1: on
// gfx950 the device-scope release/acquire fences such a tail needs write
// back and invalidate the whole L2 of each XCD, once per workgroup.
__global__ __launch_bounds__(kScanThreads) void k_frontier_scan(BlockCounts bc, FrontierInfo *__restrict__ info,
                                                                uint64_t *__restrict__ doff,
                                                                RoundCtl *__restrict__ ctl,
                                                                uint64_t *__restrict__ m_doff) {
    __shared__ uint64_t lds[8 * kScanThreads / kWave];
    scan_counts(bc, info, doff, ctl, m_doff, lds);
}

// Compaction pass 2 body: one workgroup turns the per-block counts into
// exclusive offsets (in place), publishes
// the round's summary and advances the traversal state: statistics,
// liveness, the BFS direction and, when the near set is empty but vertices
// are held back, the near-far threshold (min held-back value + delta) for
// the next compaction.  Each thread scans a run of consecutive blocks, so the
// whole pass is one multi-value block scan.  `lds` holds 8 * blockDim / 64.
__device__ void scan_counts(const BlockCounts &bc, FrontierInfo *__restrict__ info, uint64_t *__restrict__ doff,
                            RoundCtl *__restrict__ ctl, uint64_t *__restrict__ m_doff, uint64_t *lds) {
    // a round enqueued after the traversal ended (its count pass returned at
    // once): the summary and the traversal state stand
    if (ctl->halt && !ctl->live) return;
    const uint32_t nsb = bc.nsb;
    const uint64_t *sbt = bc.sbt;
    const uint32_t per = (nsb + blockDim.x - 1) / blockDim.x;
    const uint32_t b0 = min(nsb, threadIdx.x * per), b1 = min(nsb, b0 + per);
    uint64_t s01 = 0, s23 = 0, sdd = 0, zs = 0, hs = 0, fs = 0, ss = 0, ls = 0;
    uint32_t fm = 0xFFFFFFFFu;
    for (uint32_t b = b0; b < b1; ++b) {
        s01 += sbt[0 * nsb + b];
        s23 += sbt[1 * nsb + b];
        sdd += sbt[2 * nsb + b];
        zs += sbt[3 * nsb + b];
        hs += sbt[4 * nsb + b];
        fs += sbt[5 * nsb + b];
        ss += sbt[6 * nsb + b];
        ls += sbt[7 * nsb + b];
        fm = (uint32_t)min((uint64_t)fm, sbt[8 * nsb + b]);
    }
    const uint64_t x[8] = {s01, s23, sdd, zs, hs, fs, ss, ls};
    uint64_t ex[8], t[8];
    block_scan_n<8>(x, ex, t, lds);
    const uint32_t fmin = block_min_u32(fm, lds);
    // super-block prefixes for the write pass, and the totals cleared for
    // the next compaction
    uint64_t r01 = ex[0], r23 = ex[1], rdd = ex[2];
    for (uint32_t b = b0; b < b1; ++b) {
        bc.sbp[0 * nsb + b] = r01;
        bc.sbp[1 * nsb + b] = r23;
        bc.sbp[2 * nsb + b] = rdd;
        r01 += sbt[0 * nsb + b];
        r23 += sbt[1 * nsb + b];
        rdd += sbt[2 * nsb + b];
#pragma unroll
        for (int i = 0; i < kSbVals - 1; ++i) bc.sbt[i * nsb + b] = 0;
        bc.sbt[(kSbVals - 1) * nsb + b] = ~0ull;
    }
    if (threadIdx.x == 0) {
        FrontierInfo fi;
        fi.n[0] = (uint32_t)t[0];
        fi.n[1] = (uint32_t)(t[0] >> 32);
        fi.n[2] = (uint32_t)t[1];
        fi.n[3] = (uint32_t)(t[1] >> 32);
        fi.nzero = (uint32_t)t[3];
        fi.nhub = (uint32_t)t[4];
        fi.static_deg = t[6];
        fi.dyn_deg = t[2];
        fi.nfar = (uint32_t)t[5];
        fi.min_far = fmin;
        RoundCtl c = *ctl;
        const uint64_t snodes = (uint64_t)fi.n[0] + fi.n[1] + fi.nhub;
        // deferred order: while static vertices are active, the dynamic list
        // stays flagged for a later round (the write pass keeps its flags)
        const bool was_heavy = c.mode != 0;
        fi.dyn_held = (c.defer == 1 && !was_heavy && snodes && fi.n[3]) ? 1u : 0u;
        const uint32_t held = fi.dyn_held ? fi.n[3] : 0u;
        if (held) {
            fi.n[3] = 0;
            fi.dyn_deg = 0;
        }
        const uint64_t active = snodes + fi.n[3] + fi.nzero;
        // light/heavy split: light rounds count the vertices they relax (they
        // turn heavy-pending); once a bucket's light rounds find nothing near,
        // the next compaction collects the heavy-pending vertices (their
        // distances are final within the bucket) before the threshold moves
        bool heavy_next = false;
        if (c.split) {
            if (was_heavy) {
                c.hpend = 0;
            } else if (active) {
                c.hpend += active;
            } else if (c.hpend) {
                heavy_next = true;
            }
        }
        // (after a heavy round a light compaction always follows: the held-back
        // vertices sit in the other flag array)
        fi.live = (active || fi.nfar || held || heavy_next || was_heavy) ? 1u : 0u;
        // BFS direction (Beamer et al., SC'12): bottom-up once the frontier's
        // edges exceed 1/14 of the unexplored ones, back to top-down when
        // the frontier falls below V / 24.
        fi.pull = 0;
        fi.heavy = c.mode;
        if (c.dir_opt && active) {
            const uint64_t mf = fi.static_deg + fi.dyn_deg;
            if (!c.pull && mf > c.unexplored / 14) c.pull = 1;
            else if (c.pull && active < c.nverts / 24) c.pull = 0;
            c.unexplored -= min(mf, c.unexplored);
            fi.pull = c.pull;
            c.pull_rounds += c.pull;
        }
        *info = fi;
        doff[fi.n[3]] = fi.dyn_deg;
        if (m_doff) m_doff[fi.n[3]] = fi.dyn_deg;
        c.compactions++;
        c.live = fi.live;
        if (active) {
            c.rounds++;
            c.node_sum += active;
            c.static_node_sum += snodes;
            c.dyn_node_sum += fi.n[3];
            c.work_edges += fi.static_deg + fi.dyn_deg;
            c.static_work_edges += fi.static_deg;
            c.dyn_rounds += fi.n[3] ? 1 : 0;
        } else if (heavy_next) {
            // the threshold stays: the heavy round's offers are all >= it
        } else if (fi.nfar && c.delta) {
            // Near set empty.  Lagged order: every vertex below the threshold
            // has relaxed except the waiting dynamic ones, which go next (a
            // static vertex they improve is below the threshold and joins
            // them).  Otherwise every held-back value is >= the threshold and
            // it moves to the smallest of them + delta.
            if (t[7]) c.fbound = c.thresh;
            else c.thresh = (uint32_t)min((uint64_t)0xFFFFFFFFull, (uint64_t)fi.min_far + c.delta);
        }
        c.mode = heavy_next ? 1u : 0u;
        *ctl = c;
    }
}

// Compaction pass 3: emits the lists in vertex order from the codes of pass
// 1 (bins 0/1 need no reads; hub tiles and the dynamic degree prefix read
// the row bounds again, kLoadBatch at a time) with the chunk table of the
// dynamic relax (dchunk[c]: the list entry holding edge c *
// kDynEdgesPerBlock of the dynamic edge space), and rewrites the window:
// 1 for held-back vertices, 0 elsewhere.  m_dlist / m_doff, when not null,
// receive a copy of the dynamic list and its degree prefix: pinned host
// memory the staged loader gathers from without a copy back.
template <int FV>
__global__ __launch_bounds__(kBlock) void k_frontier_write(uint8_t *__restrict__ flag_a, uint8_t *__restrict__ hflag,
                                                           uint64_t Vpad, const uint64_t *__restrict__ ptr,
                                                           const uint64_t *__restrict__ pmid,
                                                           const uint32_t *__restrict__ deg,
                                                           const uint32_t *__restrict__ hdeg, Tuning tn,
                                                           BlockCounts bc, const FrontierInfo *__restrict__ info,
                                                           uint32_t *__restrict__ list, uint64_t *__restrict__ hub,
                                                           uint64_t hub_cap, uint64_t *__restrict__ doff,
                                                           uint32_t *__restrict__ m_dlist,
                                                           uint64_t *__restrict__ m_doff,
                                                           uint32_t *__restrict__ dchunk, uint64_t dchunk_cap) {
    __shared__ uint64_t lds[3 * kBlock / kWave];
    __shared__ uint32_t s_win[4 * FV * kBlock];  // word i of thread t at [i * kBlock + t]
    __shared__ uint64_t s_pre[3];                // the block's exclusive prefixes (bins 0|1, 2|3, degrees)
    const uint32_t nblk = bc.nblk;
    if (!info->live) return;  // nothing is flagged: no list, and the window is already clear
    if (!bc.bcnt[4 * nblk + blockIdx.x]) return;  // the count pass found no flag in this window
    // prefix = the super-block's + the counts of the blocks before this one
    // in its super-block (one lane each; read before the block scan below,
    // whose barriers publish s_pre)
    if (threadIdx.x < kWave) {
        static_assert(kSbBlocks <= kWave, "one lane per block of a super-block");
        const uint32_t sb = blockIdx.x / kSbBlocks, q = sb * kSbBlocks + threadIdx.x;
        uint64_t a01 = 0, a23 = 0, add = 0;
        if (q < blockIdx.x) {
            a01 = (uint64_t)bc.bcnt[0 * nblk + q] | ((uint64_t)bc.bcnt[1 * nblk + q] << 32);
            a23 = (uint64_t)bc.bcnt[2 * nblk + q] | ((uint64_t)bc.bcnt[3 * nblk + q] << 32);
            add = bc.bdeg[q];
        }
#pragma unroll
        for (int m = kWave / 2; m >= 1; m >>= 1) {
            a01 += shfl_xor_u64(a01, m);
            a23 += shfl_xor_u64(a23, m);
            add += shfl_xor_u64(add, m);
        }
        if (threadIdx.x == 0) {
            s_pre[0] = bc.sbp[0 * bc.nsb + sb] + a01;
            s_pre[1] = bc.sbp[1 * bc.nsb + sb] + a23;
            s_pre[2] = bc.sbp[2 * bc.nsb + sb] + add;
        }
    }
    // the array and row ranges the count pass used (light/heavy split)
    const bool heavy = pmid && info->heavy;
    uint8_t *flag = heavy ? hflag : flag_a;
    const bool mark = pmid && !heavy;  // light round: relaxed vertices turn heavy-pending
    const uint64_t base = ((uint64_t)blockIdx.x * kBlock + threadIdx.x) * (16 * FV);
    uint4 w[FV];
    load_flags(flag, base, Vpad, w);
#pragma unroll
    for (int k = 0; k < FV; ++k) {
        s_win[(4 * k + 0) * kBlock + threadIdx.x] = w[k].x;
        s_win[(4 * k + 1) * kBlock + threadIdx.x] = w[k].y;
        s_win[(4 * k + 2) * kBlock + threadIdx.x] = w[k].z;
        s_win[(4 * k + 3) * kBlock + threadIdx.x] = w[k].w;
    }
    // bin of slot j (codes of pass 1; each thread reads its own words)
    auto bin_at = [&](uint32_t j) -> int {
        return (int)((s_win[(j >> 2) * kBlock + threadIdx.x] >> (8 * (j & 3))) & 0xFFu) - kCode0;
    };
    const uint32_t all = window_mask(w, kCode0);
    uint32_t c0 = 0, c1 = 0, c2 = 0, c3 = 0;
    uint64_t ddeg = 0;
    // walks the set bits kLoadBatch at a time; degrees read for bins 2 / 3
    auto walk = [&](auto &&use) {
        for (uint32_t m = all; m;) {
            uint32_t jj[kLoadBatch];
            int bb[kLoadBatch];
            uint64_t dg[kLoadBatch];
#pragma unroll
            for (int k = 0; k < kLoadBatch; ++k) {
                jj[k] = (16 * FV);
                bb[k] = -1;
                dg[k] = 0;
                if (m) {
                    const uint32_t j = (uint32_t)__builtin_ctz(m);
                    m &= m - 1;
                    jj[k] = j;
                    bb[k] = bin_at(j);
                    if (bb[k] == 2 || bb[k] == 3) dg[k] = round_deg(deg, hdeg, ptr, pmid, heavy, base + j);
                }
            }
#pragma unroll
            for (int k = 0; k < kLoadBatch; ++k)
                if (jj[k] != (16 * FV)) use(jj[k], bb[k], dg[k]);
        }
    };
    walk([&](uint32_t, int b, uint64_t deg) {
        switch (b) {
            case 0: ++c0; break;
            case 1: ++c1; break;
            case 2: c2 += (uint32_t)((deg + tn.tile - 1) / tn.tile); break;
            case 3: ++c3; ddeg += deg; break;
            default: break;
        }
    });
    const uint64_t x[3] = {(uint64_t)c0 | ((uint64_t)c1 << 32), (uint64_t)c2 | ((uint64_t)c3 << 32), ddeg};
    uint64_t ex[3], t[3];
    block_scan_n<3>(x, ex, t, lds);
    if (!all) return;
    uint32_t keep = 0;  // slots that stay flagged: held back (bin 5) or a held dynamic list
    if (!info->pull) {  // bottom-up round: no lists, the frontier is its level
        const uint32_t n0 = info->n[0], n1 = info->n[1];
        uint64_t p0 = (uint64_t)(uint32_t)s_pre[0] + (uint32_t)ex[0];
        uint64_t p1 = (uint64_t)n0 + (uint32_t)(s_pre[0] >> 32) + (uint32_t)(ex[0] >> 32);
        uint64_t p2 = (uint64_t)(uint32_t)s_pre[1] + (uint32_t)ex[1];
        uint64_t p3 = (uint64_t)(uint32_t)(s_pre[1] >> 32) + (uint32_t)(ex[1] >> 32);
        uint64_t pd = s_pre[2] + ex[2];
        const uint64_t dyn_base = (uint64_t)n0 + n1;
        const bool held = info->dyn_held != 0;
        walk([&](uint32_t j, int b, uint64_t deg) {
            const uint64_t v = base + j;
            if (mark && b <= 4 && !(b == 3 && held)) hflag[v] = 1;
            switch (b) {
                case 0: list[p0++] = (uint32_t)v; break;
                case 1: list[p1++] = (uint32_t)v; break;
                case 2: {
                    const uint32_t nt = (uint32_t)((deg + tn.tile - 1) / tn.tile);
                    for (uint32_t q = 0; q < nt; ++q, ++p2)
                        if (p2 < hub_cap) hub[p2] = ((uint64_t)q << 32) | v;
                    break;
                }
                case 3:
                    if (held) {
                        keep |= 1u << j;
                        break;
                    }
                    list[dyn_base + p3] = (uint32_t)v;
                    doff[p3] = pd;
                    // the dynamic relax's chunks of kDynEdgesPerBlock edges
                    // whose first edge is one of v's: they start at entry p3
                    for (uint64_t c = (pd + kDynEdgesPerBlock - 1) / kDynEdgesPerBlock;
                         c * kDynEdgesPerBlock < pd + deg && c < dchunk_cap; ++c)
                        dchunk[c] = (uint32_t)p3;
                    if (m_dlist) {  // pinned host mirrors for the host-side gather
                        m_dlist[p3] = (uint32_t)v;
                        m_doff[p3] = pd;
                    }
                    ++p3;
                    pd += deg;
                    break;
                case 5: keep |= 1u << j; break;
                default: break;
            }
        });
    }
    uint32_t word[4 * FV];
#pragma unroll
    for (int k = 0; k < 4 * FV; ++k) word[k] = 0;
#pragma unroll
    for (int j = 0; j < (16 * FV); ++j) word[j >> 2] |= ((keep >> j) & 1u) << (8 * (j & 3));
    store_flags<FV>(flag, base, Vpad, word);
}

__global__ void k_ctl_reset(RoundCtl *ctl, uint32_t delta, uint32_t defer, uint64_t dir_opt_edges, uint64_t V,
                            uint32_t split) {
    RoundCtl c{};
    c.split = split;
    c.delta = delta;
    c.defer = defer;
    c.dir_opt = dir_opt_edges ? 1u : 0u;
    c.unexplored = dir_opt_edges;
    c.nverts = V;
    c.thresh = delta ? (uint32_t)min((uint64_t)0xFFFFFFFFull, 1ull + delta) : 0xFFFFFFFFu;
    c.fbound = 0;
    c.live = 1;
    c.halt = 1;
    *ctl = c;
}

// ---------------------------------------------------------------- relax

template <int ALGO>
__device__ __forceinline__ uint32_t offer(uint32_t du, uint32_t wt) {
    if (ALGO == ALGO_BFS) return du + 1;
    if (ALGO == ALGO_SSSP) {
        // distance + 1 saturates at 0xFFFFFFFE (0xFFFFFFFF means unreached):
        // min() is monotone, so the fixed point is min(dist + 1, 0xFFFFFFFE)
        const uint64_t s = (uint64_t)du + wt;
        return s > 0xFFFFFFFEull ? 0xFFFFFFFEu : (uint32_t)s;
    }
    return du;
}

// What a frontier vertex carries to its out-edges: its value, or for the
// multi-source BFS the mask of sources whose frontier holds it.
template <int ALGO> struct SrcOf { using T = uint32_t; };
template <> struct SrcOf<ALGO_MSBFS> { using T = uint64_t; };

template <int ALGO>
__device__ __forceinline__ typename SrcOf<ALGO>::T load_src(const RelaxArgs &a, uint32_t u) {
    if constexpr (ALGO == ALGO_MSBFS) return a.pull ? a.front[u] & ~*a.pull : a.front[u];
    else return a.val[u];
}

// Relaxing an edge (u -> to, wt) is split in two so that a lane can keep the
// reads of several edges in flight before its first atomic: peek() reads the
// target's state, commit() compares u's offer with what peek saw and applies
// it.  Values only decrease (masks only grow), so a stale peek can cost an
// atomic that finds nothing to do but never loses an update.
template <int ALGO>
__device__ __forceinline__ typename SrcOf<ALGO>::T peek(const RelaxArgs &a, uint32_t to) {
    if constexpr (ALGO == ALGO_MSBFS) return a.visit[to];
    else return a.val[to];
}

// Returns 1 when `to` improved (and is flagged for the next frontier).
// BFS writes the same level from every writer in an iteration, so a plain
// store is race-free (a visited bitmap claimed with atomicOr, one writer per
// vertex, was measured slower on R-MAT 20: 14.7 vs 17.2 GTEPS); SSSP/CC use atomicMin so no smaller offer is lost
// (the reference's check-then-write in K7/K11 can lose one).  Multi-source:
// the sources of u's frontier that have not reached `to` are OR-ed into
// next[to] (one 64-bit atomic for all 64 searches).
template <int ALGO>
__device__ __forceinline__ uint32_t commit(const RelaxArgs &a, typename SrcOf<ALGO>::T s, uint32_t to, uint32_t wt,
                                           typename SrcOf<ALGO>::T seen) {
    if constexpr (ALGO == ALGO_MSBFS) {
        const uint64_t m = s & ~seen;
        if (!m) return 0;
        atomicOr((unsigned long long *)(a.next + to), (unsigned long long)m);
        a.flag[to] = 1;
        return 1;
    } else {
        const uint32_t nd = offer<ALGO>(s, wt);
        if (nd >= seen) return 0;
        if (ALGO == ALGO_BFS) {
            a.val[to] = nd;
            a.flag[to] = 1;
            return 1;
        }
        const uint32_t old = atomicMin(a.val + to, nd);
        if (nd < old) {
            a.flag[to] = 1;
            return 1;
        }
        return 0;
    }
}

// The relax kernels read their work counts from the round's FrontierInfo on
// the device and stride over them, so a launch needs no host-side count: a
// round can be enqueued (or captured in a graph) before its compaction ran.

// Bins 0 and 1: LPV lanes per frontier vertex (LPV = 64: one wave).  BIN
// selects the list segment: bin 0 at list[0, n0), bin 1 at list[n0, n0 + n1).
// The bodies take the workgroup's index and the number of workgroups of its
// kind (bid, nb), so that one launch can run several kinds side by side.
template <int ALGO, bool W, bool PK, int LPV>
__device__ __forceinline__ uint32_t relax_group_body(const uint32_t *__restrict__ list, int bin,
                                                     const FrontierInfo *__restrict__ info,
                                                     const uint64_t *__restrict__ ptr, const void *__restrict__ edges,
                                                     const RelaxArgs &a, uint32_t bid, uint32_t nb) {
    if (info->pull) return 0;
    const bool heavy = info->heavy != 0;
    const uint32_t n = info->n[bin];
    const uint32_t *lst = list + (bin ? info->n[0] : 0u);
    const uint32_t sub = threadIdx.x % LPV;
    const uint64_t stride = (uint64_t)nb * (kBlock / LPV);
    uint32_t cnt = 0;
    for (uint64_t item = ((uint64_t)bid * kBlock + threadIdx.x) / LPV; item < n; item += stride) {
        const uint32_t u = lst[item];
        uint64_t b, e;
        row_range(ptr, a.pmid, heavy, u, b, e);
        const auto s = load_src<ALGO>(a, u);
        if (ALGO == ALGO_MSBFS && !s) continue;  // all of u's sources go bottom-up this round
        for (uint64_t i = b + sub; i < e; i += (uint64_t)kRelaxBatch * LPV) {
            uint32_t to[kRelaxBatch], wt[kRelaxBatch];
            typename SrcOf<ALGO>::T seen[kRelaxBatch];
#pragma unroll
            for (int k = 0; k < kRelaxBatch; ++k) {
                to[k] = 0;
                wt[k] = 0;
                if (i + (uint64_t)k * LPV < e) load_edge<W, PK>(edges, i + (uint64_t)k * LPV, a.wbits, to[k], wt[k]);
            }
#pragma unroll
            for (int k = 0; k < kRelaxBatch; ++k)
                if (i + (uint64_t)k * LPV < e) seen[k] = peek<ALGO>(a, to[k]);
#pragma unroll
            for (int k = 0; k < kRelaxBatch; ++k)
                if (i + (uint64_t)k * LPV < e) cnt += commit<ALGO>(a, s, to[k], wt[k], seen[k]);
        }
    }
    return cnt;
}

template <int ALGO, bool W, bool PK, int LPV>
__global__ __launch_bounds__(kBlock) void k_relax_group(const uint32_t *__restrict__ list, int bin,
                                                        const FrontierInfo *__restrict__ info,
                                                        const uint64_t *__restrict__ ptr,
                                                        const void *__restrict__ edges, RelaxArgs a,
                                                        unsigned long long *upd) {
    wave_add(upd, relax_group_body<ALGO, W, PK, LPV>(list, bin, info, ptr, edges, a, blockIdx.x, gridDim.x));
}

template <int ALGO, bool W, bool PK>
__device__ __forceinline__ uint32_t relax_hub_body(const uint64_t *__restrict__ hub, uint64_t hub_cap,
                                                   const FrontierInfo *__restrict__ info, uint32_t tile,
                                                   const uint64_t *__restrict__ ptr, const void *__restrict__ edges,
                                                   const RelaxArgs &a, uint32_t *s_edges, uint32_t bid, uint32_t nb) {
    constexpr bool WIDE = W && !PK;  // 8-byte edges
    const uint64_t ntiles = info->pull ? 0 : min((uint64_t)info->n[2], hub_cap);
    const bool heavy = info->heavy != 0;
    uint32_t c = 0;
    for (uint64_t ti = bid; ti < ntiles; ti += nb) {
        const uint64_t it = hub[ti];
        const uint32_t u = (uint32_t)it, t = (uint32_t)(it >> 32);
        uint64_t rb, re;
        row_range(ptr, a.pmid, heavy, u, rb, re);
        const uint64_t b = rb + (uint64_t)t * tile;
        const uint64_t e = min(re, b + tile);
        const uint32_t cnt = (uint32_t)(e - b);
        const auto s = load_src<ALGO>(a, u);
        if (ALGO == ALGO_MSBFS && !s) continue;  // block-uniform: u is the same for every thread
        for (uint32_t k = threadIdx.x; k < cnt; k += kBlock) {
            if (WIDE)
                ((uint2 *)s_edges)[k] = ((const uint2 *)edges)[b + k];
            else
                s_edges[k] = ((const uint32_t *)edges)[b + k];
        }
        __syncthreads();
        for (uint32_t k0 = threadIdx.x; k0 < cnt; k0 += kRelaxBatch * kBlock) {
            uint32_t to[kRelaxBatch], wt[kRelaxBatch];
            typename SrcOf<ALGO>::T seen[kRelaxBatch];
#pragma unroll
            for (int j = 0; j < kRelaxBatch; ++j) {
                to[j] = 0;
                wt[j] = 0;
                if (k0 + j * kBlock < cnt) load_edge<W, PK>(s_edges, k0 + j * kBlock, a.wbits, to[j], wt[j]);
            }
#pragma unroll
            for (int j = 0; j < kRelaxBatch; ++j)
                if (k0 + j * kBlock < cnt) seen[j] = peek<ALGO>(a, to[j]);
#pragma unroll
            for (int j = 0; j < kRelaxBatch; ++j)
                if (k0 + j * kBlock < cnt) c += commit<ALGO>(a, s, to[j], wt[j], seen[j]);
        }
        __syncthreads();  // the tile buffer is reused by the next tile
    }
    return c;
}

// Bin 2: one workgroup per tile of a hub's neighbour list (workgroups stride
// over the tiles); the tile is staged in LDS with coalesced loads before the
// relax pass.
template <int ALGO, bool W, bool PK>
__global__ __launch_bounds__(kBlock) void k_relax_hub(const uint64_t *__restrict__ hub, uint64_t hub_cap,
                                                      const FrontierInfo *__restrict__ info, uint32_t tile,
                                                      const uint64_t *__restrict__ ptr,
                                                      const void *__restrict__ edges, RelaxArgs a,
                                                      unsigned long long *upd) {
    extern __shared__ uint32_t s_edges[];
    wave_add(upd, relax_hub_body<ALGO, W, PK>(hub, hub_cap, info, tile, ptr, edges, a, s_edges, blockIdx.x, gridDim.x));
}

// Dynamic region, edges [c0, min(c1, D)) of the concatenated active dynamic
// edge space (D = sum of the dynamic list's degrees; the list sits at
// list[n0 + n1, n0 + n1 + n3)).  STAGED (DIRECT = false): they sit in a
// staging slot, edge q at q - c0.  DIRECT: edge q of vertex u is CSR edge
// x = ptr[u] + (q - doff[i]); when u's chunk of kResChunk vertices is
// resident in HBM (rdelta[u >> kResChunkShift] != kNotResident) it is read
// from `resident` at x + rdelta, else in place from the pinned host array
// through its device mapping (`stage`, at x - dbase); reads served from HBM
// are counted in *hits.  (Storing the host rows in
// order of distance from a hub instead of vertex order was measured slower
// on R-MAT: the active list stays in vertex order, so neighbouring lanes then
// read scattered rows.)
// Each workgroup takes chunks of kDynEdgesPerBlock consecutive edges, loads
// the covering slice of the degree prefix into LDS and maps each edge to its
// source with a binary search there (every listed vertex has >= 1 edge, so a
// chunk covers at most kDynEdgesPerBlock vertices).  With the compaction's
// chunk table (dchunk, zero-copy rounds: chunks start at edge 0) the slice's
// first and last entries are two loads; the staged loader's slots start
// anywhere, so it finds them by binary search over the degree prefix.
// LDS of the dynamic relax body: the degree-prefix slice and its bounds
struct DynLds {
    uint64_t off[kDynEdgesPerBlock + 2];
    uint32_t i[2];
};

template <int ALGO, bool W, bool PACKED, bool DIRECT>
__device__ __forceinline__ uint32_t relax_dyn_body(const uint32_t *__restrict__ list,
                                                   const FrontierInfo *__restrict__ info,
                                                   const uint64_t *__restrict__ doff, const void *__restrict__ stage,
                                                   uint64_t c0, uint64_t c1max, const uint64_t *__restrict__ ptr,
                                                   uint64_t dbase, const void *__restrict__ resident,
                                                   const int64_t *__restrict__ rdelta, uint32_t wbits,
                                                   const RelaxArgs &a, const uint64_t *__restrict__ lpos,
                                                   uint64_t nl, uint64_t vsplit, const uint32_t *__restrict__ dchunk,
                                                   DynLds &L, uint32_t bid, uint32_t nb, uint32_t &hits) {
    uint64_t *s_off = L.off;
    uint32_t *s_i = L.i;
    const uint32_t ndyn = info->n[3];
    const uint32_t *dlist = list + info->n[0] + info->n[1];
    const uint64_t c1 = min(c1max, info->dyn_deg);
    const bool heavy = info->heavy != 0;
    uint32_t c = 0, h = 0;
    for (uint64_t q0 = c0 + (uint64_t)bid * kDynEdgesPerBlock; q0 < c1; q0 += (uint64_t)nb * kDynEdgesPerBlock) {
        const uint64_t q1 = min(q0 + kDynEdgesPerBlock, c1);
        if (dchunk) {  // chunks aligned to the table (c0 = 0): two loads
            if (threadIdx.x < 2) {
                const uint64_t k = q0 / kDynEdgesPerBlock + threadIdx.x;
                // the entry holding q1 (or the last entry): one past the chunk's last edge at most
                s_i[threadIdx.x] = threadIdx.x == 1 && q1 >= info->dyn_deg ? ndyn - 1 : dchunk[k];
            }
        } else if (threadIdx.x < 2) {
            const uint64_t q = threadIdx.x == 0 ? q0 : q1 - 1;
            uint32_t lo = 0, hi = ndyn;  // largest i with doff[i] <= q; doff[0] = 0
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (doff[mid] <= q) lo = mid; else hi = mid;
            }
            s_i[threadIdx.x] = lo;
        }
        __syncthreads();
        const uint32_t i0 = s_i[0];
        const uint32_t nv = s_i[1] - i0 + 1;
        for (uint32_t k = threadIdx.x; k <= nv; k += kBlock) s_off[k] = doff[i0 + k];
        __syncthreads();
        // all edge (and source-state) reads of the lane first, then the
        // target peeks, then the commits: the link / HBM latencies overlap
        uint32_t to[kDynEdgesPerThread], wt[kDynEdgesPerThread];
        typename SrcOf<ALGO>::T sv[kDynEdgesPerThread], seen[kDynEdgesPerThread];
#pragma unroll
        for (int j = 0; j < kDynEdgesPerThread; ++j) {
            const uint64_t q = q0 + threadIdx.x + (uint64_t)j * kBlock;
            to[j] = 0;
            wt[j] = 0;
            sv[j] = 0;
            if (q < q1) {
                uint32_t lo = 0, hi = nv;
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (s_off[mid] <= q) lo = mid; else hi = mid;
                }
                const uint32_t u = dlist[i0 + lo];
                sv[j] = load_src<ALGO>(a, u);
                if (DIRECT) {
                    // hp: position of the edge in the host copy of the region
                    const uint64_t k = q - s_off[lo];
                    uint64_t hp;
                    uint64_t slot = u >> kResChunkShift;  // entry of the resident table
                    if (lpos) {  // split layout: light parts, then heavy parts
                        const uint64_t lp = lpos[u - vsplit];
                        const uint64_t hb = nl + (ptr[u] - dbase) - lp;
                        bool hv;
                        if (a.pmid) {  // a split round reads one part
                            hv = heavy;
                            hp = (hv ? hb : lp) + k;
                        } else {  // whole rows: the light part, then the heavy part
                            const uint64_t nlu = lpos[u - vsplit + 1] - lp;
                            hv = k >= nlu;
                            hp = hv ? hb + (k - nlu) : lp + k;
                        }
                        slot = 2 * slot + (hv ? 1 : 0);
                    } else {
                        hp = ((a.pmid && heavy) ? a.pmid[u] : ptr[u]) - dbase + k;
                    }
                    const int64_t dl = rdelta ? rdelta[slot] : kNotResident;
                    if (dl != kNotResident) {
                        load_edge<W, PACKED>(resident, (uint64_t)((int64_t)hp + dl), wbits, to[j], wt[j]);
                        ++h;
                    } else {
                        load_edge<W, PACKED>(stage, hp, wbits, to[j], wt[j]);
                    }
                } else {
                    load_edge<W, PACKED>(stage, q - c0, wbits, to[j], wt[j]);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < kDynEdgesPerThread; ++j)
            if (q0 + threadIdx.x + (uint64_t)j * kBlock < q1) seen[j] = peek<ALGO>(a, to[j]);
#pragma unroll
        for (int j = 0; j < kDynEdgesPerThread; ++j)
            if (q0 + threadIdx.x + (uint64_t)j * kBlock < q1) c += commit<ALGO>(a, sv[j], to[j], wt[j], seen[j]);
        __syncthreads();  // s_off / s_i are reloaded for the next chunk
    }
    hits = h;
    return c;
}

template <int ALGO, bool W, bool PACKED, bool DIRECT>
__global__ __launch_bounds__(kBlock) void k_relax_dyn(const uint32_t *__restrict__ list,
                                                      const FrontierInfo *__restrict__ info,
                                                      const uint64_t *__restrict__ doff,
                                                      const void *__restrict__ stage, uint64_t c0, uint64_t c1max,
                                                      const uint64_t *__restrict__ ptr, uint64_t dbase,
                                                      const void *__restrict__ resident,
                                                      const int64_t *__restrict__ rdelta,
                                                      uint32_t wbits, RelaxArgs a, const uint64_t *__restrict__ lpos,
                                                      uint64_t nl, uint64_t vsplit, const uint32_t *__restrict__ dchunk,
                                                      unsigned long long *upd, unsigned long long *hits) {
    __shared__ DynLds L;
    uint32_t h = 0;
    const uint32_t c = relax_dyn_body<ALGO, W, PACKED, DIRECT>(list, info, doff, stage, c0, c1max, ptr, dbase, resident,
                                                               rdelta, wbits, a, lpos, nl, vsplit, dchunk, L,
                                                               blockIdx.x, gridDim.x, h);
    wave_add(upd, c);
    if (DIRECT) wave_add(hits, h);
}

// One round's relax in a single launch: static bins 0 and 1, hub tiles and
// the zero-copy dynamic region side by side, so a device-driven round needs
// no second stream (the fork and join of which cost tens of microseconds per
// round in a replayed graph, most of a small round).  A workgroup takes its
// kind from blockIdx.x modulo a period of kFuseStatic + cd workgroups (2
// bin 0, 2 bin 1, 1 hub tile, cd dynamic): the kinds interleave in dispatch
// order, so the link-bound dynamic relax and the static bins progress
// together as the two streams did.  With K = gridDim.x / period, the kinds
// get 2K, 2K, K and cd * K workgroups.  A graph wholly in HBM runs it with
// cd = 0: its three static bins in one launch instead of three.
struct RoundRelax {
    const uint32_t *list;
    const FrontierInfo *info;
    const uint64_t *ptr;
    const void *sedges;         // static region
    const uint64_t *hub;
    uint64_t hub_cap;
    uint32_t tile;
    uint32_t cd;                // dynamic workgroups per period
    const uint64_t *doff;
    const void *stage;          // dynamic region in pinned host memory (device address)
    uint64_t dbase;
    const void *resident;       // resident chunks in HBM
    const int64_t *rdelta;
    const uint64_t *lpos;       // split layout of the host copy (null: rows whole)
    const uint32_t *dchunk;     // chunk table of the dynamic edge space
    uint64_t nl, vsplit;
    unsigned long long *counters;  // [0] static updates, [1] dynamic updates, [5] HBM-served dynamic edges
};
constexpr uint32_t kFuseStatic = 5;

template <int ALGO, bool W, bool PK, int LPV>
__global__ __launch_bounds__(kBlock) void k_relax_round(RoundRelax r, RelaxArgs a) {
    extern __shared__ uint64_t s_fuse[];
    const uint32_t P = kFuseStatic + r.cd;
    const uint32_t K = gridDim.x / P;
    const uint32_t k = blockIdx.x / P, j = blockIdx.x - k * P;
    if (j < 2) {
        wave_add(r.counters + 0,
                 relax_group_body<ALGO, W, PK, LPV>(r.list, 0, r.info, r.ptr, r.sedges, a, 2 * k + j, 2 * K));
    } else if (j < 4) {
        wave_add(r.counters + 0,
                 relax_group_body<ALGO, W, PK, 64>(r.list, 1, r.info, r.ptr, r.sedges, a, 2 * k + j - 2, 2 * K));
    } else if (j == 4) {
        wave_add(r.counters + 0, relax_hub_body<ALGO, W, PK>(r.hub, r.hub_cap, r.info, r.tile, r.ptr, r.sedges, a,
                                                              (uint32_t *)s_fuse, k, K));
    } else {
        uint32_t h = 0;
        const uint32_t c = relax_dyn_body<ALGO, W, PK, true>(r.list, r.info, r.doff, r.stage, 0, ~0ull, r.ptr, r.dbase,
                                                             r.resident, r.rdelta, a.wbits, a, r.lpos, r.nl, r.vsplit,
                                                             r.dchunk, *(DynLds *)s_fuse, r.cd * k + (j - kFuseStatic),
                                                             r.cd * K, h);
        wave_add(r.counters + 1, c);
        wave_add(r.counters + 5, h);
    }
}

// Served BFS (the query server's 64-slot engines): bottom-up scan for the
// pulling slots P = *pull.  Every vertex that some slot of P has not reached
// scans its in-edges, ORing the frontier masks of its in-neighbours
// (restricted to those slots), and stops once every such slot is found.
// Found slots are offered through next[v] like a top-down relax.
__device__ __forceinline__ void srv_pull_body(const SrvPull &p, uint32_t bid, uint32_t nb) {
    const uint64_t P = *p.pull;
    if (!P) return;
    for (uint64_t v = (uint64_t)bid * kBlock + threadIdx.x; v < p.V; v += (uint64_t)nb * kBlock) {
        const uint64_t need = P & ~p.visit[v];
        if (!need) continue;
        uint64_t acc = 0;
        const uint64_t b = p.in_ptr[v], e = p.in_ptr[v + 1];
        for (uint64_t i = b; i < e && acc != need; i += kRelaxBatch) {
            uint32_t u[kRelaxBatch];
#pragma unroll
            for (int k = 0; k < kRelaxBatch; ++k) u[k] = i + k < e ? p.in_cols[i + k] : 0u;
#pragma unroll
            for (int k = 0; k < kRelaxBatch; ++k)
                if (i + k < e) acc |= p.front[u[k]] & need;
        }
        if (acc) {
            atomicOr((unsigned long long *)(p.next + v), (unsigned long long)acc);
            p.flag[v] = 1;
        }
    }
}

// A served BFS round's relax in one launch: the top-down slots through the
// static bins (as k_relax_round with cd = 0) and, beside them, the bottom-up
// scan of the pulling slots (cp workgroups per period of kFuseStatic + cp).
// The two only read visit / front and only OR into next and set flags, so
// they may run side by side.
template <bool W, bool PK, int LPV>
__global__ __launch_bounds__(kBlock) void k_srv_round(RoundRelax r, RelaxArgs a, SrvPull p, uint32_t cp) {
    extern __shared__ uint64_t s_fuse[];
    const uint32_t P = kFuseStatic + cp;
    const uint32_t K = gridDim.x / P;
    const uint32_t k = blockIdx.x / P, j = blockIdx.x - k * P;
    if (j < 2) {
        wave_add(r.counters + 0, relax_group_body<ALGO_MSBFS, W, PK, LPV>(r.list, 0, r.info, r.ptr, r.sedges, a,
                                                                         2 * k + j, 2 * K));
    } else if (j < 4) {
        wave_add(r.counters + 0, relax_group_body<ALGO_MSBFS, W, PK, 64>(r.list, 1, r.info, r.ptr, r.sedges, a,
                                                                        2 * k + j - 2, 2 * K));
    } else if (j == 4) {
        wave_add(r.counters + 0, relax_hub_body<ALGO_MSBFS, W, PK>(r.hub, r.hub_cap, r.info, r.tile, r.ptr, r.sedges,
                                                                    a, (uint32_t *)s_fuse, k, K));
    } else {
        srv_pull_body(p, cp * k + (j - kFuseStatic), cp * K);
    }
}

// Bottom-up BFS round (direction-optimising BFS, Beamer et al., SC'12): the
// frontier is the current level L = rounds so far; every unvisited vertex
// scans its in-edges, kRelaxBatch parents' levels at a time, and takes level
// L + 1 at the first parent on level L.  Each vertex is written by its own
// lane only, and a fresh L + 1 never reads as L, so no atomics are needed.
__global__ __launch_bounds__(kBlock) void k_bfs_pull(const uint64_t *__restrict__ in_ptr,
                                                     const uint32_t *__restrict__ in_cols, uint64_t V,
                                                     const FrontierInfo *__restrict__ info,
                                                     const RoundCtl *__restrict__ ctl, uint32_t *val,
                                                     uint8_t *__restrict__ flag, unsigned long long *upd) {
    if (!info->pull) return;
    const uint32_t L = ctl->rounds;
    uint32_t c = 0;
    for (uint64_t v = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v < V; v += (uint64_t)gridDim.x * kBlock) {
        if (val[v] != OG_UNREACHED) continue;
        const uint64_t b = in_ptr[v], e = in_ptr[v + 1];
        for (uint64_t i = b; i < e; i += kRelaxBatch) {
            uint32_t lv[kRelaxBatch];
#pragma unroll
            for (int k = 0; k < kRelaxBatch; ++k) lv[k] = i + k < e ? val[in_cols[i + k]] : 0u;
            bool hit = false;
#pragma unroll
            for (int k = 0; k < kRelaxBatch; ++k) hit |= lv[k] == L;
            if (hit) {
                val[v] = L + 1;
                flag[v] = 1;
                ++c;
                break;
            }
        }
    }
    wave_add(upd, c);
}

// Order-independent digest of a result array: the sum (mod 2^64) of
// mix64(v << 32 | value[v]) over reached vertices.
__device__ __forceinline__ uint64_t d_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// reached = #{v : val[v] != UNREACHED}, traversed = sum of their degrees,
// checksum = digest of val.
// deg / hdeg: the degree table (hdeg: heavy degrees with the split, else
// null), or null for the row bounds.  With the table, each lane reads four
// values and four degrees per 16-byte load.
__global__ __launch_bounds__(kBlock) void k_reach_stats(const uint32_t *__restrict__ val,
                                                        const uint64_t *__restrict__ ptr,
                                                        const uint32_t *__restrict__ deg,
                                                        const uint32_t *__restrict__ hdeg, uint64_t V,
                                                        unsigned long long *out) {
    __shared__ uint64_t lds[kBlock / kWave];
    uint64_t r = 0, t = 0, h = 0;
    const uint64_t tid = (uint64_t)blockIdx.x * kBlock + threadIdx.x, nth = (uint64_t)gridDim.x * kBlock;
    auto add = [&](uint64_t v, uint32_t x, uint64_t dv) {
        if (x != OG_UNREACHED) {
            ++r;
            t += dv;
            h += d_mix64((v << 32) | x);
        }
    };
    uint64_t v0 = 0;  // first vertex of the scalar tail
    if (deg) {
        // val and the table are 16-byte aligned allocations; the table is
        // padded to a multiple of 16 entries
        const uint64_t n4 = V / 4;
        // two chunks per lane per step, their loads issued together
        for (uint64_t q0 = tid; q0 < n4; q0 += 2 * nth) {
            uint4 x[2], dl[2], dh[2];
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const uint64_t q = q0 + k * nth;
                const bool in = q < n4;
                x[k] = in ? ((const uint4 *)val)[q] : make_uint4(OG_UNREACHED, OG_UNREACHED, OG_UNREACHED, OG_UNREACHED);
                dl[k] = in ? ((const uint4 *)deg)[q] : make_uint4(0, 0, 0, 0);
                dh[k] = in && hdeg ? ((const uint4 *)hdeg)[q] : make_uint4(0, 0, 0, 0);
            }
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const uint64_t q = q0 + k * nth;
                add(4 * q + 0, x[k].x, (uint64_t)dl[k].x + dh[k].x);
                add(4 * q + 1, x[k].y, (uint64_t)dl[k].y + dh[k].y);
                add(4 * q + 2, x[k].z, (uint64_t)dl[k].z + dh[k].z);
                add(4 * q + 3, x[k].w, (uint64_t)dl[k].w + dh[k].w);
            }
        }
        v0 = 4 * n4;
    }
    for (uint64_t v = v0 + tid; v < V; v += nth) {
        const uint32_t x = val[v];
        add(v, x, deg ? (uint64_t)deg[v] + (hdeg ? hdeg[v] : 0u) : (x != OG_UNREACHED ? ptr[v + 1] - ptr[v] : 0));
    }
    r = block_sum(r, lds);
    t = block_sum(t, lds);
    h = block_sum(h, lds);
    if (threadIdx.x == 0) {
        atomicAdd(out + 0, (unsigned long long)r);
        atomicAdd(out + 1, (unsigned long long)t);
        atomicAdd(out + 2, (unsigned long long)h);
    }
}

// ------------------------------------------------------- multi-source BFS

// Seeds up to 64 sources: bit j of visit/front at source j, level 1.
__global__ void k_ms_init(const uint32_t *__restrict__ src, uint32_t n, uint64_t *visit, uint64_t *front,
                          uint8_t *flag, uint8_t *lvl) {
    const uint32_t j = threadIdx.x;
    if (j < n) {
        const uint32_t s = src[j];
        atomicOr((unsigned long long *)(visit + s), 1ull << j);
        atomicOr((unsigned long long *)(front + s), 1ull << j);
        flag[s] = 1;
        lvl[(uint64_t)s * kMsWidth + j] = 1;
    }
}

// After a round: for every flagged vertex, the offered sources it had not
// seen become its new frontier mask and get the round's level (rounds with
// active vertices so far + 1, read from the device state; a level past 255
// sets ctl->overflow); vertices with nothing new leave the frontier.
__global__ __launch_bounds__(kBlock) void k_ms_commit(uint8_t *__restrict__ flag, uint64_t V, uint64_t *visit,
                                                      uint64_t *front, uint64_t *next, uint8_t *lvl, RoundCtl *ctl) {
    const uint32_t lv32 = ctl->rounds + 1;
    const uint8_t level = (uint8_t)min(lv32, 255u);
    for (uint64_t v = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v < V; v += (uint64_t)gridDim.x * kBlock) {
        if (!flag[v]) continue;
        const uint64_t vis = visit[v];
        const uint64_t nw = next[v] & ~vis;
        next[v] = 0;
        front[v] = nw;
        if (nw) {
            if (lv32 > 255) ctl->overflow = 1;
            visit[v] = vis | nw;
            uint64_t m = nw;
            while (m) {
                const int j = __builtin_ctzll(m);
                m &= m - 1;
                lvl[v * kMsWidth + j] = level;
            }
        } else {
            flag[v] = 0;
        }
    }
}

// Per-source reached count, traversed edges and digest over the vertex-major
// level array: lane j of each wave handles source j (one 64-byte line per
// vertex), partial sums leave with one atomic per lane per block.
__global__ __launch_bounds__(kBlock) void k_ms_digest(const uint8_t *__restrict__ lvl,
                                                      const uint64_t *__restrict__ ptr, uint64_t V,
                                                      unsigned long long *out /* 3 x 64 */) {
    const int lane = threadIdx.x & (kWave - 1);
    const uint64_t wave = ((uint64_t)blockIdx.x * kBlock + threadIdx.x) / kWave;
    const uint64_t nwaves = (uint64_t)gridDim.x * kBlock / kWave;
    uint64_t r = 0, t = 0, h = 0;
    for (uint64_t v = wave; v < V; v += nwaves) {
        const uint32_t x = lvl[v * kMsWidth + lane];
        if (x) {
            ++r;
            t += ptr[v + 1] - ptr[v];
            h += d_mix64((v << 32) | x);
        }
    }
    atomicAdd(out + lane, (unsigned long long)h);
    atomicAdd(out + kMsWidth + lane, (unsigned long long)r);
    atomicAdd(out + 2 * kMsWidth + lane, (unsigned long long)t);
}

// Levels of source j as u32 in the single-source convention.
__global__ void k_ms_extract(const uint8_t *__restrict__ lvl, uint64_t V, uint32_t j, uint32_t *out) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t x = lvl[v * kMsWidth + j];
        out[v] = x ? x : OG_UNREACHED;
    }
}

// ============================================================ host driver

static unsigned fill_grid(uint64_t n) { return (unsigned)std::min<uint64_t>(8192, std::max<uint64_t>(1, (n + 255) / 256)); }

void device_state_free(DeviceState *d) {
    if (!d) return;
    if (d->s_compute) (void)hipStreamSynchronize(d->s_compute);
    if (d->s_copy) (void)hipStreamSynchronize(d->s_copy);
    if (d->s_static) (void)hipStreamSynchronize(d->s_static);
    for (hipEvent_t e : d->ev_h2d) (void)hipEventDestroy(e);
    for (hipEvent_t e : d->ev_relax) (void)hipEventDestroy(e);
    d->mem.release();
    // (a lane runs everything on one stream: s_copy and s_static alias it)
    if (d->s_copy && d->s_copy != d->s_compute) (void)hipStreamDestroy(d->s_copy);
    if (d->s_static && d->s_static != d->s_compute) (void)hipStreamDestroy(d->s_static);
    if (d->s_compute) (void)hipStreamDestroy(d->s_compute);
    if (d->ev_static) (void)hipEventDestroy(d->ev_static);
    if (d->ev_fork) (void)hipEventDestroy(d->ev_fork);
    for (hipEvent_t e : d->ev_out)
        if (e) (void)hipEventDestroy(e);
    delete d;
}

constexpr size_t kOutChunk = 8u << 20;  // bytes per pinned staging chunk of a result copy-back

void copy_to_host(DeviceState *d, void *dst, const void *src, size_t n, hipStream_t s) {
    if (n <= kOutChunk) {  // one chunk: nothing to overlap
        OG_HIP(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, s));
        OG_HIP(hipStreamSynchronize(s));
        return;
    }
    if (!d->h_out[0]) {
        for (auto &p : d->h_out) p = d->palloc(kOutChunk);
        for (auto &e : d->ev_out) OG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    if (!d->pool) d->pool.reset(new WorkerPool(d->host_threads));
    const size_t nch = (n + kOutChunk - 1) / kOutChunk;
    auto issue = [&](size_t k) {
        const size_t off = k * kOutChunk, len = std::min(kOutChunk, n - off);
        OG_HIP(hipMemcpyAsync(d->h_out[k & 1], (const char *)src + off, len, hipMemcpyDeviceToHost, s));
        OG_HIP(hipEventRecord(d->ev_out[k & 1], s));
    };
    issue(0);
    for (size_t k = 0; k < nch; ++k) {
        // chunk k + 1 reuses the buffer chunk k - 1 left: its host copy ended
        // in the previous iteration
        if (k + 1 < nch) issue(k + 1);
        OG_HIP(hipEventSynchronize(d->ev_out[k & 1]));
        const size_t off = k * kOutChunk, len = std::min(kOutChunk, n - off);
        const char *from = (const char *)d->h_out[k & 1];
        char *to = (char *)dst + off;
        const unsigned T = d->pool->size();
        d->pool->run([&](unsigned t) {
            const size_t a = len * t / T, b = len * (t + 1) / T;
            std::memcpy(to + a, from + a, b - a);
        });
    }
}

// Largest number of hub tiles any frontier can produce for the current tuning.
static uint64_t hub_capacity(const og_graph *g, uint64_t vsplit, const Tuning &tn) {
    uint64_t cap = 0;
    for (uint64_t v = 0; v < vsplit; ++v) {
        uint64_t deg = g->ptr[v + 1] - g->ptr[v];
        if (deg > tn.t2) cap += (deg + tn.tile - 1) / tn.tile;
    }
    return cap;
}

void frontier_alloc(DeviceState *d, const og_graph *g, Frontier &f, MemPool &m) {
    const uint64_t V = d->V;
    f.flag = (uint8_t *)m.dalloc(d->Vpad);
    if (d->d_pmid) {
        f.hflag = (uint8_t *)m.dalloc(d->Vpad);
        OG_HIP(hipMemset(f.hflag, 0, d->Vpad));
    }
    f.list = (uint32_t *)m.dalloc(V * sizeof(uint32_t));
    f.doff = (uint64_t *)m.dalloc((V + 1) * sizeof(uint64_t));
    // chunk table of the dynamic relax: the dynamic edge space of a round is
    // at most the dynamic region's edges
    const uint64_t dyn_edges = d->plan.vsplit < V ? g->E - g->ptr[d->plan.vsplit] : 0;
    f.dchunk_cap = dyn_edges / kDynEdgesPerBlock + 2;
    f.dchunk = (uint32_t *)m.dalloc(f.dchunk_cap * sizeof(uint32_t));
    f.fv = (uint32_t)frontier_fv(V);
    f.nblk = (uint32_t)((d->Vpad + verts_per_block(f.fv) - 1) / verts_per_block(f.fv));
    f.V = V;
    f.bcnt = (uint32_t *)m.dalloc((uint64_t)kBcntRows * f.nblk * sizeof(uint32_t));
    f.bdeg = (uint64_t *)m.dalloc((uint64_t)f.nblk * sizeof(uint64_t));
    f.nsb = (f.nblk + kSbBlocks - 1) / kSbBlocks;
    f.sbt = (uint64_t *)m.dalloc((uint64_t)kSbVals * f.nsb * sizeof(uint64_t));
    f.sbp = (uint64_t *)m.dalloc(3ull * f.nsb * sizeof(uint64_t));
    OG_HIP(hipMemset(f.sbt, 0, (uint64_t)(kSbVals - 1) * f.nsb * sizeof(uint64_t)));
    OG_HIP(hipMemset(f.sbt + (uint64_t)(kSbVals - 1) * f.nsb, 0xFF, (uint64_t)f.nsb * sizeof(uint64_t)));
    f.info = (FrontierInfo *)m.dalloc(sizeof(FrontierInfo));
    f.ctl = (RoundCtl *)m.dalloc(sizeof(RoundCtl));
    f.counters = (unsigned long long *)m.dalloc(16 * sizeof(unsigned long long));
    f.h_info = (FrontierInfo *)m.palloc(sizeof(FrontierInfo));
    f.h_ctl = (RoundCtl *)m.palloc(sizeof(RoundCtl));
    f.h_counters = (unsigned long long *)m.palloc(16 * sizeof(unsigned long long));
    const uint64_t cap = std::max<uint64_t>(hub_capacity(g, d->plan.vsplit, d->tune), 1);
    f.hub = (uint64_t *)m.dalloc(cap * sizeof(uint64_t));
    f.hub_cap = cap;
    OG_HIP(hipMemset(f.flag, 0, d->Vpad));
    RoundCtl c{};
    c.thresh = 0xFFFFFFFFu;
    OG_HIP(hipMemcpy(f.ctl, &c, sizeof c, hipMemcpyHostToDevice));
}

void frontier_begin(Frontier &f, hipStream_t s, uint32_t delta, uint32_t defer, uint64_t dir_opt_edges,
                    bool split) {
    hipLaunchKernelGGL(k_ctl_reset, dim3(1), dim3(1), 0, s, f.ctl, delta, defer, dir_opt_edges, f.V,
                       split ? 1u : 0u);
    OG_HIP(hipGetLastError());
}

void frontier_compact(DeviceState *d, Frontier &f, hipStream_t s, const uint32_t *val, bool to_host,
                      const uint64_t *pmid) {
    // (pinned mirrors of the dynamic list written by pass 3 cost one small
    // PCIe write per vertex: slower than one bulk copy back, so unused)
    uint32_t *m_dlist = nullptr;
    uint64_t *m_doff = nullptr;
    // the degree table (a split round needs its light and heavy parts)
    const uint32_t *deg = d->d_deg && (!pmid || d->deg_split) ? d->d_deg : nullptr;
    const uint32_t *hdeg = deg ? d->deg_heavy() : nullptr;
    if (f.fv == 1)
        hipLaunchKernelGGL(k_frontier_count<1>, dim3(f.nblk), dim3(kBlock), 0, s, f.flag, f.hflag, d->Vpad,
                           d->plan.vsplit, d->d_ptr, pmid, deg, hdeg, d->tune, val, f.ctl, f.counts());
    else
        hipLaunchKernelGGL(k_frontier_count<2>, dim3(f.nblk), dim3(kBlock), 0, s, f.flag, f.hflag, d->Vpad,
                           d->plan.vsplit, d->d_ptr, pmid, deg, hdeg, d->tune, val, f.ctl, f.counts());
    // one lane per super-block up to kScanThreads lanes (one wave up to
    // 2^25 vertices)
    const unsigned scan_threads =
        (unsigned)std::min<uint64_t>(kScanThreads, (std::max<uint64_t>(f.nsb, 1) + kWave - 1) / kWave * kWave);
    hipLaunchKernelGGL(k_frontier_scan, dim3(1), dim3(scan_threads), 0, s, f.counts(), f.info, f.doff, f.ctl, m_doff);
    if (f.fv == 1)
        hipLaunchKernelGGL(k_frontier_write<1>, dim3(f.nblk), dim3(kBlock), 0, s, f.flag, f.hflag, d->Vpad, d->d_ptr,
                           pmid, deg, hdeg, d->tune, f.counts(), f.info, f.list, f.hub, f.hub_cap, f.doff, m_dlist,
                           m_doff, f.dchunk, f.dchunk_cap);
    else
        hipLaunchKernelGGL(k_frontier_write<2>, dim3(f.nblk), dim3(kBlock), 0, s, f.flag, f.hflag, d->Vpad, d->d_ptr,
                           pmid, deg, hdeg, d->tune, f.counts(), f.info, f.list, f.hub, f.hub_cap, f.doff, m_dlist,
                           m_doff, f.dchunk, f.dchunk_cap);
    OG_HIP(hipGetLastError());
    if (to_host) OG_HIP(hipMemcpyAsync(f.h_info, f.info, sizeof(FrontierInfo), hipMemcpyDeviceToHost, s));
}

// Bits needed by the largest edge weight (at least 1).
static uint32_t weight_bits(const og_graph *g) {
    std::atomic<uint32_t> maxw{0};
    parallel_for(g->E, 1 << 22, 0, [&](uint64_t b, uint64_t e, unsigned) {
        uint32_t m = 0;
        for (uint64_t i = b; i < e; ++i) m = std::max(m, g->wedges[2 * i + 1]);
        uint32_t cur = maxw.load();
        while (m > cur && !maxw.compare_exchange_weak(cur, m)) {
        }
    });
    uint32_t wb = 1;
    while (wb < 32 && (maxw.load() >> wb)) ++wb;
    return wb;
}

// Rows [v0, v1) of a weighted graph as packed words (to << wb | w) at
// dst[ptr[v] - ptr[v0] ...]; with the light/heavy split (sw > 0) each row
// lists its edges of weight < sw first, then the others, each part in CSR
// order.
static void pack_rows(const og_graph *g, uint64_t v0, uint64_t v1, uint32_t wb, uint32_t sw, uint32_t *dst) {
    const uint64_t e0 = g->ptr[v0];
    parallel_for(v1 - v0, 1 << 12, 0, [&](uint64_t b, uint64_t e, unsigned) {
        for (uint64_t v = v0 + b; v < v0 + e; ++v) {
            const uint32_t *src = g->wedges + 2 * g->ptr[v];
            const uint64_t n = g->ptr[v + 1] - g->ptr[v];
            uint32_t *out = dst + (g->ptr[v] - e0);
            if (!sw) {
                for (uint64_t i = 0; i < n; ++i) out[i] = (src[2 * i] << wb) | src[2 * i + 1];
                continue;
            }
            uint64_t k = 0;
            for (uint64_t i = 0; i < n; ++i)
                if (src[2 * i + 1] < sw) out[k++] = (src[2 * i] << wb) | src[2 * i + 1];
            for (uint64_t i = 0; i < n; ++i)
                if (src[2 * i + 1] >= sw) out[k++] = (src[2 * i] << wb) | src[2 * i + 1];
        }
    });
}

// Dynamic rows [v0, V) in the split layout: the light part of v at
// dst + lpos[v - v0], its heavy part at dst + nl + (ptr[v] - ptr[v0]) -
// lpos[v - v0] (nl = lpos[V - v0], all light parts first).
static void pack_split_rows(const og_graph *g, uint64_t v0, uint32_t wb, uint32_t sw, const uint64_t *lpos,
                            uint32_t *dst) {
    const uint64_t e0 = g->ptr[v0], nl = lpos[g->V - v0];
    parallel_for(g->V - v0, 1 << 12, 0, [&](uint64_t b, uint64_t e, unsigned) {
        for (uint64_t v = v0 + b; v < v0 + e; ++v) {
            const uint32_t *src = g->wedges + 2 * g->ptr[v];
            const uint64_t n = g->ptr[v + 1] - g->ptr[v], lp = lpos[v - v0];
            uint32_t *lo = dst + lp, *hi = dst + nl + (g->ptr[v] - e0) - lp;
            for (uint64_t i = 0; i < n; ++i) {
                const uint32_t w = (src[2 * i] << wb) | src[2 * i + 1];
                if (src[2 * i + 1] < sw) *lo++ = w;
                else *hi++ = w;
            }
        }
    });
}

// Uploads the packed rows [0, vend) into dst (device): host threads pack
// row-aligned chunks of about 64 MiB into two pinned bounce buffers while
// the previous chunk is in flight on s_compute.
static void upload_packed(DeviceState *d, const og_graph *g, uint64_t vend, uint32_t *dst) {
    const uint64_t n = g->ptr[vend];
    if (!n) return;
    struct Bounce {
        hipStream_t s;
        uint32_t *b[2] = {nullptr, nullptr};
        hipEvent_t ev[2] = {nullptr, nullptr};
        ~Bounce() {
            (void)hipStreamSynchronize(s);
            for (uint32_t *p : b)
                if (p) (void)hipHostFree(p);
            for (hipEvent_t e : ev)
                if (e) (void)hipEventDestroy(e);
        }
    } bb{d->s_compute};
    uint64_t maxrow = 0;
    for (uint64_t v = 0; v < vend; ++v) maxrow = std::max(maxrow, g->ptr[v + 1] - g->ptr[v]);
    const uint64_t C = std::max<uint64_t>(std::min<uint64_t>(n, 16ull << 20), maxrow);
    for (auto &p : bb.b) OG_HIP(hipHostMalloc((void **)&p, C * 4, 0));
    for (auto &e : bb.ev) OG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (uint64_t v0 = 0, k = 0; v0 < vend; ++k) {
        uint64_t v1 = v0;
        while (v1 < vend && g->ptr[v1 + 1] - g->ptr[v0] <= C) ++v1;
        const int s = (int)(k & 1);
        if (k >= 2) OG_HIP(hipEventSynchronize(bb.ev[s]));
        const uint64_t m = g->ptr[v1] - g->ptr[v0];
        pack_rows(g, v0, v1, d->wbits, d->split_weight, bb.b[s]);
        OG_HIP(hipMemcpyAsync(dst + g->ptr[v0], bb.b[s], m * 4, hipMemcpyHostToDevice, d->s_compute));
        OG_HIP(hipEventRecord(bb.ev[s], d->s_compute));
        v0 = v1;
    }
    OG_HIP(hipStreamSynchronize(d->s_compute));
}

// Zero-copy loading reads dynamic rows in place, so the staging share of the
// edge budget (cap edges) holds part of the dynamic region in HBM: blocks of
// kResChunk consecutive vertices, chosen by the link lines their rows would
// cost per edge.  A row of b bytes at an arbitrary 4-byte offset touches
// (b + kLinkLine - 4) / kLinkLine lines of kLinkLine bytes on average, each
// fetched whole, so blocks of short rows are taken first (ties: lower ids).
// With the light/heavy split each chunk has two blocks, its light parts and
// its heavy parts, ranked by the same measure and chosen separately.
// (Ranking every light block first, for the light rounds' repeated reads,
// measured slower on R-MAT: the short heavy rows it leaves on the host cost
// the heavy round more.)  The chunk table (8 bytes per block) is charged to
// the same share.
static void stage_resident(DeviceState *d, const og_graph *g, uint64_t vsplit, uint64_t cap,
                           const std::vector<uint64_t> &lpos) {
    const uint64_t V = g->V, eb = d->eb;
    const bool split = !lpos.empty();  // two blocks per chunk: light parts, heavy parts
    const uint64_t nb = split ? 2 : 1;  // blocks per chunk
    const uint64_t c0 = vsplit >> kResChunkShift, c1 = (V + kResChunk - 1) >> kResChunkShift;
    const uint64_t table_edges = (nb * c1 * sizeof(int64_t) + eb - 1) / eb;
    if (vsplit >= V || cap <= table_edges) return;
    cap -= table_edges;
    const uint64_t nch = c1 - c0;
    auto first = [&](uint64_t i) { return std::max<uint64_t>((c0 + i) << kResChunkShift, vsplit); };
    auto last = [&](uint64_t i) { return std::min<uint64_t>((c0 + i + 1) << kResChunkShift, V); };
    // block k = nb * i + part of chunk i (part 0: light or whole rows, 1: heavy)
    const uint64_t nblk = nb * nch;
    std::vector<uint64_t> bedges(nblk);
    std::vector<double> score(nblk);
    auto lines_of = [&](uint64_t n) { return n ? (double)(n * eb + kLinkLine - 4) / kLinkLine : 0.0; };
    parallel_for(nch, 1 << 12, 0, [&](uint64_t b, uint64_t e, unsigned) {
        for (uint64_t i = b; i < e; ++i) {
            double lines[2] = {0, 0};
            uint64_t n[2] = {0, 0};
            for (uint64_t v = first(i); v < last(i); ++v) {
                const uint64_t deg = g->ptr[v + 1] - g->ptr[v];
                const uint64_t nl = split ? lpos[v - vsplit + 1] - lpos[v - vsplit] : deg;
                lines[0] += lines_of(nl);
                n[0] += nl;
                lines[1] += lines_of(deg - nl);
                n[1] += deg - nl;
            }
            for (uint64_t k = 0; k < nb; ++k) {
                bedges[nb * i + k] = n[k];
                score[nb * i + k] = n[k] ? lines[k] / (double)n[k] : 0.0;
            }
        }
    });
    std::vector<uint32_t> order(nblk);  // nblk <= 2^28
    for (uint64_t i = 0; i < nblk; ++i) order[i] = (uint32_t)i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return score[x] > score[y]; });
    std::vector<uint8_t> take(nblk, 0);
    uint64_t used = 0;
    for (uint32_t k : order) {
        if (!bedges[k] || used + bedges[k] > cap) continue;
        take[k] = 1;
        used += bedges[k];
    }
    if (!used) return;
    // The resident buffer holds the chosen blocks in block order (a chunk's
    // light block before its heavy block); runs of host positions that follow
    // each other are gathered in one piece.  Positions are in the host copy
    // of the dynamic region (with the split: light parts, then heavy parts).
    struct Run {
        uint64_t src, dst, n;
    };
    std::vector<Run> runs;
    auto add_run = [&](uint64_t src, uint64_t dst, uint64_t n) {
        if (!n) return;
        if (!runs.empty() && runs.back().src + runs.back().n == src && runs.back().dst + runs.back().n == dst)
            runs.back().n += n;
        else
            runs.push_back({src, dst, n});
    };
    const uint64_t dbase = d->dyn_base, nl = split ? lpos[V - vsplit] : 0;
    std::vector<int64_t> delta(nb * c1, kNotResident);
    uint64_t pos = 0;
    for (uint64_t k = 0; k < nblk; ++k) {
        if (!take[k]) continue;
        const uint64_t i = k / nb, v0 = first(i);
        uint64_t h0;  // host position of the block's first edge
        if (!split) h0 = g->ptr[v0] - dbase;
        else if (k % nb == 0) h0 = lpos[v0 - vsplit];
        else h0 = nl + (g->ptr[v0] - dbase) - lpos[v0 - vsplit];
        delta[nb * c0 + k] = (int64_t)pos - (int64_t)h0;
        add_run(h0, pos, bedges[k]);
        pos += bedges[k];
    }
    MemPool tmp;
    uint8_t *hb = (uint8_t *)tmp.palloc(used * eb);
    const uint8_t *src = (const uint8_t *)d->h_dyn;
    parallel_for(runs.size(), 64, 0, [&](uint64_t b, uint64_t e, unsigned) {
        for (uint64_t r = b; r < e; ++r) memcpy(hb + runs[r].dst * eb, src + runs[r].src * eb, runs[r].n * eb);
    });
    d->dyn_resident = used;
    d->d_staging = d->dalloc(used * eb);
    d->d_rdelta = (int64_t *)d->dalloc(delta.size() * sizeof(int64_t));
    OG_HIP(hipMemcpyAsync(d->d_staging, hb, used * eb, hipMemcpyHostToDevice, d->s_compute));
    OG_HIP(hipMemcpyAsync(d->d_rdelta, delta.data(), delta.size() * sizeof(int64_t), hipMemcpyHostToDevice,
                          d->s_compute));
    OG_HIP(hipStreamSynchronize(d->s_compute));
}

// The compaction's degree table (DeviceState::d_deg): 4 bytes per vertex
// and range in place of two 8-byte row bounds.  Not built when a degree
// exceeds 32 bits (the passes then read the row bounds).
static void upload_degrees(DeviceState *d, const og_graph *g) {
    const uint64_t V = g->V, Vp = d->Vpad;
    const bool split = d->split_weight != 0;
    std::atomic<bool> wide{false};
    std::vector<uint32_t> dg((split ? 2 : 1) * Vp, 0);
    parallel_for(V, 1 << 16, 0, [&](uint64_t b, uint64_t e, unsigned) {
        bool w = false;
        for (uint64_t v = b; v < e; ++v) {
            const uint64_t b0 = g->ptr[v], e0 = g->ptr[v + 1];
            const uint64_t m = split ? d->h_pmid[v] : e0;
            w |= (e0 - b0) >> 32 != 0;
            dg[v] = (uint32_t)(m - b0);
            if (split) dg[Vp + v] = (uint32_t)(e0 - m);
        }
        if (w) wide = true;
    });
    if (wide) return;
    d->d_deg = (uint32_t *)d->dalloc(dg.size() * sizeof(uint32_t));
    OG_HIP(hipMemcpyAsync(d->d_deg, dg.data(), dg.size() * sizeof(uint32_t), hipMemcpyHostToDevice, d->s_compute));
    OG_HIP(hipStreamSynchronize(d->s_compute));
    d->deg_split = split;
}

static void prepare_impl(og_graph *g, og_approach approach, const og_plan_params *params, og_plan_info *info) {
    if (!hip_available()) raise(OG_EHIP, "no HIP device is visible");
    if (approach != OG_ASCETIC && approach != OG_SUBWAY && approach != OG_USM)
        raise(OG_EINVAL, "unknown approach %d", (int)approach);
    og_plan_params p{};
    if (params) p = *params;
    if (g->dev) {
        device_state_free(g->dev);
        g->dev = nullptr;
    }
    std::unique_ptr<DeviceState, void (*)(DeviceState *)> d(new DeviceState, device_state_free);
    OG_HIP(hipGetDevice(&d->device));
    d->approach = approach;
    d->V = g->V;
    d->E = g->E;
    d->Vpad = (g->V + 15) / 16 * 16;
    d->eb = g->edge_bytes();
    d->flags = p.flags;
    d->host_threads = p.host_threads ? p.host_threads : default_threads();
    // Compact weighted format: when ids and weights fit one 32-bit word, the
    // static region, the staging slots and the link all carry 4 bytes per
    // edge instead of 8, so the same byte budget holds twice the edges.  The
    // planner sizes its regions in these device bytes.  USM reads the CSR as
    // it lies in host memory and keeps the 8-byte form.
    if (g->kind == OG_KIND_WEIGHTED && approach != OG_USM && !(p.flags & OG_PLAN_NO_PACK) && g->E) {
        const uint32_t wb = weight_bits(g);
        if (wb < 32 && g->V - 1 < (1ull << (32 - wb))) {
            d->packed = true;
            d->wbits = wb;
            d->eb = 4;
        }
    }
    d->deb = d->eb;
    const uint64_t V = g->V, E = g->E, eb = d->eb;
    // Light/heavy split: rows of the packed device / link copies list light
    // edges first; pmid[v] is where row v's heavy edges start.
    if (p.split_weight) {
        if (!d->packed)
            raise(OG_EUNSUPPORTED, "the light/heavy split needs packed weighted edges (Ascetic or Subway, ids and "
                                   "weights in one 32-bit word, no OG_PLAN_NO_PACK)");
        d->split_weight = p.split_weight;
        d->h_pmid.resize(V);
        parallel_for(V, 1 << 14, 0, [&](uint64_t b, uint64_t e, unsigned) {
            for (uint64_t v = b; v < e; ++v) {
                uint64_t k = g->ptr[v];
                for (uint64_t i = g->ptr[v]; i < g->ptr[v + 1]; ++i) k += g->wedges[2 * i + 1] < p.split_weight;
                d->h_pmid[v] = k;
            }
        });
    }

    // ---- plan
    og_plan_info &pl = d->plan;
    if (approach == OG_ASCETIC) {
        og_status st = og_plan_compute(V, E, g->ptr, (uint32_t)eb, &p, &pl);
        if (st != OG_OK) raise(st, "%s", og_last_error());
    } else if (approach == OG_SUBWAY) {
        // Subway keeps no static region; the whole budget is the transfer
        // buffer (reference: 0.9 * (12 GiB - 32 V) / 4, BFS/sbw_BFS/main.cpp:334).
        uint64_t total = p.edge_budget_bytes ? p.edge_budget_bytes / eb
                                             : (uint64_t)(0.9 * (double)((12ull << 30) - 32ull * V) / (double)eb);
        std::memset(&pl, 0, sizeof pl);
        pl.total_budget_edges = total;
        pl.static_edges = 0;
        pl.staging_edges = total;
        pl.overload_edges = E;
        pl.static_vertices = 0;
        pl.vsplit = 0;
        pl.advisek_used = 0.0f;
    } else {
        std::memset(&pl, 0, sizeof pl);
        pl.total_budget_edges = E;
        pl.static_edges = E;
        pl.static_vertices = V;
        pl.vsplit = V;
        pl.advisek_used = 1.0f;
    }
    pl.approach = (uint32_t)approach;

    OG_HIP(hipStreamCreateWithFlags(&d->s_compute, hipStreamNonBlocking));
    OG_HIP(hipStreamCreateWithFlags(&d->s_copy, hipStreamNonBlocking));
    OG_HIP(hipStreamCreateWithFlags(&d->s_static, hipStreamNonBlocking));
    OG_HIP(hipEventCreateWithFlags(&d->ev_static, hipEventDisableTiming));
    OG_HIP(hipEventCreateWithFlags(&d->ev_fork, hipEventDisableTiming));

    // ---- graph placement
    if (approach == OG_USM) {
        // Graph arrays stay in pinned host memory and the kernels read them
        // through the device mapping: what SYCL malloc_shared gives on a GPU
        // without XNACK page migration.
        if (!g->pinned_ptr || !(eb == 8 ? g->pinned_wedges : g->pinned_cols))
            raise(OG_EUNSUPPORTED, "USM approach needs the graph in pinned host memory");
        void *dp = nullptr, *de = nullptr;
        OG_HIP(hipHostGetDevicePointer(&dp, g->ptr, 0));
        OG_HIP(hipHostGetDevicePointer(&de, (void *)g->edge_data(), 0));
        d->d_ptr = (uint64_t *)dp;
        d->d_static = de;
        d->static_used = E;
        d->managed = true;
    } else {
        d->d_ptr = (uint64_t *)d->dalloc((V + 1) * sizeof(uint64_t));
        OG_HIP(hipMemcpyAsync(d->d_ptr, g->ptr, (V + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, d->s_compute));
        if (d->split_weight) {
            d->d_pmid = (uint64_t *)d->dalloc(std::max<uint64_t>(V, 1) * sizeof(uint64_t));
            OG_HIP(hipMemcpyAsync(d->d_pmid, d->h_pmid.data(), V * sizeof(uint64_t), hipMemcpyHostToDevice,
                                  d->s_compute));
        }
        upload_degrees(d.get(), g);
        if (pl.static_edges) {
            d->d_static = d->dalloc(pl.static_edges * eb);
            d->static_used = g->ptr[pl.vsplit];
            if (d->packed)
                upload_packed(d.get(), g, pl.vsplit, (uint32_t *)d->d_static);
            else if (d->static_used)
                OG_HIP(hipMemcpyAsync(d->d_static, g->edge_data(), d->static_used * eb, hipMemcpyHostToDevice,
                                      d->s_compute));
        }
        if (pl.vsplit < V) {
            // PageRank streams its dynamic region through the slots (vertex-
            // aligned chunks), so a pull-layout graph always gets them
            const bool zc = (p.flags & OG_PLAN_ZERO_COPY) != 0 && g->kind != OG_KIND_PULL;
            // Staging slots: gather, copy and relax of consecutive chunks
            // overlap.  Chunks of ~kSlotTargetBytes: smaller ones cost more
            // in per-chunk gather dispatch and launches than they save in
            // pipeline fill.  Never smaller than the largest dynamic row,
            // which the vertex-aligned PageRank chunks need whole.
            uint64_t maxdyn = 0;
            for (uint64_t v = pl.vsplit; v < V; ++v) maxdyn = std::max(maxdyn, g->ptr[v + 1] - g->ptr[v]);
            uint32_t ns = (uint32_t)std::min<uint64_t>(kMaxSlots, std::max<uint64_t>(2, pl.staging_edges * eb /
                                                                                          kSlotTargetBytes));
            while (ns > 2 && pl.staging_edges / ns < maxdyn) --ns;
            d->nslots = ns;
            d->slot_edges = pl.staging_edges / ns;
            if (!zc) {
                if (d->slot_edges == 0)
                    raise(OG_EBUDGET, "edge budget leaves %llu staging edges for %llu dynamic-region edges",
                          (unsigned long long)pl.staging_edges, (unsigned long long)(E - g->ptr[pl.vsplit]));
                d->d_staging = d->dalloc((uint64_t)ns * d->slot_edges * eb);
                d->h_stage.resize(ns);
                d->ev_h2d.assign(ns, nullptr);
                d->ev_relax.assign(ns, nullptr);
                d->slot_h2d_pending.assign(ns, 0);
                d->slot_relax_pending.assign(ns, 0);
                for (uint32_t s = 0; s < ns; ++s) {
                    d->h_stage[s] = d->palloc(d->slot_edges * eb);
                    OG_HIP(hipEventCreateWithFlags(&d->ev_h2d[s], hipEventDisableTiming));
                    OG_HIP(hipEventCreateWithFlags(&d->ev_relax[s], hipEventDisableTiming));
                }
                d->h_dlist = (uint32_t *)d->palloc(V * sizeof(uint32_t));
                d->h_doff = (uint64_t *)d->palloc((V + 1) * sizeof(uint64_t));
                d->pool.reset(new WorkerPool(d->host_threads));
            }
            // Host source of the dynamic region in the device format: the
            // CSR itself, or (packed) a pinned packed copy built once here.
            d->dyn_base = g->ptr[pl.vsplit];
            const uint64_t nd = E - d->dyn_base;
            // zero-copy reads with the light/heavy split: all light parts
            // first, so that a light round's reads of neighbouring rows share
            // the link's 128-byte lines (staged chunks keep rows whole)
            std::vector<uint64_t> lpos;
            if (zc && d->split_weight) {
                lpos.resize(V - pl.vsplit + 1);
                lpos[0] = 0;
                for (uint64_t v = pl.vsplit; v < V; ++v)
                    lpos[v - pl.vsplit + 1] = lpos[v - pl.vsplit] + (d->h_pmid[v] - g->ptr[v]);
                d->split_nl = lpos.back();
                d->d_lpos = (uint64_t *)d->dalloc(lpos.size() * sizeof(uint64_t));
                OG_HIP(hipMemcpyAsync(d->d_lpos, lpos.data(), lpos.size() * sizeof(uint64_t), hipMemcpyHostToDevice,
                                      d->s_compute));
                OG_HIP(hipStreamSynchronize(d->s_compute));
            }
            if (d->packed) {
                uint32_t *pk = (uint32_t *)d->palloc(std::max<uint64_t>(nd, 1) * 4);
                if (!lpos.empty()) pack_split_rows(g, pl.vsplit, d->wbits, d->split_weight, lpos.data(), pk);
                else pack_rows(g, pl.vsplit, V, d->wbits, d->split_weight, pk);
                d->h_dyn = pk;
            } else {
                d->h_dyn = (const uint8_t *)g->edge_data() + d->dyn_base * eb;
            }
            d->slot_dyn = d->slot_edges;
            if (zc) {
                if (!d->packed && !(eb == 8 ? g->pinned_wedges : g->pinned_cols))
                    raise(OG_EUNSUPPORTED, "zero-copy loading needs the graph in pinned host memory");
                // map the allocation's base, then offset to the dynamic region
                const void *hbase = d->packed ? d->h_dyn : g->edge_data();
                void *hd = nullptr;
                OG_HIP(hipHostGetDevicePointer(&hd, (void *)hbase, 0));
                d->hd_dyn = (const uint8_t *)hd + ((const uint8_t *)d->h_dyn - (const uint8_t *)hbase);
                // Reading in place needs no transit slots, so the staging
                // share of the budget keeps part of the dynamic region
                // resident in HBM instead.
                stage_resident(d.get(), g, pl.vsplit, pl.staging_edges, lpos);
            }
        }
    }

    // ---- in-edge CSR for direction-optimising BFS (counting sort by target)
    if (p.flags & OG_PLAN_DIRECTION_OPT) {
        if (approach != OG_ASCETIC || pl.vsplit < V)
            raise(OG_EUNSUPPORTED, "direction-optimising BFS needs the whole graph in HBM (Ascetic, vsplit = V)");
        if (g->kind == OG_KIND_PULL) raise(OG_EUNSUPPORTED, "pull-layout graphs serve PageRank only");
        const bool wt = g->kind == OG_KIND_WEIGHTED;
        auto col = [&](uint64_t i) { return wt ? g->wedges[2 * i] : g->cols[i]; };
        std::vector<uint64_t> iptr(V + 1, 0);
        std::vector<std::atomic<uint64_t>> cnt(V);
        parallel_for(V, 1 << 16, 0, [&](uint64_t b, uint64_t e, unsigned) {
            for (uint64_t v = b; v < e; ++v) cnt[v].store(0, std::memory_order_relaxed);
        });
        parallel_for(E, 1 << 20, 0, [&](uint64_t b, uint64_t e, unsigned) {
            for (uint64_t i = b; i < e; ++i) cnt[col(i)].fetch_add(1, std::memory_order_relaxed);
        });
        for (uint64_t v = 0; v < V; ++v) iptr[v + 1] = iptr[v] + cnt[v].load(std::memory_order_relaxed);
        parallel_for(V, 1 << 16, 0, [&](uint64_t b, uint64_t e, unsigned) {
            for (uint64_t v = b; v < e; ++v) cnt[v].store(iptr[v], std::memory_order_relaxed);
        });
        std::vector<uint32_t> icol(std::max<uint64_t>(E, 1));
        parallel_for(V, 1 << 12, 0, [&](uint64_t b, uint64_t e, unsigned) {
            for (uint64_t u = b; u < e; ++u)
                for (uint64_t i = g->ptr[u]; i < g->ptr[u + 1]; ++i)
                    icol[cnt[col(i)].fetch_add(1, std::memory_order_relaxed)] = (uint32_t)u;
        });
        d->d_in_ptr = (uint64_t *)d->dalloc((V + 1) * sizeof(uint64_t));
        d->d_in_cols = (uint32_t *)d->dalloc(std::max<uint64_t>(E, 1) * sizeof(uint32_t));
        OG_HIP(hipMemcpy(d->d_in_ptr, iptr.data(), (V + 1) * sizeof(uint64_t), hipMemcpyHostToDevice));
        OG_HIP(hipMemcpy(d->d_in_cols, icol.data(), std::max<uint64_t>(E, 1) * sizeof(uint32_t),
                         hipMemcpyHostToDevice));
    }

    // ---- vertex state
    d->d_val = (uint32_t *)d->dalloc(V * sizeof(uint32_t));
    OG_HIP(hipStreamSynchronize(d->s_compute));
    frontier_alloc(d.get(), g, d->fr, d->mem);
    pl.device_bytes = d->device_bytes();
    if (info) *info = pl;
    g->dev = d.release();
}

// ------------------------------------------------------- frontier driver

enum { kClsFront = 0, kClsStat = 1, kClsDyn = 2, kClsCopy = 3 };

struct RunTimers {
    TimedClass front, stat, dyn, copy;
    TimedClass &at(int c) { return c == kClsFront ? front : c == kClsStat ? stat : c == kClsDyn ? dyn : copy; }
};

// A captured batch of device-driven rounds, replayed with one hipGraphLaunch.
// Its timing marks are external event-record nodes, read after each replay.
struct RoundGraph {
    struct Mark {
        int cls;
        hipEvent_t a, b;
    };
    uint64_t key = ~0ull;
    hipGraphExec_t exec = nullptr;
    std::vector<hipEvent_t> events;
    std::vector<Mark> marks;
    ~RoundGraph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
    }
};

// Timing marks of enqueued work.  EagerSpan: an EventSpan of the run per
// launch sequence.  GraphSpan: a pair of event-record nodes in the graph
// being captured; with `chain`, a span that follows another on the same
// stream with nothing enqueued between them starts at that span's end event
// (each event node costs the replay about ten microseconds).
struct EagerSpan {
    DeviceState *d;
    RunTimers &tm;
    template <class F>
    void operator()(int cls, hipStream_t s, F &&fn) const {
        EventSpan span(d, &tm.at(cls), s);
        fn();
    }
};

struct GraphSpan {
    RoundGraph *rg;
    bool chain = false;
    mutable hipEvent_t last = nullptr;
    mutable hipStream_t last_s = nullptr;
    template <class F>
    void operator()(int cls, hipStream_t s, F &&fn) const {
        hipEvent_t a = nullptr, b = nullptr;
        if (chain && last && last_s == s) {
            a = last;
        } else {
            OG_HIP(hipEventCreate(&a));
            rg->events.push_back(a);
            OG_HIP(hipEventRecordWithFlags(a, s, hipEventRecordExternal));
        }
        OG_HIP(hipEventCreate(&b));
        rg->events.push_back(b);
        fn();
        OG_HIP(hipEventRecordWithFlags(b, s, hipEventRecordExternal));
        rg->marks.push_back({cls, a, b});
        last = b;
        last_s = s;
    }
};

// Captured batches of rounds of one preparation, by key (algorithm, layout,
// rounds per batch); the oldest is dropped beyond kGraphCache.
constexpr size_t kGraphCache = 8;
struct RoundGraphCache {
    std::vector<std::unique_ptr<RoundGraph>> g;
};

// The cached graph for `key`, captured on the compute stream by body(rg)
// when the cache does not hold it.
template <class Body>
static RoundGraph *round_graph_for(DeviceState *d, uint64_t key, Body body) {
    if (!d->round_graph)
        d->round_graph = std::shared_ptr<void>(new RoundGraphCache,
                                               [](void *p) { delete static_cast<RoundGraphCache *>(p); });
    auto &cache = static_cast<RoundGraphCache *>(d->round_graph.get())->g;
    for (auto &c : cache)
        if (c->key == key) return c.get();
    std::unique_ptr<RoundGraph> rg(new RoundGraph);
    hipStream_t s = d->s_compute;
    OG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipGraph_t graph = nullptr;
    try {
        body(rg.get());
    } catch (...) {
        (void)hipStreamEndCapture(s, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
    }
    OG_HIP(hipStreamEndCapture(s, &graph));
    const hipError_t e = hipGraphInstantiate(&rg->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    OG_HIP(e);
    rg->key = key;
    if (cache.size() >= kGraphCache) cache.erase(cache.begin());
    cache.push_back(std::move(rg));
    return cache.back().get();
}

// Host gather of dynamic edges [c0, c1) into a pinned slot (reference
// fillEdgeArrByMultiThread, GraphMeta.h:470-506, with the copy offsets taken
// from the device-built degree prefix instead of a serial host scan).  The
// source is the dynamic region in its transfer format (d->h_dyn, whose first
// element is edge d->dyn_base of the CSR).
static void gather(DeviceState *d, const og_graph *g, uint32_t ndyn, uint64_t c0, uint64_t c1, uint8_t *dst,
                   bool heavy) {
    const uint64_t eb = d->deb;
    const uint8_t *src = (const uint8_t *)d->h_dyn;
    const uint64_t base = d->dyn_base;
    const uint64_t n = c1 - c0;
    const unsigned T = d->pool->size();
    const unsigned use = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(T, n / 16384));
    const uint64_t L = (n + use - 1) / use;
    d->pool->run([&](unsigned t) {
        if (t >= use) return;
        const uint64_t p0 = c0 + (uint64_t)t * L;
        const uint64_t p1 = std::min(c1, p0 + L);
        if (p0 >= p1) return;
        uint64_t i = (uint64_t)(std::upper_bound(d->h_doff, d->h_doff + ndyn + 1, p0) - d->h_doff) - 1;
        uint64_t p = p0;
        while (p < p1) {
            const uint32_t v = d->h_dlist[i];
            const uint64_t vb = d->h_doff[i], ve = d->h_doff[i + 1];
            const uint64_t stop = std::min(ve, p1);
            const uint64_t rb = heavy ? d->h_pmid[v] : g->ptr[v];  // row start of the round's range
            std::memcpy(dst + (p - c0) * eb, src + (rb - base + (p - vb)) * eb, (stop - p) * eb);
            p = stop;
            ++i;
        }
    });
}

static RelaxArgs relax_args(DeviceState *d) {
    return RelaxArgs{d->d_val, d->fr.flag, d->d_visit, d->d_front, d->d_next, d->wbits, nullptr, d->run_pmid};
}

// Grids of device-count launches: the kernels stride over whatever count
// the round's compaction produced, so without a host-side count a fixed
// grid that fills the chip (256 CUs) is launched.
constexpr unsigned kGroupGridCap = 2048;  // workgroups of k_relax_group
constexpr unsigned kHubGridCap = 1024;    // workgroups of k_relax_hub
constexpr unsigned kDynGridCap = 8192;    // workgroups of k_relax_dyn (zero-copy reads want many in flight)

static unsigned capped(uint64_t blocks, unsigned cap) {
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(blocks, cap));
}

template <int ALGO, bool W, bool PK>
static void launch_static_relax_fmt(DeviceState *d, Frontier &f, const FrontierInfo *hfi, const RelaxArgs &a,
                                    hipStream_t s) {
    const void *edges = d->d_static;
    unsigned long long *upd = f.counters + 0;
    const uint32_t lpv0 = d->tune.lpv0;
    if (!hfi || hfi->n[0]) {
        const dim3 gr(hfi ? capped(((uint64_t)hfi->n[0] * lpv0 + kBlock - 1) / kBlock, kGroupGridCap) : kGroupGridCap);
        switch (lpv0) {
#define OG_LPV(L)                                                                                                 \
    case L:                                                                                                       \
        hipLaunchKernelGGL((k_relax_group<ALGO, W, PK, L>), gr, dim3(kBlock), 0, s, f.list, 0, f.info, d->d_ptr, \
                           edges, a, upd);                                                                        \
        break;
            OG_LPV(1) OG_LPV(2) OG_LPV(4) OG_LPV(8) OG_LPV(16) OG_LPV(32)
            default: OG_LPV(64)
#undef OG_LPV
        }
    }
    if (!hfi || hfi->n[1])
        hipLaunchKernelGGL((k_relax_group<ALGO, W, PK, 64>),
                           dim3(hfi ? capped(((uint64_t)hfi->n[1] * 64 + kBlock - 1) / kBlock, kGroupGridCap)
                                    : kGroupGridCap),
                           dim3(kBlock), 0, s, f.list, 1, f.info, d->d_ptr, edges, a, upd);
    if (!hfi || hfi->n[2])
        hipLaunchKernelGGL((k_relax_hub<ALGO, W, PK>), dim3(hfi ? capped(hfi->n[2], kHubGridCap) : kHubGridCap),
                           dim3(kBlock), d->tune.tile * d->eb, s, f.hub, f.hub_cap, f.info, d->tune.tile, d->d_ptr,
                           edges, a, upd);
    OG_HIP(hipGetLastError());
}

template <int ALGO, bool W>
void launch_static_relax(DeviceState *d, Frontier &f, const FrontierInfo *hfi, const RelaxArgs &a, hipStream_t s) {
    if (W && d->packed) launch_static_relax_fmt<ALGO, W, true>(d, f, hfi, a, s);
    else launch_static_relax_fmt<ALGO, W, false>(d, f, hfi, a, s);
}
template void launch_static_relax<ALGO_MSBFS, false>(DeviceState *, Frontier &, const FrontierInfo *, const RelaxArgs &, hipStream_t);
template void launch_static_relax<ALGO_MSBFS, true>(DeviceState *, Frontier &, const FrontierInfo *, const RelaxArgs &, hipStream_t);

// Static-region relax of a round on s_static, so that it overlaps the
// dynamic region (link bound): forked from the compute stream after the
// compaction; returns true when the caller must join ev_static before the
// next compaction.  All relax updates are monotone (atomicMin / same-level
// stores / atomicOr), so the two regions may interleave.  NO_OVERLAP runs it
// (and a plan without a dynamic region) runs it on the compute stream
// instead.  hfi: the round's summary when the host has
// it (sizes the grids, skips empty bins), else null.
template <int ALGO, bool W, class Span>
static bool launch_static(DeviceState *d, const FrontierInfo *hfi, const Span &span) {
    if (hfi && !(hfi->n[0] | hfi->n[1] | hfi->n[2])) return false;
    if ((d->flags & OG_PLAN_NO_OVERLAP) || d->plan.vsplit >= d->V) {  // nothing to overlap with
        span(kClsStat, d->s_compute,
             [&] { launch_static_relax<ALGO, W>(d, d->fr, hfi, relax_args(d), d->s_compute); });
        return false;
    }
    OG_HIP(hipEventRecord(d->ev_fork, d->s_compute));
    OG_HIP(hipStreamWaitEvent(d->s_static, d->ev_fork, 0));
    span(kClsStat, d->s_static, [&] { launch_static_relax<ALGO, W>(d, d->fr, hfi, relax_args(d), d->s_static); });
    OG_HIP(hipEventRecord(d->ev_static, d->s_static));
    return true;
}

template <int ALGO, bool W, bool PACKED, bool DIRECT>
static void launch_dyn(DeviceState *d, unsigned grid, const void *src, uint64_t c0, uint64_t c1) {
    hipLaunchKernelGGL((k_relax_dyn<ALGO, W, PACKED, DIRECT>), dim3(grid), dim3(kBlock), 0, d->s_compute, d->fr.list,
                       d->fr.info, d->fr.doff, src, c0, c1, d->d_ptr, d->dyn_base, (const void *)d->d_staging,
                       DIRECT ? (const int64_t *)d->d_rdelta : nullptr, d->wbits, relax_args(d),
                       DIRECT ? (const uint64_t *)d->d_lpos : nullptr, d->split_nl, d->plan.vsplit,
                       DIRECT && c0 == 0 ? (const uint32_t *)d->fr.dchunk : nullptr, d->fr.counters + 1,
                       d->fr.counters + 5);
}

// The fused round relax (k_relax_round) of a device-driven round on the
// compute stream: the static grids of the separate launches (kGroupGridCap
// twice, kHubGridCap) and, with_dyn (zero-copy), the dynamic grid rounded up
// to a multiple of kHubGridCap.
template <int ALGO, bool W, bool PK, int LPV>
static void launch_round_lpv(DeviceState *d, const RoundRelax &r, unsigned grid, size_t lds) {
    hipLaunchKernelGGL((k_relax_round<ALGO, W, PK, LPV>), dim3(grid), dim3(kBlock), lds, d->s_compute, r,
                       relax_args(d));
}

template <int ALGO, bool W, bool PK>
static void launch_round_fmt(DeviceState *d, bool with_dyn) {
    static_assert(2 * kHubGridCap == kGroupGridCap, "period layout: 2 + 2 + 1 static workgroups per K");
    const unsigned dyn = d->tune.dyn_grid ? d->tune.dyn_grid : kDynGridCap;
    RoundRelax r{};
    r.list = d->fr.list;
    r.info = d->fr.info;
    r.ptr = d->d_ptr;
    r.sedges = d->d_static;
    r.hub = d->fr.hub;
    r.hub_cap = d->fr.hub_cap;
    r.tile = d->tune.tile;
    r.cd = with_dyn ? (dyn + kHubGridCap - 1) / kHubGridCap : 0;
    r.doff = d->fr.doff;
    r.stage = d->hd_dyn;
    r.dbase = d->dyn_base;
    r.resident = d->d_staging;
    r.rdelta = d->d_rdelta;
    r.lpos = d->d_lpos;
    r.dchunk = d->fr.dchunk;
    r.nl = d->split_nl;
    r.vsplit = d->plan.vsplit;
    r.counters = d->fr.counters;
    const unsigned grid = (kFuseStatic + r.cd) * kHubGridCap;
    const size_t lds = std::max<size_t>(with_dyn ? sizeof(DynLds) : 0, (size_t)d->tune.tile * d->eb);
    switch (d->tune.lpv0) {
        case 1: launch_round_lpv<ALGO, W, PK, 1>(d, r, grid, lds); break;
        case 2: launch_round_lpv<ALGO, W, PK, 2>(d, r, grid, lds); break;
        case 4: launch_round_lpv<ALGO, W, PK, 4>(d, r, grid, lds); break;
        case 8: launch_round_lpv<ALGO, W, PK, 8>(d, r, grid, lds); break;
        case 16: launch_round_lpv<ALGO, W, PK, 16>(d, r, grid, lds); break;
        case 32: launch_round_lpv<ALGO, W, PK, 32>(d, r, grid, lds); break;
        default: launch_round_lpv<ALGO, W, PK, 64>(d, r, grid, lds); break;
    }
    OG_HIP(hipGetLastError());
}

template <int ALGO, bool W>
static void launch_round_relax(DeviceState *d, bool with_dyn) {
    if (W && d->packed) launch_round_fmt<ALGO, W, true>(d, with_dyn);
    else launch_round_fmt<ALGO, W, false>(d, with_dyn);
}

template <bool W, bool PK>
static void launch_srv_round_fmt(DeviceState *d, Frontier &f, const RelaxArgs &a, hipStream_t s, const SrvPull *pull) {
    RoundRelax r{};
    r.list = f.list;
    r.info = f.info;
    r.ptr = d->d_ptr;
    r.sedges = d->d_static;
    r.hub = f.hub;
    r.hub_cap = f.hub_cap;
    r.tile = d->tune.tile;
    r.counters = f.counters;
    SrvPull p{};
    if (pull) p = *pull;
    const uint32_t cp = pull ? kDynGridCap / kHubGridCap : 0;  // the separate pull's 8192 workgroups
    const unsigned grid = (kFuseStatic + cp) * kHubGridCap;
    const size_t lds = (size_t)d->tune.tile * d->eb;
    switch (d->tune.lpv0) {
#define OG_SRV(L)                                                                                                  \
    case L:                                                                                                        \
        hipLaunchKernelGGL((k_srv_round<W, PK, L>), dim3(grid), dim3(kBlock), lds, s, r, a, p, cp);               \
        break;
        OG_SRV(1) OG_SRV(2) OG_SRV(4) OG_SRV(8) OG_SRV(16) OG_SRV(32)
        default: OG_SRV(64)
#undef OG_SRV
    }
    OG_HIP(hipGetLastError());
}

void launch_srv_round(DeviceState *d, Frontier &f, const RelaxArgs &a, hipStream_t s, const SrvPull *pull,
                      bool weighted) {
    if (weighted && d->packed) launch_srv_round_fmt<true, true>(d, f, a, s, pull);
    else if (weighted) launch_srv_round_fmt<true, false>(d, f, a, s, pull);
    else launch_srv_round_fmt<false, false>(d, f, a, s, pull);
}

// Dynamic relax of edges [c0, min(c1, D)); `known` = c1 - c0 when the host
// knows the edge count (exact grid), else 0 (fixed grid).
template <int ALGO, bool W>
static void launch_dyn_any(DeviceState *d, const void *src, uint64_t c0, uint64_t c1, uint64_t known, bool direct) {
    const unsigned cap = d->tune.dyn_grid ? d->tune.dyn_grid : kDynGridCap;
    const unsigned grid = known ? capped((known + kDynEdgesPerBlock - 1) / kDynEdgesPerBlock, cap) : cap;
    if (d->packed) {
        if (direct) launch_dyn<ALGO, W, true, true>(d, grid, src, c0, c1);
        else launch_dyn<ALGO, W, true, false>(d, grid, src, c0, c1);
    } else {
        if (direct) launch_dyn<ALGO, W, false, true>(d, grid, src, c0, c1);
        else launch_dyn<ALGO, W, false, false>(d, grid, src, c0, c1);
    }
    OG_HIP(hipGetLastError());
}

// Staged dynamic region of a round whose summary the host holds: the host
// gathers the active edges into pinned slots and the copy stream moves them
// while the other slot is relaxed.
template <int ALGO, bool W>
static void run_dynamic(DeviceState *d, const og_graph *g, const FrontierInfo &fi, RunTimers &tm, og_run_stats &st,
                        double &gather_ms) {
    const uint32_t ndyn = fi.n[3];
    const uint64_t D = fi.dyn_deg;
    if (!ndyn || !D) return;
    const uint32_t *d_dlist = d->fr.list + fi.n[0] + fi.n[1];
    const uint64_t eb = d->deb;
    OG_HIP(hipMemcpyAsync(d->h_dlist, d_dlist, (size_t)ndyn * 4, hipMemcpyDeviceToHost, d->s_copy));
    OG_HIP(hipMemcpyAsync(d->h_doff, d->fr.doff, ((size_t)ndyn + 1) * 8, hipMemcpyDeviceToHost, d->s_copy));
    OG_HIP(hipStreamSynchronize(d->s_copy));
    const uint64_t C = d->slot_dyn;
    const bool serial = (d->flags & OG_PLAN_NO_OVERLAP) != 0;
    for (uint64_t c0 = 0, c = 0; c0 < D; c0 += C, ++c) {
        const uint64_t c1 = std::min(D, c0 + C);
        const int slot = (int)(c % d->nslots);
        if (d->slot_h2d_pending[slot]) OG_HIP(hipEventSynchronize(d->ev_h2d[slot]));
        Stopwatch sw;
        gather(d, g, ndyn, c0, c1, (uint8_t *)d->h_stage[slot], d->run_pmid && fi.heavy);
        gather_ms += sw.ms();
        uint8_t *dslot = (uint8_t *)d->d_staging + (uint64_t)slot * d->slot_edges * d->eb;
        if (d->slot_relax_pending[slot]) OG_HIP(hipStreamWaitEvent(d->s_copy, d->ev_relax[slot], 0));
        {
            EventSpan span(d, &tm.copy, d->s_copy);
            OG_HIP(hipMemcpyAsync(dslot, d->h_stage[slot], (c1 - c0) * eb, hipMemcpyHostToDevice, d->s_copy));
        }
        OG_HIP(hipEventRecord(d->ev_h2d[slot], d->s_copy));
        d->slot_h2d_pending[slot] = true;
        OG_HIP(hipStreamWaitEvent(d->s_compute, d->ev_h2d[slot], 0));
        {
            EventSpan span(d, &tm.dyn, d->s_compute);
            launch_dyn_any<ALGO, W>(d, (const void *)dslot, c0, c1, c1 - c0, false);
        }
        OG_HIP(hipEventRecord(d->ev_relax[slot], d->s_compute));
        d->slot_relax_pending[slot] = true;
        if (serial) OG_HIP(hipStreamSynchronize(d->s_compute));
        st.h2d_bytes += (c1 - c0) * eb;
        st.chunks++;
    }
}

// Algorithmic bytes per relaxed edge: column (+ weight) and the target state.
template <int ALGO, bool W>
constexpr uint64_t edge_bytes_alg() {
    return ALGO == ALGO_MSBFS ? 12ull : (ALGO == ALGO_SSSP ? 12ull : 8ull);
}

constexpr int kRoundBatch = 4;     // device-driven rounds enqueued per host check
constexpr int kRoundBatchMax = 64;  // largest first batch (sized from the previous run)

// Iterates compaction -> static relax || dynamic relax -> after() until no
// vertex is active or held back.  Accumulates into st / tm.
//
// Without a host-side loader (whole graph in HBM, zero-copy dynamic region,
// USM) the rounds are device-driven: every kernel reads its counts from the
// round's FrontierInfo and the near-far threshold advances on the device, so
// rounds are enqueued a batch at a time and the host reads the traversal
// state once per batch (rounds enqueued past the end find no flags and do no
// work).  The first batch holds as many rounds as the previous run of the
// algorithm on this preparation took, rounded up to a multiple of
// kRoundBatch (a repeated traversal then needs one launch and no host round
// trip: each costs about 50 us); later batches kRoundBatch.  The staged loader needs
// the host in every round (gather, copies), so it synchronises on each
// round's summary.
// quiet_after: after() enqueues nothing (its rounds' timing spans may chain).
template <int ALGO, bool W, class After>
static void frontier_loop(og_graph *g, DeviceState *d, og_run_stats &st, RunTimers &tm, double &gather_ms,
                          double &overload_wall, After after, bool quiet_after) {
    hipStream_t s = d->s_compute;
    // SSSP near-far order: only vertices below the threshold relax in a
    // round; when none is left it moves to (smallest held-back value + delta).
    const uint32_t delta = ALGO == ALGO_SSSP ? d->tune.delta : 0;
    const uint32_t *nf_val = delta ? d->d_val : nullptr;
    const bool zero_copy = (d->flags & OG_PLAN_ZERO_COPY) != 0;
    const bool has_dyn = d->plan.vsplit < d->V;
    const bool staged = has_dyn && !zero_copy;
    // deferring dynamic vertices reorders relaxations across levels, which
    // only the atomicMin relaxes (SSSP, CC) tolerate
    const uint32_t defer = (ALGO == ALGO_SSSP || ALGO == ALGO_CC) && has_dyn ? d->tune.defer : 0;
    const bool dir_opt = ALGO == ALGO_BFS && d->d_in_ptr != nullptr;
    // light/heavy split: near-far SSSP on a graph prepared with it
    const bool split = ALGO == ALGO_SSSP && delta && d->d_pmid;
    d->run_pmid = split ? d->d_pmid : nullptr;
    const uint64_t *pm = d->run_pmid;
    frontier_begin(d->fr, s, delta, defer, dir_opt ? std::max<uint64_t>(d->E, 1) : 0, split);
    if (!staged) {
        const EagerSpan eager{d, tm};
        eager(kClsFront, s, [&] { frontier_compact(d, d->fr, s, nf_val, false, pm); });
        // one batch: kRoundBatch x (static || dynamic relax, after, compaction),
        // then the traversal state to the host
        // With both regions the static and dynamic relax of a round run as
        // one launch (k_relax_round); with OG_PLAN_NO_FUSE on two streams,
        // the dynamic relax enqueued before the static fork (a replay submits
        // the graph's kernels in capture order, and the link-bound relax
        // should not wait for the static launches).
        // A graph wholly in HBM runs its three static bins as one fused launch.
        const bool overlap = has_dyn && !(d->flags & OG_PLAN_NO_OVERLAP);
        const bool fuse = overlap && !(d->flags & OG_PLAN_NO_FUSE);
        const bool fuse_static = !has_dyn && !(d->flags & OG_PLAN_NO_FUSE);
        auto batch = [&](const auto &span, int rounds) {
            for (int r = 0; r < rounds; ++r) {
                bool joined = false;
                if (fuse) {
                    span(kClsDyn, s, [&] { launch_round_relax<ALGO, W>(d, true); });
                } else if (fuse_static) {
                    span(kClsStat, s, [&] { launch_round_relax<ALGO, W>(d, false); });
                } else if (overlap) {
                    OG_HIP(hipEventRecord(d->ev_fork, s));
                    span(kClsDyn, s, [&] { launch_dyn_any<ALGO, W>(d, d->hd_dyn, 0, ~0ull, 0, true); });
                    OG_HIP(hipStreamWaitEvent(d->s_static, d->ev_fork, 0));
                    span(kClsStat, d->s_static,
                         [&] { launch_static_relax<ALGO, W>(d, d->fr, nullptr, relax_args(d), d->s_static); });
                    OG_HIP(hipEventRecord(d->ev_static, d->s_static));
                    joined = true;
                } else {
                    joined = launch_static<ALGO, W>(d, nullptr, span);
                    if (has_dyn) span(kClsDyn, s, [&] { launch_dyn_any<ALGO, W>(d, d->hd_dyn, 0, ~0ull, 0, true); });
                }
                if (dir_opt)
                    span(kClsStat, s, [&] {
                        hipLaunchKernelGGL(k_bfs_pull, dim3(capped((d->V + kBlock - 1) / kBlock, kGroupGridCap)),
                                           dim3(kBlock), 0, s, d->d_in_ptr, d->d_in_cols, d->V, d->fr.info,
                                           d->fr.ctl, d->d_val, d->fr.flag, d->fr.counters + 0);
                        OG_HIP(hipGetLastError());
                    });
                if (joined) OG_HIP(hipStreamWaitEvent(s, d->ev_static, 0));
                after();
                span(kClsFront, s, [&] { frontier_compact(d, d->fr, s, nf_val, false, pm); });
            }
            OG_HIP(hipMemcpyAsync(d->fr.h_ctl, d->fr.ctl, sizeof(RoundCtl), hipMemcpyDeviceToHost, s));
        };
        if (!(d->flags & OG_PLAN_NO_GRAPH)) {
            // first batch: exactly the pairs the previous run took (a 6-round
            // BFS replayed 8 rounds when this was rounded up to a multiple of
            // kRoundBatch, two empty relax + compaction pairs per run)
            const uint32_t guess = std::min<uint32_t>(d->last_pairs[ALGO], kRoundBatchMax);
            int rounds = guess ? (int)guess : kRoundBatch;
            for (;;) {
                const uint64_t key = (uint64_t)ALGO | (uint64_t)W << 4 | (uint64_t)(delta != 0) << 5 |
                                     (uint64_t)has_dyn << 6 | (uint64_t)dir_opt << 7 | (uint64_t)split << 16 |
                                     (uint64_t)rounds << 8;
                RoundGraph *rg = round_graph_for(d, key, [&](RoundGraph *n) {
                    GraphSpan gs{n};
                    gs.chain = (fuse || fuse_static) && quiet_after;  // one stream: relax span(s), then compaction
                    batch(gs, rounds);
                });
                rounds = kRoundBatch;
                OG_HIP(hipGraphLaunch(rg->exec, s));
                OG_HIP(hipStreamSynchronize(s));
                for (const auto &m : rg->marks) {
                    float ms = 0;
                    if (hipEventElapsedTime(&ms, m.a, m.b) == hipSuccess) {
                        tm.at(m.cls).ms += ms;
                        tm.at(m.cls).launches++;
                    } else {
                        (void)hipGetLastError();
                    }
                }
                if (!d->fr.h_ctl->live) break;
            }
            // relax + compaction pairs of this run, for the next run's first batch
            d->last_pairs[ALGO] = d->fr.h_ctl->compactions ? d->fr.h_ctl->compactions - 1 : 0;
        } else {
            for (;;) {
                batch(eager, kRoundBatch);
                OG_HIP(hipStreamSynchronize(s));
                if (!d->fr.h_ctl->live) break;
            }
        }
    } else {
        {
            EventSpan span(d, &tm.front, s);
            frontier_compact(d, d->fr, s, nf_val, true, pm);
        }
        OG_HIP(hipStreamSynchronize(s));
        for (;;) {
            const FrontierInfo fi = *d->fr.h_info;
            if (!fi.live) break;
            const uint64_t active = (uint64_t)fi.n[0] + fi.n[1] + fi.nhub + fi.n[3] + fi.nzero;
            if (active) {  // else only the near-far threshold moved
                if (fi.n[2] > d->fr.hub_cap)
                    raise(OG_EHIP, "hub tile list overflow (%u > %llu)", fi.n[2], (unsigned long long)d->fr.hub_cap);
                const bool joined = launch_static<ALGO, W>(d, &fi, EagerSpan{d, tm});
                if (fi.n[3]) {
                    Stopwatch ow;
                    run_dynamic<ALGO, W>(d, g, fi, tm, st, gather_ms);
                    overload_wall += ow.ms();
                }
                if (joined) OG_HIP(hipStreamWaitEvent(s, d->ev_static, 0));
                after();
            }
            {
                EventSpan span(d, &tm.front, s);
                frontier_compact(d, d->fr, s, nf_val, true, pm);
            }
            OG_HIP(hipStreamSynchronize(s));
        }
        OG_HIP(hipMemcpyAsync(d->fr.h_ctl, d->fr.ctl, sizeof(RoundCtl), hipMemcpyDeviceToHost, s));
        OG_HIP(hipStreamSynchronize(s));
    }
    const RoundCtl c = *d->fr.h_ctl;
    if (c.overflow) raise(OG_EUNSUPPORTED, "multi-source BFS deeper than 254 levels");
    st.iterations += c.rounds;
    st.node_sum += c.node_sum;
    st.static_node_sum += c.static_node_sum;
    st.work_edges += c.work_edges;
    st.static_work_edges += c.static_work_edges;
    // per edge: the edge in the device format (d->eb bytes) + the target state
    const uint64_t per_edge = edge_bytes_alg<ALGO, W>() - (W ? 8ull : 4ull) + d->eb;
    const uint64_t dyn_edges = c.work_edges - c.static_work_edges;
    st.relax_static_bytes += 24ull * c.static_node_sum + per_edge * c.static_work_edges;
    st.relax_dynamic_bytes += 24ull * c.dyn_node_sum + per_edge * dyn_edges;
    st.frontier_bytes += 2ull * d->Vpad * c.compactions + 4ull * c.node_sum;
    if (has_dyn && zero_copy) {  // every active dynamic edge read in place (HBM-resident ones subtracted later)
        st.h2d_bytes += dyn_edges * d->deb;
        st.chunks += c.dyn_rounds;
    }
}

static void finish_stats(og_run_stats &st, RunTimers &tm, double gather_ms, double overload_wall,
                         const unsigned long long *ctr, uint64_t upd_bytes, uint64_t link_eb) {
    st.updates = ctr[0] + ctr[1];
    // dynamic edges served from the HBM-resident head did not cross the link
    st.h2d_bytes -= std::min<uint64_t>(st.h2d_bytes, ctr[5] * link_eb);
    st.relax_static_bytes += upd_bytes * ctr[0];
    st.relax_dynamic_bytes += upd_bytes * ctr[1];
    st.frontier_kernel_ms = sum_ms(tm.front);
    st.frontier_launches = tm.front.launches;
    st.relax_static_kernel_ms = sum_ms(tm.stat);
    st.relax_static_launches = tm.stat.launches;
    st.relax_dynamic_kernel_ms = sum_ms(tm.dyn);
    st.relax_dynamic_launches = tm.dyn.launches;
    st.copy_ms = sum_ms(tm.copy);
    st.pre_ms = st.frontier_kernel_ms;
    st.static_ms = st.relax_static_kernel_ms;
    st.overload_ms = overload_wall;
    st.overload_move_ms = st.copy_ms;
    st.gather_ms = gather_ms;
}

template <int ALGO, bool W>
static void run_frontier(og_graph *g, uint32_t src, uint32_t *out, og_run_stats *stats, DeviceState *dv = nullptr) {
    DeviceState *d = dv ? dv : g->dev;
    const uint64_t V = d->V;
    og_run_stats st{};
    st.approach = (uint32_t)d->approach;
    RunTimers tm;
    d->events.reset();
    double gather_ms = 0, overload_wall = 0;
    hipStream_t s = d->s_compute;

    OG_HIP(hipStreamSynchronize(d->s_copy));
    OG_HIP(hipStreamSynchronize(d->s_static));
    OG_HIP(hipStreamSynchronize(s));
    Stopwatch total;
    OG_HIP(hipMemsetAsync(d->fr.counters, 0, 16 * sizeof(unsigned long long), s));
    if (ALGO == ALGO_CC) {
        hipLaunchKernelGGL(k_iota, dim3(fill_grid(V)), dim3(256), 0, s, d->d_val, V);
        hipLaunchKernelGGL(k_fill_u8, dim3(fill_grid(V)), dim3(256), 0, s, d->fr.flag, V, (uint8_t)1);
    } else {
        hipLaunchKernelGGL(k_fill_u8, dim3(fill_grid(d->Vpad)), dim3(256), 0, s, d->fr.flag, d->Vpad, (uint8_t)0);
        hipLaunchKernelGGL(k_fill_u32, dim3(fill_grid(V)), dim3(256), 0, s, d->d_val, V, OG_UNREACHED);
        hipLaunchKernelGGL(k_set_source, dim3(1), dim3(1), 0, s, d->d_val, d->fr.flag, src);
    }
    OG_HIP(hipGetLastError());
    frontier_loop<ALGO, W>(g, d, st, tm, gather_ms, overload_wall, []() {}, true);
    OG_HIP(hipStreamSynchronize(s));
    st.total_ms = total.ms();

    {
        // one lane per four vertices with the degree table, at most 2048
        // workgroups (three atomics each on the same three counters: 16384
        // workgroups spent 0.6 ms on them)
        const uint64_t lanes = d->d_deg ? (V + 3) / 4 : V;
        const unsigned grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(2048, (lanes + kBlock - 1) / kBlock));
        hipLaunchKernelGGL(k_reach_stats, dim3(grid), dim3(kBlock), 0, s, d->d_val, d->d_ptr, d->d_deg,
                           d->deg_heavy(), V,
                           d->fr.counters + 2);
    }
    OG_HIP(hipGetLastError());
    OG_HIP(hipMemcpyAsync(d->fr.h_counters, d->fr.counters, 16 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    if (out) copy_to_host(d, out, d->d_val, V * sizeof(uint32_t), s);
    OG_HIP(hipStreamSynchronize(s));
    st.reached = d->fr.h_counters[2];
    st.traversed_edges = d->fr.h_counters[3];
    st.checksum = d->fr.h_counters[4];
    finish_stats(st, tm, gather_ms, overload_wall, d->fr.h_counters, 8, d->deb);
    if (stats) *stats = st;
}

// Multi-source BFS: sources in batches of 64; per vertex, u64 masks of the
// sources that reached it (visit), whose frontier holds it (front) and that
// were offered to it this round (next); levels as u8 per (vertex, source).
template <bool W>
static void run_msbfs(og_graph *g, const uint32_t *sources, uint32_t n, uint32_t *levels, uint64_t *checksums,
                      uint64_t *reached, og_run_stats *stats) {
    DeviceState *d = g->dev;
    const uint64_t V = d->V;
    hipStream_t s = d->s_compute;
    if (!d->d_visit) {
        d->d_visit = (uint64_t *)d->dalloc(V * 8);
        d->d_front = (uint64_t *)d->dalloc(V * 8);
        d->d_next = (uint64_t *)d->dalloc(V * 8);
        d->d_mlvl = (uint8_t *)d->dalloc(V * kMsWidth);
        d->d_msrc = (uint32_t *)d->dalloc(kMsWidth * 4);
        d->d_msum = (unsigned long long *)d->dalloc(3 * kMsWidth * 8);
    }
    std::vector<unsigned long long> hsum(3 * kMsWidth);
    og_run_stats st{};
    st.approach = (uint32_t)d->approach;
    RunTimers tm;
    d->events.reset();
    double gather_ms = 0, overload_wall = 0;
    OG_HIP(hipStreamSynchronize(d->s_copy));
    OG_HIP(hipStreamSynchronize(d->s_static));
    OG_HIP(hipStreamSynchronize(s));
    Stopwatch total;
    OG_HIP(hipMemsetAsync(d->fr.counters, 0, 16 * sizeof(unsigned long long), s));
    double device_ms = 0;
    for (uint32_t b0 = 0; b0 < n; b0 += kMsWidth) {
        const uint32_t nb = std::min<uint32_t>(kMsWidth, n - b0);
        OG_HIP(hipMemsetAsync(d->d_visit, 0, V * 8, s));
        OG_HIP(hipMemsetAsync(d->d_front, 0, V * 8, s));
        OG_HIP(hipMemsetAsync(d->d_next, 0, V * 8, s));
        OG_HIP(hipMemsetAsync(d->d_mlvl, 0, V * kMsWidth, s));
        OG_HIP(hipMemsetAsync(d->fr.flag, 0, d->Vpad, s));
        OG_HIP(hipMemcpyAsync(d->d_msrc, sources + b0, nb * 4, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_ms_init, dim3(1), dim3(kMsWidth), 0, s, d->d_msrc, nb, d->d_visit, d->d_front,
                           d->fr.flag, d->d_mlvl);
        OG_HIP(hipGetLastError());
        const unsigned cgrid = (unsigned)std::min<uint64_t>(4096, (V + kBlock - 1) / kBlock);
        frontier_loop<ALGO_MSBFS, W>(g, d, st, tm, gather_ms, overload_wall, [&]() {
            hipLaunchKernelGGL(k_ms_commit, dim3(cgrid), dim3(kBlock), 0, s, d->fr.flag, V, d->d_visit, d->d_front,
                               d->d_next, d->d_mlvl, d->fr.ctl);
            OG_HIP(hipGetLastError());
        }, false);
        OG_HIP(hipMemsetAsync(d->d_msum, 0, 3 * kMsWidth * 8, s));
        hipLaunchKernelGGL(k_ms_digest, dim3((unsigned)std::min<uint64_t>(1024, (V + kBlock - 1) / kBlock)),
                           dim3(kBlock), 0, s, d->d_mlvl, d->d_ptr, V, d->d_msum);
        OG_HIP(hipGetLastError());
        OG_HIP(hipMemcpyAsync(hsum.data(), d->d_msum, 3 * kMsWidth * 8, hipMemcpyDeviceToHost, s));
        OG_HIP(hipStreamSynchronize(s));
        device_ms = total.ms();
        for (uint32_t j = 0; j < nb; ++j) {
            if (checksums) checksums[b0 + j] = hsum[j];
            if (reached) reached[b0 + j] = hsum[kMsWidth + j];
            st.reached += hsum[kMsWidth + j];
            st.traversed_edges += hsum[2 * kMsWidth + j];
            st.checksum += hsum[j];
        }
        if (levels) {
            if (!d->d_extract) d->d_extract = (uint32_t *)d->dalloc(V * 4);
            for (uint32_t j = 0; j < nb; ++j) {
                hipLaunchKernelGGL(k_ms_extract, dim3(fill_grid(V)), dim3(256), 0, s, d->d_mlvl, V, j, d->d_extract);
                OG_HIP(hipGetLastError());
                copy_to_host(d, levels + (uint64_t)(b0 + j) * V, d->d_extract, V * 4, s);
            }
            OG_HIP(hipStreamSynchronize(s));
        }
    }
    st.total_ms = device_ms;
    OG_HIP(hipMemcpyAsync(d->fr.h_counters, d->fr.counters, 16 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    OG_HIP(hipStreamSynchronize(s));
    finish_stats(st, tm, gather_ms, overload_wall, d->fr.h_counters, 9, d->deb);
    if (stats) *stats = st;
}

DeviceState *lane_state_create(og_graph *g, uint32_t delta) {
    const DeviceState *d = g->dev;
    if (d->plan.vsplit < d->V || d->managed) raise(OG_EUNSUPPORTED, "a lane needs the whole graph in device memory");
    std::unique_ptr<DeviceState, void (*)(DeviceState *)> l(new DeviceState, device_state_free);
    // the graph as prepared (read only: every array below belongs to g->dev)
    l->approach = d->approach;
    l->device = d->device;
    l->V = d->V;
    l->E = d->E;
    l->Vpad = d->Vpad;
    l->eb = d->eb;
    l->deb = d->deb;
    l->plan = d->plan;
    l->tune = d->tune;
    if (delta) l->tune.delta = delta;
    l->host_threads = d->host_threads;
    l->flags = d->flags;
    l->d_ptr = d->d_ptr;
    l->d_pmid = d->d_pmid;
    l->d_deg = d->d_deg;
    l->deg_split = d->deg_split;
    l->split_weight = d->split_weight;
    l->d_static = d->d_static;
    l->static_used = d->static_used;
    l->packed = d->packed;
    l->wbits = d->wbits;
    l->d_in_ptr = d->d_in_ptr;
    l->d_in_cols = d->d_in_cols;
    // the lane's own run state, on one stream: a graph wholly in HBM has no
    // dynamic region to copy or overlap, and every stream of a process takes
    // a hardware queue (or shares one) while the server runs
    OG_HIP(hipSetDevice(d->device));
    OG_HIP(hipStreamCreateWithFlags(&l->s_compute, hipStreamNonBlocking));
    l->s_copy = l->s_static = l->s_compute;
    OG_HIP(hipEventCreateWithFlags(&l->ev_static, hipEventDisableTiming));
    OG_HIP(hipEventCreateWithFlags(&l->ev_fork, hipEventDisableTiming));
    l->d_val = (uint32_t *)l->dalloc(l->V * sizeof(uint32_t));
    frontier_alloc(l.get(), g, l->fr, l->mem);
    OG_HIP(hipDeviceSynchronize());
    return l.release();
}

uint64_t lane_state_bytes(const og_graph *g) {
    const DeviceState *d = g->dev;
    const uint64_t V = d->V, Vpad = d->Vpad;
    const uint64_t nblk = (Vpad + verts_per_block(frontier_fv(V)) - 1) / verts_per_block(frontier_fv(V));
    const uint64_t nsb = (nblk + kSbBlocks - 1) / kSbBlocks;
    const uint64_t hub = std::max<uint64_t>(hub_capacity(g, d->plan.vsplit, d->tune), 1);
    // what lane_state_create and frontier_alloc request: values, flags (and
    // the heavy-round flags with the split), bin lists, degree prefix, the
    // (empty) dynamic chunk table, block counts, super-block sums and
    // prefixes, round summary, control block, counters and hub tiles
    return 4 * V + Vpad * (d->d_pmid ? 2 : 1) + 4 * V + 8 * (V + 1) + 4 * 2 + 4ull * kBcntRows * nblk + 8 * nblk +
           8ull * kSbVals * nsb + 24 * nsb + sizeof(FrontierInfo) + sizeof(RoundCtl) + 16 * 8 + 8 * hub;
}

void lane_run_sssp(og_graph *g, DeviceState *lane, uint32_t src, uint32_t *out, og_run_stats *st) {
    run_frontier<ALGO_SSSP, true>(g, src, out, st, lane);
}

static og_status run_algo(og_graph *g, int algo, uint32_t src, uint32_t *out, og_run_stats *stats) {
    clear_error();
    if (!g) return fail(OG_EINVAL, "graph is NULL");
    if (!g->dev) return fail(OG_ESTATE, "og_prepare has not been called for this graph");
    if (algo != ALGO_CC && src >= g->V) return fail(OG_EINVAL, "source %u >= V=%llu", src, (unsigned long long)g->V);
    if (g->kind == OG_KIND_PULL) return fail(OG_EUNSUPPORTED, "pull-layout graphs serve PageRank only");
    if (algo == ALGO_SSSP && g->kind != OG_KIND_WEIGHTED) return fail(OG_EUNSUPPORTED, "SSSP needs a weighted graph");
    int expect = 0;
    if (!g->busy.compare_exchange_strong(expect, 1)) return fail(OG_EBUSY, "graph is running another call");
    og_status st = guard([&]() -> og_status {
        OG_HIP(hipSetDevice(g->dev->device));
        const bool w = g->kind == OG_KIND_WEIGHTED;
        if (algo == ALGO_BFS) {
            if (w) run_frontier<ALGO_BFS, true>(g, src, out, stats);
            else run_frontier<ALGO_BFS, false>(g, src, out, stats);
        } else if (algo == ALGO_SSSP) {
            run_frontier<ALGO_SSSP, true>(g, src, out, stats);
        } else {
            if (w) run_frontier<ALGO_CC, true>(g, 0, out, stats);
            else run_frontier<ALGO_CC, false>(g, 0, out, stats);
        }
        return OG_OK;
    });
    g->busy.store(0);
    return st;
}

}  // namespace og

using namespace og;

extern "C" og_status og_prepare(og_graph *g, og_approach approach, const og_plan_params *params, og_plan_info *info) {
    clear_error();
    if (!g) return fail(OG_EINVAL, "graph is NULL");
    int expect = 0;
    if (!g->busy.compare_exchange_strong(expect, 1)) return fail(OG_EBUSY, "graph is running another call");
    og_status st = guard([&]() -> og_status {
        prepare_impl(g, approach, params, info);
        return OG_OK;
    });
    g->busy.store(0);
    return st;
}

extern "C" og_status og_release(og_graph *g) {
    clear_error();
    if (!g) return fail(OG_EINVAL, "graph is NULL");
    int expect = 0;
    if (!g->busy.compare_exchange_strong(expect, 1)) return fail(OG_EBUSY, "graph is running another call");
    device_state_free(g->dev);
    g->dev = nullptr;
    g->busy.store(0);
    return OG_OK;
}

extern "C" og_status og_bfs(og_graph *g, uint32_t source, uint32_t *levels, og_run_stats *stats) {
    return run_algo(g, ALGO_BFS, source, levels, stats);
}

extern "C" og_status og_sssp(og_graph *g, uint32_t source, uint32_t *dist, og_run_stats *stats) {
    return run_algo(g, ALGO_SSSP, source, dist, stats);
}

extern "C" og_status og_cc(og_graph *g, uint32_t *labels, og_run_stats *stats) {
    return run_algo(g, ALGO_CC, 0, labels, stats);
}

extern "C" og_status og_msbfs(og_graph *g, const uint32_t *sources, uint32_t num_sources, uint32_t *levels,
                              uint64_t *checksums, uint64_t *reached, og_run_stats *stats) {
    clear_error();
    if (!g || (!sources && num_sources)) return fail(OG_EINVAL, "graph/sources is NULL");
    if (!g->dev) return fail(OG_ESTATE, "og_prepare has not been called for this graph");
    if (g->kind == OG_KIND_PULL) return fail(OG_EUNSUPPORTED, "pull-layout graphs serve PageRank only");
    for (uint32_t i = 0; i < num_sources; ++i)
        if (sources[i] >= g->V) return fail(OG_EINVAL, "source %u >= V", sources[i]);
    int expect = 0;
    if (!g->busy.compare_exchange_strong(expect, 1)) return fail(OG_EBUSY, "graph is running another call");
    og_status st = guard([&]() -> og_status {
        OG_HIP(hipSetDevice(g->dev->device));
        if (g->kind == OG_KIND_WEIGHTED) run_msbfs<true>(g, sources, num_sources, levels, checksums, reached, stats);
        else run_msbfs<false>(g, sources, num_sources, levels, checksums, reached, stats);
        return OG_OK;
    });
    g->busy.store(0);
    return st;
}

// ------------------------------------------------------------ link probe

namespace og {

__global__ __launch_bounds__(kBlock) void k_probe_contig(const uint32_t *__restrict__ p, uint64_t n, uint32_t *out) {
    uint32_t acc = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) acc ^= p[i];
    if (acc == 0x9E3779B9u) out[0] = acc;  // keeps the reads alive
}

// Rows r = [off[r], off[r+1]) of the concatenated space start at base[r] in p;
// lane q finds its row by binary search (as k_relax_dyn does).
__global__ __launch_bounds__(kBlock) void k_probe_rows(const uint32_t *__restrict__ p,
                                                       const uint64_t *__restrict__ off,
                                                       const uint64_t *__restrict__ base, uint32_t nrows, uint64_t m,
                                                       uint32_t *out) {
    uint32_t acc = 0;
    for (uint64_t q = (uint64_t)blockIdx.x * kBlock + threadIdx.x; q < m; q += (uint64_t)gridDim.x * kBlock) {
        uint32_t lo = 0, hi = nrows;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (off[mid] <= q) lo = mid; else hi = mid;
        }
        acc ^= p[base[lo] + (q - off[lo])];
    }
    if (acc == 0x9E3779B9u) out[0] = acc;
}

}  // namespace og

namespace og {

// Random 4-byte gathers from a table of n words: each lane draws its indices
// with SplitMix64 from its own counter and keeps kGatherIlp loads in flight.
constexpr int kGatherIlp = 8;
__global__ __launch_bounds__(kBlock) void k_probe_gather(const uint32_t *__restrict__ t, uint64_t n, uint64_t per,
                                                         uint64_t seed, uint32_t *out) {
    const uint64_t tid = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    uint32_t acc = 0;
    for (uint64_t i = 0; i < per; i += kGatherIlp) {
        uint32_t x[kGatherIlp];
#pragma unroll
        for (int k = 0; k < kGatherIlp; ++k) x[k] = t[d_mix64(seed + tid * per + i + k) % n];
#pragma unroll
        for (int k = 0; k < kGatherIlp; ++k) acc ^= x[k];
    }
    if (acc == 0x9E3779B9u) out[0] = acc;  // keeps the reads alive
}

}  // namespace og

// Random-gather probe (no reference counterpart: a measurement helper for the
// relax kernels, whose per-edge target reads are random 4-byte accesses).
extern "C" og_status og_probe_hbm_gather(uint64_t table_bytes, double *gather_gbs, double *mgathers_per_s) {
    clear_error();
    if (table_bytes < (1u << 16) || table_bytes > (64ull << 30)) return fail(OG_EINVAL, "table must be 64 KiB .. 64 GiB");
    return guard([&]() -> og_status {
        if (!hip_available()) raise(OG_EHIP, "no HIP device is visible");
        MemPool m;
        const uint64_t n = table_bytes / 4;
        uint32_t *t = (uint32_t *)m.dalloc(n * 4);
        uint32_t *dout = (uint32_t *)m.dalloc(4);
        hipLaunchKernelGGL(k_iota, dim3(fill_grid(n)), dim3(256), 0, 0, t, n);
        const unsigned grid = 8192;
        const uint64_t per = 8 * kGatherIlp;  // gathers per lane
        const double total = (double)grid * kBlock * per;
        hipEvent_t a, b;
        OG_HIP(hipEventCreate(&a));
        OG_HIP(hipEventCreate(&b));
        hipLaunchKernelGGL(k_probe_gather, dim3(grid), dim3(kBlock), 0, 0, t, n, per, 1ull, dout);
        OG_HIP(hipDeviceSynchronize());
        OG_HIP(hipEventRecord(a, 0));
        for (int r = 0; r < 3; ++r)
            hipLaunchKernelGGL(k_probe_gather, dim3(grid), dim3(kBlock), 0, 0, t, n, per, 2ull + r, dout);
        OG_HIP(hipEventRecord(b, 0));
        OG_HIP(hipEventSynchronize(b));
        OG_HIP(hipGetLastError());
        float ms = 0;
        OG_HIP(hipEventElapsedTime(&ms, a, b));
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        const double per_s = 3.0 * total / (ms * 1e-3);
        if (gather_gbs) *gather_gbs = per_s * 4 / 1e9;
        if (mgathers_per_s) *mgathers_per_s = per_s / 1e6;
        return OG_OK;
    });
}

extern "C" og_status og_probe_host_link(uint64_t bytes, double *contig_gbs, double *scattered_gbs,
                                        double *memcpy_gbs) {
    clear_error();
    if (bytes < (1u << 20) || bytes > (64ull << 30)) return fail(OG_EINVAL, "probe size must be 1 MiB .. 64 GiB");
    return guard([&]() -> og_status {
        if (!hip_available()) raise(OG_EHIP, "no HIP device is visible");
        MemPool m;
        const uint64_t n = bytes / 4;
        uint32_t *h = (uint32_t *)m.palloc(n * 4);
        parallel_for(n, 1 << 20, 0, [&](uint64_t b, uint64_t e, unsigned) {
            for (uint64_t i = b; i < e; ++i) h[i] = (uint32_t)i;
        });
        void *hd = nullptr;
        OG_HIP(hipHostGetDevicePointer(&hd, h, 0));
        // rows of 1..19 words with gaps of 0..39 words (SplitMix64 draws)
        std::vector<uint64_t> off{0}, base;
        uint64_t pos = 0, key = rng_key(7);
        for (uint64_t r = 0;; ++r) {
            const uint64_t x = rng_draw(key, r);
            const uint64_t len = 1 + x % 19, gap = (x >> 32) % 40;
            if (pos + gap + len > n) break;
            pos += gap;
            base.push_back(pos);
            pos += len;
            off.push_back(off.back() + len);
        }
        const uint32_t nrows = (uint32_t)base.size();
        const uint64_t mwords = off.back();
        uint64_t *doff = (uint64_t *)m.dalloc(off.size() * 8), *dbase = (uint64_t *)m.dalloc(base.size() * 8 + 8);
        uint32_t *dout = (uint32_t *)m.dalloc(4), *ddst = (uint32_t *)m.dalloc(n * 4);
        OG_HIP(hipMemcpy(doff, off.data(), off.size() * 8, hipMemcpyHostToDevice));
        OG_HIP(hipMemcpy(dbase, base.data(), base.size() * 8, hipMemcpyHostToDevice));
        hipEvent_t a, b;
        OG_HIP(hipEventCreate(&a));
        OG_HIP(hipEventCreate(&b));
        auto timed = [&](double useful, auto fn) {
            fn();
            OG_HIP(hipDeviceSynchronize());
            OG_HIP(hipEventRecord(a, 0));
            for (int r = 0; r < 3; ++r) fn();
            OG_HIP(hipEventRecord(b, 0));
            OG_HIP(hipEventSynchronize(b));
            float ms = 0;
            OG_HIP(hipEventElapsedTime(&ms, a, b));
            return 3.0 * useful / (ms * 1e-3) / 1e9;
        };
        const double c = timed((double)n * 4, [&] {
            hipLaunchKernelGGL(k_probe_contig, dim3(4096), dim3(kBlock), 0, 0, (const uint32_t *)hd, n, dout);
        });
        const double sc = timed((double)mwords * 4, [&] {
            hipLaunchKernelGGL(k_probe_rows, dim3(4096), dim3(kBlock), 0, 0, (const uint32_t *)hd, doff, dbase, nrows,
                               mwords, dout);
        });
        const double mc = timed((double)n * 4, [&] { OG_HIP(hipMemcpyAsync(ddst, h, n * 4, hipMemcpyHostToDevice, 0)); });
        OG_HIP(hipGetLastError());
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        if (contig_gbs) *contig_gbs = c;
        if (scattered_gbs) *scattered_gbs = sc;
        if (memcpy_gbs) *memcpy_gbs = mc;
        return OG_OK;
    });
}

// Host digest of a result array, equal to the one the device computes:
// sum (mod 2^64) of mix64(v << 32 | value[v]) over v with value != UNREACHED.
extern "C" og_status og_levels_checksum(const uint32_t *values, uint64_t num_vertices, uint64_t *checksum) {
    if (!values || !checksum) return fail(OG_EINVAL, "null argument");
    uint64_t h = 0;
    for (uint64_t v = 0; v < num_vertices; ++v)
        if (values[v] != OG_UNREACHED) h += mix64((v << 32) | values[v]);
    *checksum = h;
    return OG_OK;
}

// ------------------------------------------------------------------ tuning

namespace og {

static og_status check_tuning(const og_tuning *t) {
    const uint32_t l = t->lpv0;
    if (!(l == 1 || l == 2 || l == 4 || l == 8 || l == 16 || l == 32 || l == 64))
        return fail(OG_EINVAL, "lpv0 must be a power of two in [1, 64]");
    if (t->t1 < 1 || t->t2 < t->t1) return fail(OG_EINVAL, "need 1 <= t1 <= t2");
    if (t->tile < 64 || t->tile > 4096) return fail(OG_EINVAL, "tile must be in [64, 4096]");
    if (t->defer > 2) return fail(OG_EINVAL, "defer must be 0, 1 or 2");
    if (t->dyn_grid > 65535) return fail(OG_EINVAL, "dyn_grid must be at most 65535");
    return OG_OK;
}

static void apply_tuning(og_graph *g, const og_tuning *t) {
    DeviceState *d = g->dev;
    d->tune.t1 = t->t1;
    d->tune.lpv0 = t->lpv0;
    d->tune.t2 = t->t2;
    d->tune.tile = t->tile;
    d->tune.delta = t->delta;
    d->tune.defer = t->defer;
    d->tune.dyn_grid = t->dyn_grid;
    const uint64_t cap = hub_capacity(g, d->plan.vsplit, d->tune);
    if (cap > d->fr.hub_cap) {
        d->fr.hub = (uint64_t *)d->dalloc(cap * sizeof(uint64_t));
        d->fr.hub_cap = cap;
    }
    d->pr_plan.reset();      // PageRank bins follow t1 / t2
    d->round_graph.reset();  // captured rounds hold the tuning and the hub list
}

static og_tuning tuning_of(const DeviceState *d) {
    return og_tuning{d->tune.t1, d->tune.lpv0, d->tune.t2, d->tune.tile, d->tune.delta, d->tune.defer,
                     d->tune.dyn_grid};
}

// The multi-source tuning batch: 64 sources drawn uniformly (SplitMix64
// seeded with `seed`) among the vertices with out-edges.
static std::vector<uint32_t> tune_sources(const og_graph *g, uint32_t seed) {
    std::vector<uint32_t> cand;
    for (uint64_t v = 0; v < g->V; ++v)
        if (g->ptr[v + 1] > g->ptr[v]) cand.push_back((uint32_t)v);
    std::vector<uint32_t> out;
    if (cand.empty()) return out;
    const uint64_t key = rng_key(seed);
    for (uint32_t i = 0; i < (uint32_t)kMsWidth; ++i) out.push_back(cand[rng_draw(key, i) % cand.size()]);
    return out;
}

static double time_algo(og_graph *g, uint32_t algo, uint32_t src, uint32_t runs) {
    std::vector<double> ts;
    const bool w = g->kind == OG_KIND_WEIGHTED;
    const std::vector<uint32_t> batch = algo == OG_TUNE_MSBFS ? tune_sources(g, src) : std::vector<uint32_t>();
    for (uint32_t r = 0; r < runs; ++r) {
        og_run_stats st;
        if (algo == OG_TUNE_MSBFS) {
            if (w) run_msbfs<true>(g, batch.data(), (uint32_t)batch.size(), nullptr, nullptr, nullptr, &st);
            else run_msbfs<false>(g, batch.data(), (uint32_t)batch.size(), nullptr, nullptr, nullptr, &st);
        } else if (algo == OG_TUNE_BFS) {
            if (w) run_frontier<ALGO_BFS, true>(g, src, nullptr, &st);
            else run_frontier<ALGO_BFS, false>(g, src, nullptr, &st);
        } else if (algo == OG_TUNE_SSSP) {
            run_frontier<ALGO_SSSP, true>(g, src, nullptr, &st);
        } else {
            if (w) run_frontier<ALGO_CC, true>(g, 0, nullptr, &st);
            else run_frontier<ALGO_CC, false>(g, 0, nullptr, &st);
        }
        ts.push_back(st.total_ms);
    }
    std::sort(ts.begin(), ts.end());
    return ts[ts.size() / 2];
}

}  // namespace og

extern "C" og_status og_get_tuning(const og_graph *g, og_tuning *t) {
    clear_error();
    if (!g || !t) return fail(OG_EINVAL, "null argument");
    if (!g->dev) return fail(OG_ESTATE, "og_prepare has not been called for this graph");
    *t = tuning_of(g->dev);
    return OG_OK;
}

extern "C" og_status og_set_tuning(og_graph *g, const og_tuning *t) {
    clear_error();
    if (!g || !t) return fail(OG_EINVAL, "null argument");
    if (!g->dev) return fail(OG_ESTATE, "og_prepare has not been called for this graph");
    og_status st = check_tuning(t);
    if (st != OG_OK) return st;
    int expect = 0;
    if (!g->busy.compare_exchange_strong(expect, 1)) return fail(OG_EBUSY, "graph is running another call");
    st = guard([&]() -> og_status {
        apply_tuning(g, t);
        return OG_OK;
    });
    g->busy.store(0);
    return st;
}

extern "C" og_status og_autotune(og_graph *g, uint32_t algo, uint32_t source, uint32_t runs, og_tuning *best,
                                 double *best_ms) {
    clear_error();
    if (!g) return fail(OG_EINVAL, "graph is NULL");
    if (!g->dev) return fail(OG_ESTATE, "og_prepare has not been called for this graph");
    if (algo > OG_TUNE_MSBFS)
        return fail(OG_EINVAL, "algo must be OG_TUNE_BFS, OG_TUNE_SSSP, OG_TUNE_CC or OG_TUNE_MSBFS");
    if (g->kind == OG_KIND_PULL) return fail(OG_EUNSUPPORTED, "pull-layout graphs serve PageRank only");
    if (algo == OG_TUNE_SSSP && g->kind != OG_KIND_WEIGHTED) return fail(OG_EUNSUPPORTED, "SSSP needs weights");
    if ((algo == OG_TUNE_BFS || algo == OG_TUNE_SSSP) && source >= g->V) return fail(OG_EINVAL, "source out of range");
    if (runs == 0) runs = 1;
    int expect = 0;
    if (!g->busy.compare_exchange_strong(expect, 1)) return fail(OG_EBUSY, "graph is running another call");
    og_status st = guard([&]() -> og_status {
        OG_HIP(hipSetDevice(g->dev->device));
        og_tuning cur = tuning_of(g->dev);
        (void)time_algo(g, algo, source, 1);  // warm-up
        double cur_ms = time_algo(g, algo, source, runs);
        auto trial = [&](og_tuning cand) {
            if (check_tuning(&cand) != OG_OK) return;
            apply_tuning(g, &cand);
            const double ms = time_algo(g, algo, source, runs);
            if (ms < cur_ms) {
                cur_ms = ms;
                cur = cand;
            }
            apply_tuning(g, &cur);
        };
        const bool has_dyn = g->dev->plan.vsplit < g->V;
        for (int pass = 0; pass < 2; ++pass) {
            if ((algo == OG_TUNE_SSSP || algo == OG_TUNE_CC) && has_dyn)
                for (uint32_t df : {0u, 1u, 2u}) {
                    if (df == 2 && algo != OG_TUNE_SSSP) continue;
                    og_tuning c = cur;
                    c.defer = df;
                    trial(c);
                }
            if (algo == OG_TUNE_SSSP)
                for (uint32_t dl : {0u, 8u, 16u, 32u, 64u, 128u, 256u, 512u}) {
                    og_tuning c = cur;
                    c.delta = dl;
                    trial(c);
                }
            for (uint32_t t1 : {4u, 8u, 16u, 32u, 64u})
                for (uint32_t l : {2u, 4u, 8u, 16u}) {
                    og_tuning c = cur;
                    c.t1 = t1;
                    c.lpv0 = l;
                    if (c.t2 < t1) c.t2 = t1;
                    trial(c);
                }
            for (uint32_t t2 : {512u, 1024u, 2048u, 4096u, 16384u, 65536u}) {
                og_tuning c = cur;
                c.t2 = std::max(t2, c.t1);
                trial(c);
            }
            for (uint32_t tile : {512u, 1024u, 2048u, 4096u}) {
                og_tuning c = cur;
                c.tile = tile;
                trial(c);
            }
            if (has_dyn)
                for (uint32_t dg : {512u, 1024u, 2048u, 4096u, 8192u}) {
                    og_tuning c = cur;
                    c.dyn_grid = dg;
                    trial(c);
                }
        }
        apply_tuning(g, &cur);
        if (best) *best = cur;
        if (best_ms) *best_ms = cur_ms;
        return OG_OK;
    });
    g->busy.store(0);
    return st;
}

// ==================================================== allocation-free BFS
//
// og_bfs_device_noalloc: og_bfs over a CSR that already sits in device
// memory, for callers that cannot allocate (hipMalloc / hipFree / an
// allocating library call) at run time.  Every buffer beyond the caller's
// inputs and output -- next-frontier flags, bin lists, hub tiles, per-block
// counts, the round's summary and control block, counters -- lives in the
// file-scope device array `scratch` below, sized for V <= kNoAllocMaxV and
// E <= kNoAllocMaxE.  Rounds are enqueued directly (no captured hipGraph, no
// events) and cleared with fill kernels, so the call allocates nothing.
// Each GPU holds its own copy of `scratch` until the process ends; calls
// that overlap on one GPU share it and race.

namespace og {

constexpr uint64_t kNoAllocMaxV = 1ull << 21;  // vertices
constexpr uint64_t kNoAllocMaxE = 1ull << 26;  // edges
constexpr uint64_t kNoAlign = 256;
constexpr uint64_t na_up(uint64_t x) { return (x + kNoAlign - 1) / kNoAlign * kNoAlign; }
// Hub tiles of any graph within the limits under the default tuning: a vertex
// of degree d > t2 takes ceil(d / tile) <= d / tile + 1 tiles, and at most
// E / (t2 + 1) vertices exceed t2.
constexpr uint64_t kNoAllocHubCap = kNoAllocMaxE / 2048 + kNoAllocMaxE / 4097 + 1;
static_assert(frontier_fv(kNoAllocMaxV) == 1, "the allocation-free BFS runs the one-vector compaction");
constexpr uint64_t kNoAllocMaxBlk = (kNoAllocMaxV + verts_per_block(1) - 1) / verts_per_block(1);
constexpr uint64_t kNaFlag = 0;
constexpr uint64_t kNaList = kNaFlag + na_up(kNoAllocMaxV);
constexpr uint64_t kNaHub = kNaList + na_up(4 * kNoAllocMaxV);
constexpr uint64_t kNaDoff = kNaHub + na_up(8 * kNoAllocHubCap);
constexpr uint64_t kNoAllocMaxSb = (kNoAllocMaxBlk + kSbBlocks - 1) / kSbBlocks;
constexpr uint64_t kNaBcnt = kNaDoff + na_up(8 * 16);
constexpr uint64_t kNaBdeg = kNaBcnt + na_up(4 * kBcntRows * kNoAllocMaxBlk);
constexpr uint64_t kNaSbt = kNaBdeg + na_up(8 * kNoAllocMaxBlk);
constexpr uint64_t kNaSbp = kNaSbt + na_up(8 * kSbVals * kNoAllocMaxSb);
constexpr uint64_t kNaInfo = kNaSbp + na_up(8 * 3 * kNoAllocMaxSb);
constexpr uint64_t kNaCtl = kNaInfo + na_up(sizeof(FrontierInfo));
constexpr uint64_t kNaCounters = kNaCtl + na_up(sizeof(RoundCtl));
constexpr uint64_t kNoAllocScratchBytes = kNaCounters + na_up(16 * 8);

// The scratch memory of og_bfs_device_noalloc.  In host code `scratch` is a
// stand-in for the symbol, not a device address: kernels get the address
// that hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

constexpr int kNoAllocMaxDevices = 64;
static std::mutex g_na_mu;
static void *g_na_ptr[kNoAllocMaxDevices];  // device address of `scratch`, per device

// Device address of `scratch` on the current device, looked up once per
// device with that device current.
static uint8_t *noalloc_scratch(int dev) {
    if (dev < 0 || dev >= kNoAllocMaxDevices) raise(OG_EUNSUPPORTED, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            raise(OG_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        }
        g_na_ptr[dev] = p;
    }
    return (uint8_t *)g_na_ptr[dev];
}

static void noalloc_bfs(const uint64_t *d_ptr, const uint32_t *d_cols, uint64_t V, uint64_t E, uint32_t src,
                        uint32_t *d_levels, hipStream_t s, og_run_stats *stats) {
    int dev = 0;
    OG_HIP(hipGetDevice(&dev));
    uint8_t *base = noalloc_scratch(dev);
    uint64_t ends[2] = {1, 1};
    OG_HIP(hipMemcpyAsync(&ends[0], d_ptr, 8, hipMemcpyDeviceToHost, s));
    OG_HIP(hipMemcpyAsync(&ends[1], d_ptr + V, 8, hipMemcpyDeviceToHost, s));
    OG_HIP(hipStreamSynchronize(s));
    if (ends[0] != 0 || ends[1] != E)
        raise(OG_EINVAL, "row offsets must run from 0 to E=%llu over V+1 entries (got %llu .. %llu)",
              (unsigned long long)E, (unsigned long long)ends[0], (unsigned long long)ends[1]);
    const uint64_t Vpad = (V + 15) / 16 * 16;
    const uint32_t nblk = (uint32_t)((Vpad + verts_per_block(1) - 1) / verts_per_block(1));
    uint8_t *flag = base + kNaFlag;
    uint32_t *list = (uint32_t *)(base + kNaList);
    uint64_t *hub = (uint64_t *)(base + kNaHub);
    uint64_t *doff = (uint64_t *)(base + kNaDoff);
    const uint32_t nsb = (nblk + kSbBlocks - 1) / kSbBlocks;
    const BlockCounts bc{(uint32_t *)(base + kNaBcnt), (uint64_t *)(base + kNaBdeg), (uint64_t *)(base + kNaSbt),
                         (uint64_t *)(base + kNaSbp), nblk, nsb};
    FrontierInfo *info = (FrontierInfo *)(base + kNaInfo);
    RoundCtl *ctl = (RoundCtl *)(base + kNaCtl);
    unsigned long long *ctr = (unsigned long long *)(base + kNaCounters);
    const Tuning tn;  // default bins: the hub bound above assumes its t2 and tile
    static_assert(Tuning{}.t2 == 4096 && Tuning{}.tile == 2048, "hub capacity bound follows the default tuning");
    Stopwatch total;
    // clear what must start at zero: fill kernels (no memset nodes)
    hipLaunchKernelGGL(k_fill_u8, dim3(fill_grid(Vpad)), dim3(256), 0, s, flag, Vpad, (uint8_t)0);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, (uint32_t *)ctr, (uint64_t)32, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(fill_grid(V)), dim3(256), 0, s, d_levels, V, OG_UNREACHED);
    hipLaunchKernelGGL(k_set_source, dim3(1), dim3(1), 0, s, d_levels, flag, src);
    hipLaunchKernelGGL(k_ctl_reset, dim3(1), dim3(1), 0, s, ctl, 0u, 0u, (uint64_t)0, V, 0u);
    hipLaunchKernelGGL(k_sb_init, dim3(1), dim3(64), 0, s, bc.sbt, nsb);
    auto compact = [&] {
        hipLaunchKernelGGL(k_frontier_count<1>, dim3(nblk), dim3(kBlock), 0, s, flag, (uint8_t *)nullptr, Vpad, V,
                           d_ptr, (const uint64_t *)nullptr, (const uint32_t *)nullptr, (const uint32_t *)nullptr, tn,
                           (const uint32_t *)nullptr, ctl, bc);
        hipLaunchKernelGGL(k_frontier_scan, dim3(1), dim3(kWave), 0, s, bc, info, doff, ctl, (uint64_t *)nullptr);
        hipLaunchKernelGGL(k_frontier_write<1>, dim3(nblk), dim3(kBlock), 0, s, flag, (uint8_t *)nullptr, Vpad, d_ptr,
                           (const uint64_t *)nullptr, (const uint32_t *)nullptr, (const uint32_t *)nullptr, tn, bc,
                           info, list, hub,
                           kNoAllocHubCap, doff, (uint32_t *)nullptr, (uint64_t *)nullptr, (uint32_t *)nullptr,
                           (uint64_t)0);
    };
    RoundRelax r{};
    r.list = list;
    r.info = info;
    r.ptr = d_ptr;
    r.sedges = d_cols;
    r.hub = hub;
    r.hub_cap = kNoAllocHubCap;
    r.tile = tn.tile;
    r.cd = 0;  // the whole graph is static: three bins in one launch
    r.counters = ctr;
    const RelaxArgs a{d_levels, flag, nullptr, nullptr, nullptr, 0, nullptr, nullptr};
    static_assert(Tuning{}.lpv0 == 8, "the relax below is instantiated for 8 lanes per bin-0 vertex");
    compact();
    OG_HIP(hipGetLastError());
    RoundCtl h{};
    for (;;) {
        for (int k = 0; k < kRoundBatch; ++k) {
            hipLaunchKernelGGL((k_relax_round<ALGO_BFS, false, false, 8>), dim3(kFuseStatic * kHubGridCap),
                               dim3(kBlock), (size_t)tn.tile * 4, s, r, a);
            compact();
        }
        OG_HIP(hipGetLastError());
        OG_HIP(hipMemcpyAsync(&h, ctl, sizeof h, hipMemcpyDeviceToHost, s));
        OG_HIP(hipStreamSynchronize(s));
        if (!h.live) break;
    }
    const double ms = total.ms();
    hipLaunchKernelGGL(k_reach_stats, dim3((unsigned)std::min<uint64_t>(2048, (V + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, s, d_levels, d_ptr, (const uint32_t *)nullptr, (const uint32_t *)nullptr, V,
                       ctr + 2);
    OG_HIP(hipGetLastError());
    unsigned long long hc[16];
    OG_HIP(hipMemcpyAsync(hc, ctr, sizeof hc, hipMemcpyDeviceToHost, s));
    OG_HIP(hipStreamSynchronize(s));
    if (stats) {
        og_run_stats st{};
        st.approach = OG_ASCETIC;
        st.iterations = h.rounds;
        st.node_sum = h.node_sum;
        st.static_node_sum = h.static_node_sum;
        st.work_edges = h.work_edges;
        st.static_work_edges = h.static_work_edges;
        st.updates = hc[0];
        st.reached = hc[2];
        st.traversed_edges = hc[3];
        st.checksum = hc[4];
        st.total_ms = ms;
        *stats = st;
    }
}

}  // namespace og

extern "C" og_status og_bfs_device_noalloc(const uint64_t *row_ptr, const uint32_t *cols, uint64_t num_vertices,
                                           uint64_t num_edges, uint32_t source, uint32_t *levels, void *stream,
                                           og_run_stats *stats) {
    clear_error();
    if (!row_ptr || !levels || (num_edges && !cols)) return fail(OG_EINVAL, "null device pointer");
    if (num_vertices == 0 || source >= num_vertices)
        return fail(OG_EINVAL, "source %u out of range for V=%llu", source, (unsigned long long)num_vertices);
    if (num_vertices > kNoAllocMaxV || num_edges > kNoAllocMaxE)
        return fail(OG_EINVAL,
                    "og_bfs_device_noalloc handles V <= %llu and E <= %llu (its scratch is fixed at %llu bytes); "
                    "got V=%llu E=%llu",
                    (unsigned long long)kNoAllocMaxV, (unsigned long long)kNoAllocMaxE,
                    (unsigned long long)kNoAllocScratchBytes, (unsigned long long)num_vertices,
                    (unsigned long long)num_edges);
    return guard([&]() -> og_status {
        if (!hip_available()) raise(OG_EHIP, "no HIP device is visible");
        noalloc_bfs(row_ptr, cols, num_vertices, num_edges, source, levels, (hipStream_t)stream, stats);
        return OG_OK;
    });
}
