/*
 * oracle.c -- CPU restatements of the reference algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (libonegraph.so) never links or calls it.
 *
 * Arrays follow the reference CSR layout: ptr has V entries (no sentinel;
 * row V-1 ends at E, BFS/cpu_BFS/main.cpp:127-132).
 *
 * Pinning (see tests/test_oracle.py):
 *   og_oracle_bfs      -- closed forms (path, star, complete, binary tree,
 *                         grid), edge invariants, scipy breadth_first_order
 *                         fixtures in tests/golden.
 *   og_oracle_sssp     -- closed forms, edge/tightness invariants, scipy
 *                         dijkstra fixtures; unit weights must equal BFS.
 *   og_oracle_cc       -- union-find min-id on undirected inputs, scipy
 *                         connected_components fixtures, and the update count
 *                         against the reference serial CC program compiled
 *                         from its own source (oracle/_ref/cc_serial).
 *   og_oracle_sssp     -- also: distance + 1 saturates at 0xFFFFFFFE (a
 *                         fixture with weights near 2^31 checks the cap).
 *   og_oracle_pagerank -- closed forms (cycle, complete graph, star), networkx
 *                         pagerank fixtures, mass conservation.
 *   og_oracle_pagerank_par -- equal to og_oracle_pagerank bit for bit (ranks,
 *                         iterations, residual) at 1, 2 and 16 threads.
 *
 * og_oracle_loop_seconds() returns the wall time of the last call's compute
 * loop alone (allocation and initialisation excluded), the boundary the
 * reference CPU programs time (CPUCalTime, BFS/cpu_BFS/main.cpp:156-208).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double g_loop_s;

static double now_s(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

double og_oracle_loop_seconds(void) { return g_loop_s; }

static uint64_t row_end(uint64_t V, uint64_t E, const uint64_t *ptr, uint64_t v) {
    return v + 1 < V ? ptr[v + 1] : E;
}

/* runCPUBFS (BFS/cpu_BFS/main.cpp:114-209): value[] = UINT_MAX, value[src] =
 * 1; every sweep visits all vertices in index order and relaxes the out-edges
 * of those active in the previous sweep (finalValue = value + 1, written when
 * smaller, marking next[]); the sweep ends by copying next -> active and
 * counting.  *node_sum accumulates the active counts (activeSum, :162). */
void og_oracle_bfs(uint64_t V, uint64_t E, const uint64_t *ptr, const uint32_t *cols, uint32_t src, uint32_t *value,
                   uint64_t *node_sum, uint32_t *iterations) {
    unsigned char *act = (unsigned char *)calloc(V, 1);
    unsigned char *next = (unsigned char *)calloc(V, 1);
    for (uint64_t i = 0; i < V; ++i) value[i] = 0xFFFFFFFFu;
    act[src] = 1;
    value[src] = 1;
    uint64_t active = 1, sum = 0;
    uint32_t iter = 0;
    const double t0 = now_s();
    while (active > 0) {
        sum += active;
        for (uint64_t u = 0; u < V; ++u) {
            if (!act[u]) continue;
            const uint32_t fv = value[u] + 1;
            for (uint64_t i = ptr[u]; i < row_end(V, E, ptr, u); ++i) {
                const uint32_t v = cols[i];
                if (fv < value[v]) {
                    value[v] = fv;
                    next[v] = 1;
                }
            }
        }
        active = 0;
        for (uint64_t i = 0; i < V; ++i) {
            act[i] = next[i];
            active += act[i];
            next[i] = 0;
        }
        ++iter;
    }
    g_loop_s = now_s() - t0;
    if (node_sum) *node_sum = sum;
    if (iterations) *iterations = iter;
    free(act);
    free(next);
}

/* ---- binary min-heap of (dist, vertex) for Dijkstra ---- */
typedef struct {
    uint64_t d;
    uint32_t v;
} HeapItem;

static void heap_push(HeapItem *h, uint64_t *n, HeapItem x) {
    uint64_t i = (*n)++;
    while (i > 0) {
        uint64_t p = (i - 1) / 2;
        if (h[p].d <= x.d) break;
        h[i] = h[p];
        i = p;
    }
    h[i] = x;
}

static HeapItem heap_pop(HeapItem *h, uint64_t *n) {
    HeapItem top = h[0], last = h[--(*n)];
    uint64_t i = 0;
    for (;;) {
        uint64_t c = 2 * i + 1;
        if (c >= *n) break;
        if (c + 1 < *n && h[c + 1].d < h[c].d) ++c;
        if (h[c].d >= last.d) break;
        h[i] = h[c];
        i = c;
    }
    if (*n) h[i] = last;
    return top;
}

/* Textbook Dijkstra (the reference has no CPU SSSP).  Output follows the
 * reference SSSP convention (SSSP/oneGraph/GraphMeta.h:383-389): source = 1,
 * a vertex at distance d holds d + 1, unreached 0xFFFFFFFF. */
void og_oracle_sssp(uint64_t V, uint64_t E, const uint64_t *ptr, const uint32_t *to, const uint32_t *w, uint32_t src,
                    uint32_t *value) {
    uint64_t *dist = (uint64_t *)malloc(V * sizeof(uint64_t));
    unsigned char *done = (unsigned char *)calloc(V, 1);
    HeapItem *heap = (HeapItem *)malloc((E + 1) * sizeof(HeapItem));
    uint64_t n = 0;
    for (uint64_t i = 0; i < V; ++i) dist[i] = UINT64_MAX;
    dist[src] = 0;
    HeapItem s0 = {0, src};
    const double t0 = now_s();
    heap_push(heap, &n, s0);
    while (n) {
        HeapItem it = heap_pop(heap, &n);
        if (done[it.v]) continue;
        done[it.v] = 1;
        for (uint64_t i = ptr[it.v]; i < row_end(V, E, ptr, it.v); ++i) {
            const uint32_t v = to[i];
            const uint64_t nd = it.d + w[i];
            if (nd < dist[v]) {
                dist[v] = nd;
                HeapItem x = {nd, v};
                heap_push(heap, &n, x);
            }
        }
    }
    g_loop_s = now_s() - t0;
    /* distance + 1 saturates at 0xFFFFFFFE, the largest value that is not
     * "unreached" (the engine's convention for paths longer than 2^32 - 2) */
    for (uint64_t i = 0; i < V; ++i)
        value[i] = dist[i] == UINT64_MAX ? 0xFFFFFFFFu
                                         : (dist[i] >= 0xFFFFFFFDull ? 0xFFFFFFFEu : (uint32_t)(dist[i] + 1));
    free(dist);
    free(done);
    free(heap);
}

/* Serial CC (CC/serial/serial.cpp:164-213): value[i] = i, all vertices
 * active; each sweep visits vertices in index order and, for active ones,
 * reads value[u] once and lowers every out-neighbour whose value is larger
 * (updates are seen by later vertices of the same sweep); next[] becomes the
 * active set.  Returns the number of successful lowerings (the program's
 * printed "nodeSum", serial.cpp:195,225). */
uint64_t og_oracle_cc(uint64_t V, uint64_t E, const uint64_t *ptr, const uint32_t *cols, uint32_t *value) {
    unsigned char *act = (unsigned char *)malloc(V);
    unsigned char *next = (unsigned char *)calloc(V, 1);
    for (uint64_t i = 0; i < V; ++i) {
        value[i] = (uint32_t)i;
        act[i] = 1;
    }
    uint64_t updates = 0, active = 1;
    const double t0 = now_s();
    for (;;) {
        if (active == 0) break;
        for (uint64_t u = 0; u < V; ++u) {
            if (!act[u]) continue;
            const uint32_t sv = value[u];
            for (uint64_t i = ptr[u]; i < row_end(V, E, ptr, u); ++i) {
                const uint32_t v = cols[i];
                if (sv < value[v]) {
                    ++updates;
                    value[v] = sv;
                    next[v] = 1;
                }
            }
        }
        active = 0;
        for (uint64_t i = 0; i < V; ++i) {
            act[i] = next[i];
            active += act[i];
            next[i] = 0;
        }
    }
    g_loop_s = now_s() - t0;
    free(act);
    free(next);
    return updates;
}

/* Dangling mass and L1 residual are reduced over fixed blocks of
 * PR_BLOCK vertices: each block sums in vertex order and the block partials
 * are added in block order.  Both PageRank functions below use exactly this
 * order, so their ranks, iteration counts and residuals are identical for
 * every thread count. */
#define PR_BLOCK 4096u

/* Normalised PageRank by FP64 power iteration over the pull layout (the loop
 * shape of pr_cal, PR/cpu/pr.cpp:136-149; the numerics are the textbook
 * definition, not the reference's 0.15 + 0.85*sum):
 *   r0 = 1/n;  new[v] = (1-d)/n + d*(sum_{u in in(v)} r[u]/outdeg[u] + dangling/n)
 *   stop after the first iteration whose L1 change is < tol, or max_it. */
void og_oracle_pagerank(uint64_t V, uint64_t E, const uint32_t *outdeg, const uint64_t *ptr, const uint32_t *in_cols,
                        double damping, double tol, uint32_t max_it, double *rank, uint32_t *iterations,
                        double *final_l1) {
    double *next = (double *)malloc(V * sizeof(double));
    double *contrib = (double *)malloc(V * sizeof(double));
    const double inv_n = 1.0 / (double)V;
    for (uint64_t i = 0; i < V; ++i) rank[i] = inv_n;
    uint32_t it = 0;
    double l1 = 0;
    const double t0 = now_s();
    while (it < max_it) {
        double dang = 0;
        for (uint64_t b0 = 0; b0 < V; b0 += PR_BLOCK) {
            const uint64_t b1 = b0 + PR_BLOCK < V ? b0 + PR_BLOCK : V;
            double d = 0;
            for (uint64_t u = b0; u < b1; ++u) {
                contrib[u] = outdeg[u] ? rank[u] / (double)outdeg[u] : 0.0;
                if (!outdeg[u]) d += rank[u];
            }
            dang += d;
        }
        const double base = (1.0 - damping) * inv_n + damping * dang * inv_n;
        l1 = 0;
        for (uint64_t b0 = 0; b0 < V; b0 += PR_BLOCK) {
            const uint64_t b1 = b0 + PR_BLOCK < V ? b0 + PR_BLOCK : V;
            double l = 0;
            for (uint64_t v = b0; v < b1; ++v) {
                double s = 0;
                for (uint64_t i = ptr[v]; i < row_end(V, E, ptr, v); ++i) s += contrib[in_cols[i]];
                next[v] = base + damping * s;
                l += fabs(next[v] - rank[v]);
            }
            l1 += l;
        }
        memcpy(rank, next, V * sizeof(double));
        ++it;
        if (l1 < tol) break;
    }
    g_loop_s = now_s() - t0;
    if (iterations) *iterations = it;
    if (final_l1) *final_l1 = l1;
    free(next);
    free(contrib);
}

/* og_oracle_pagerank with the blocks of each vertex loop spread over
 * `threads` OpenMP threads (the CPU baseline's all-core variant).  Every
 * rank is summed in the same in-edge order and the block partials of the
 * dangling mass and the residual are stored per block and added in block
 * order afterwards, so the results equal og_oracle_pagerank's bit for bit. */
void og_oracle_pagerank_par(uint64_t V, uint64_t E, const uint32_t *outdeg, const uint64_t *ptr,
                            const uint32_t *in_cols, double damping, double tol, uint32_t max_it, double *rank,
                            uint32_t *iterations, double *final_l1, int threads) {
    const uint64_t nblk = (V + PR_BLOCK - 1) / PR_BLOCK;
    double *next = (double *)malloc(V * sizeof(double));
    double *contrib = (double *)malloc(V * sizeof(double));
    double *part = (double *)malloc((nblk ? nblk : 1) * sizeof(double));
    const double inv_n = 1.0 / (double)V;
    for (uint64_t i = 0; i < V; ++i) rank[i] = inv_n;
    uint32_t it = 0;
    double l1 = 0;
    if (threads < 1) threads = 1;
    const double t0 = now_s();
    while (it < max_it) {
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int64_t k = 0; k < (int64_t)nblk; ++k) {
            const uint64_t b0 = (uint64_t)k * PR_BLOCK, b1 = b0 + PR_BLOCK < V ? b0 + PR_BLOCK : V;
            double d = 0;
            for (uint64_t u = b0; u < b1; ++u) {
                contrib[u] = outdeg[u] ? rank[u] / (double)outdeg[u] : 0.0;
                if (!outdeg[u]) d += rank[u];
            }
            part[k] = d;
        }
        double dang = 0;
        for (uint64_t k = 0; k < nblk; ++k) dang += part[k];
        const double base = (1.0 - damping) * inv_n + damping * dang * inv_n;
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int64_t k = 0; k < (int64_t)nblk; ++k) {
            const uint64_t b0 = (uint64_t)k * PR_BLOCK, b1 = b0 + PR_BLOCK < V ? b0 + PR_BLOCK : V;
            double l = 0;
            for (uint64_t v = b0; v < b1; ++v) {
                double s = 0;
                for (uint64_t i = ptr[v]; i < row_end(V, E, ptr, v); ++i) s += contrib[in_cols[i]];
                next[v] = base + damping * s;
                l += fabs(next[v] - rank[v]);
            }
            part[k] = l;
        }
        l1 = 0;
        for (uint64_t k = 0; k < nblk; ++k) l1 += part[k];
        memcpy(rank, next, V * sizeof(double));
        ++it;
        if (l1 < tol) break;
    }
    g_loop_s = now_s() - t0;
    if (iterations) *iterations = it;
    if (final_l1) *final_l1 = l1;
    free(next);
    free(contrib);
    free(part);
}
