/*
 * onegraph.h -- C ABI of the MI355X-native out-of-GPU-memory graph engine.
 *
 * The reference (NKU-EmbeddedSystem/OneGraph) has no library or FFI: each of
 * its 15 programs reads a binary CSR file and calls one free function from
 * main().  Every entry point below names the reference function or code block
 * it replaces (paths relative to the reference repository root).
 *
 * Conventions shared by every entry point:
 *   - plain pointers and sizes only; the caller owns every output array;
 *   - every call returns an og_status; nothing throws or exits across the ABI;
 *   - og_last_error() returns a thread-local message for the last failure;
 *   - vertex ids are uint32, edge offsets uint64 (the reference's
 *     EDGE_POINTER_TYPE, BFS/oneGraph_BFS/globals.h:14);
 *   - BFS/SSSP results use the reference CPU convention: the source holds 1,
 *     a vertex at hop h (distance d) holds h+1 (d+1), unreached holds
 *     0xFFFFFFFF (BFS/cpu_BFS/main.cpp:141-145, SSSP/oneGraph/GraphMeta.h:374-381);
 *     an SSSP distance d with d + 1 >= 0xFFFFFFFE holds 0xFFFFFFFE (the
 *     reference's u32 sum wraps there instead, SSSP/oneGraph/gpu_kernels.h:
 *     186-198; weights in [1, 63] never reach it below 2^26 hops);
 *   - CC labels: min-label propagation from value[i]=i
 *     (CC/serial/serial.cpp:164-213).
 */
#ifndef ONEGRAPH_H
#define ONEGRAPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OG_UNREACHED 0xFFFFFFFFu

typedef enum og_status {
    OG_OK = 0,
    OG_EINVAL = 1,          /* bad argument: null pointer, source >= V, unknown enum */
    OG_EIO = 2,             /* open / read / write failed or file truncated        */
    OG_EFORMAT = 3,         /* CSR header or offsets inconsistent                  */
    OG_EBUDGET = 4,         /* edge budget cannot hold what the request needs      */
    OG_ENOMEM_HOST = 5,     /* host or pinned allocation failed                    */
    OG_ENOMEM_DEVICE = 6,   /* device allocation failed                            */
    OG_EHIP = 7,            /* HIP runtime error, including "no GPU present"       */
    OG_EBUSY = 8,           /* object already in use / server queue full           */
    OG_EUNSUPPORTED = 9,    /* algorithm and graph kind do not fit together        */
    OG_ESTATE = 10,         /* call order violated (e.g. run before og_prepare)    */
    OG_ETIMEDOUT = 11       /* served query not started within its deadline         */
} og_status;

/* Graph kinds = the three on-disk CSR layouts of the reference. */
typedef enum og_graph_kind {
    OG_KIND_PLAIN = 0,      /* u64 V, u64 E, u64 ptr[V], u32 col[E]
                               (BFS/oneGraph_BFS/GraphMeta.h:125-145)              */
    OG_KIND_WEIGHTED = 1,   /* u64 V, u64 E, u64 ptr[V], {u32 to,u32 w}[E]
                               (SSSP/subway/subway.cpp:144-152, globals.h:18-21)   */
    OG_KIND_PULL = 2        /* u64 V, u64 E, u32 outDegree[V], u64 ptr[V],
                               u32 inNeighbour[E]  (PR/cpu/Graph.h readDataFromFile) */
} og_graph_kind;

/* The reference's three out-of-memory approaches. */
typedef enum og_approach {
    OG_ASCETIC = 0,         /* static prefix + dynamic region (BFS/oneGraph_BFS)   */
    OG_SUBWAY = 1,          /* active-subgraph generation (BFS/sbw_BFS)             */
    OG_USM = 2              /* graph in managed memory (BFS/usm_BFS)                */
} og_approach;

typedef struct og_graph og_graph;

typedef struct og_graph_info {
    uint64_t num_vertices;
    uint64_t num_edges;
    uint32_t kind;          /* og_graph_kind */
    uint32_t max_degree;
} og_graph_info;

/* ---------------------------------------------------------------- errors */

/* Thread-local text of the last failing call on this thread ("" if none). */
const char *og_last_error(void);
/* Library version string. */
const char *og_version(void);
/* Number of visible HIP devices (0 on a host without a GPU). */
og_status og_device_count(int *count);

/* ---------------------------------------------------------------- graphs */

/* Replaces GraphMeta::readDataFromFile (BFS/oneGraph_BFS/GraphMeta.h:125-145),
 * its weighted twin (SSSP/oneGraph/GraphMeta.h) and the PR reader
 * (PR/cpu/Graph.h).  Unlike the reference it checks sizes, offsets and
 * column ids (OG_EIO / OG_EFORMAT) and loads into pinned host memory. */
og_status og_graph_load(const char *path, og_graph_kind kind, og_graph **out);

/* Writes the graph in the reference byte format of its kind. */
og_status og_graph_save(const og_graph *g, const char *path);

/* Builds a graph from caller arrays in the reference layout: ptr has V
 * entries (no sentinel; the last vertex ends at E, GraphMeta.h:222).
 * weights may be NULL unless kind == OG_KIND_WEIGHTED; out_degree is used
 * only for OG_KIND_PULL. The arrays are copied. */
og_status og_graph_from_csr(uint64_t num_vertices, uint64_t num_edges,
                            const uint64_t *ptr, const uint32_t *cols,
                            const uint32_t *weights, const uint32_t *out_degree,
                            og_graph_kind kind, og_graph **out);

/* Edge-list conversion (no reference counterpart: the reference relies on
 * the external Ascetic converter, README.md:7, to turn a text graph into its
 * CSR).  From arcs src[i] -> dst[i] (weights[i] when weights != NULL) the
 * graph gets deduplicated, row-sorted adjacency: a repeated arc keeps its
 * smallest weight, self-loops are dropped unless OG_EDGES_KEEP_SELF_LOOPS,
 * OG_EDGES_SYMMETRIZE adds v -> u for every u -> v (same weight).
 * num_vertices = 0 takes max id + 1; ids >= num_vertices are OG_EINVAL. */
#define OG_EDGES_SYMMETRIZE 1u
#define OG_EDGES_KEEP_SELF_LOOPS 2u
#define OG_EDGES_ONE_BASED 4u   /* text ids start at 1 (MatrixMarket sets it) */
#define OG_EDGES_WEIGHTED 8u    /* text: the third column is a u32 weight     */
og_status og_graph_from_edges(uint64_t num_vertices, uint64_t num_edges,
                              const uint32_t *src, const uint32_t *dst,
                              const uint32_t *weights, uint32_t flags, og_graph **out);

/* Reads a text edge list: one arc "u v [w]" per line, separated by spaces,
 * tabs or commas; lines starting with '#' or '%' are comments (SNAP style).
 * A MatrixMarket coordinate file ("%%MatrixMarket" first line) is read with
 * its size line, 1-based ids and, when "symmetric", both directions.  Builds
 * the graph as og_graph_from_edges does (plain, or weighted with
 * OG_EDGES_WEIGHTED).  OG_EFORMAT names the first bad line. */
og_status og_graph_load_edgelist(const char *path, uint32_t flags, og_graph **out);

/* Synthetic inputs (no reference counterpart: the reference relies on the
 * external Ascetic converter, README.md:7).  Directed R-MAT with quadrant
 * probabilities a,b,c (d = 1-a-b-c), self-loops and duplicates removed,
 * rows sorted.  weight_seed != 0 attaches u32 weights uniform in [1, 63]. */
og_status og_graph_generate_rmat(uint32_t scale, uint32_t edge_factor,
                                 double a, double b, double c, uint64_t seed,
                                 uint64_t weight_seed, og_graph **out);

/* Undirected Erdos-Renyi G(n, m = n*avg_degree/2), symmetrised, no
 * self-loops or duplicates; returned as OG_KIND_PLAIN out-edge CSR. */
og_status og_graph_generate_er(uint64_t num_vertices, uint32_t avg_degree,
                               uint64_t seed, og_graph **out);

/* Builds the pull layout used by PageRank (outDegree + in-edge CSR) from a
 * plain or weighted out-edge graph (weights dropped). */
og_status og_graph_to_pull(const og_graph *g, og_graph **out);

/* Adds u32 weights uniform in [1, 63] (seeded) to a plain graph. */
og_status og_graph_add_weights(const og_graph *g, uint64_t weight_seed, og_graph **out);

og_status og_graph_get_info(const og_graph *g, og_graph_info *info);

/* Copies the CSR out in the reference layout (any pointer may be NULL). */
og_status og_graph_export(const og_graph *g, uint64_t *ptr, uint32_t *cols,
                          uint32_t *weights, uint32_t *out_degree);

/* Reference findbignode (BFS/oneGraph_BFS/GraphMeta.h:149-165): first vertex
 * after `last` whose degree exceeds 3000, else `last`. */
og_status og_graph_find_big_node(const og_graph *g, uint32_t last, uint32_t *out);

/* Frees the graph and its device state.  OG_EBUSY (and the graph left
 * intact) while another call runs on it or an og_server holds it; NULL is
 * OG_OK. */
og_status og_graph_free(og_graph *g);

/* ------------------------------------------------------------- planning */

typedef struct og_plan_params {
    uint64_t edge_budget_bytes; /* device bytes for static region + staging;
                                   0 = reference default: 12 GiB minus
                                   paramSize*4*V (GraphMeta.h:304-309)        */
    float advisek;              /* static share of the budget; <= 0 selects
                                   the reference auto rule (GraphMeta.h:311-324) */
    uint32_t param_size;        /* reference paramSize (17 BFS/SSSP/CC, 19 PR);
                                   only used when edge_budget_bytes == 0      */
    uint32_t host_threads;      /* gather threads; 0 = hardware concurrency   */
    uint32_t flags;             /* OG_PLAN_* bits                             */
    uint32_t split_weight;      /* SSSP light/heavy split (delta-stepping,
                                   Meyer & Sanders): > 0 stores each row of
                                   the device copy light edges (weight <
                                   split_weight) first, and the zero-copy
                                   host copy all light parts before all heavy
                                   parts; near-far SSSP rounds then relax
                                   light edges only and a vertex's heavy
                                   edges once, after its bucket has settled.
                                   With weights in [1, 63], 8 with a bucket
                                   width (tuning delta) of 256 does the least
                                   work per round.  Packed weighted graphs,
                                   Ascetic and Subway; 0 = no split          */
} og_plan_params;

#define OG_PLAN_NO_OVERLAP 1u   /* serialise copies, static and dynamic relax
                                   (for A/B timing)                              */
#define OG_PLAN_ZERO_COPY 2u    /* dynamic region: the relax reads active edges
                                   straight from pinned host memory over the link
                                   (no staging copies, no host gather); the
                                   staging share of the budget keeps the chunks
                                   of 32 vertices whose rows cost the most link
                                   lines per edge resident in HBM instead      */
#define OG_PLAN_NO_PACK 4u      /* keep 8-byte weighted edges in HBM and on the
                                   link (no 4-byte (to << wbits | w) format)     */
#define OG_PLAN_NO_GRAPH 8u     /* enqueue device-driven rounds one API call at
                                   a time instead of replaying a captured
                                   hipGraph of them (for A/B timing)             */
#define OG_PLAN_NO_FUSE 32u     /* zero-copy rounds: relax the static and the
                                   dynamic region in two launches on two
                                   streams instead of one (for A/B timing)     */
#define OG_PLAN_DIRECTION_OPT 16u /* BFS on a graph wholly in HBM: rounds with a
                                   large frontier run bottom-up (every
                                   unvisited vertex scans its in-edges for a
                                   parent, stopping at the first); needs the
                                   in-edge CSR on the device (E * 4 + V * 8
                                   bytes beyond the edge budget)               */

typedef struct og_plan_info {
    uint64_t total_budget_edges;  /* reference total_gpu_size               */
    uint64_t static_edges;        /* reference max_partition_size           */
    uint64_t staging_edges;       /* reference partOverloadSize             */
    uint64_t overload_edges;      /* reference overloadSize                 */
    uint64_t static_vertices;     /* number of vertices with isInStatic     */
    uint64_t vsplit;              /* vertices [0, vsplit) are static        */
    uint64_t device_bytes;        /* device bytes allocated by the plan     */
    float advisek_used;
    uint32_t approach;
} og_plan_info;

/* Restatement of GraphMeta::getMaxPartitionSize (GraphMeta.h:299-360), pure
 * host computation, no device needed.  ptr has V entries (reference layout,
 * no sentinel).  edge_bytes = sizeof(Edge) (4 or 8). */
og_status og_plan_compute(uint64_t num_vertices, uint64_t num_edges,
                          const uint64_t *ptr, uint32_t edge_bytes,
                          const og_plan_params *params, og_plan_info *info);

/* Places the graph on the device for one approach: replaces
 * GraphMeta::initGraphHost/initGraphDevice (GraphMeta.h:203-291) for
 * Ascetic, the set-up of runGPUbfs_subway (BFS/sbw_BFS/main.cpp:323-392) for
 * Subway and the malloc_shared graph of BFS/usm_BFS/Graph.h:58-62 for USM.
 * Frees any previous preparation of this graph. */
og_status og_prepare(og_graph *g, og_approach approach,
                     const og_plan_params *params, og_plan_info *info);

/* Releases the device state of og_prepare. */
og_status og_release(og_graph *g);

/* --------------------------------------------------------------- runs */

typedef struct og_run_stats {
    uint32_t iterations;          /* frontier rounds                             */
    uint32_t approach;
    uint64_t reached;             /* vertices with a value != OG_UNREACHED        */
    uint64_t traversed_edges;     /* sum of out-degree over reached vertices      */
    uint64_t work_edges;          /* edges relaxed, summed over iterations (D)    */
    uint64_t updates;             /* successful improvements (U)                  */
    uint64_t node_sum;            /* reference nodeSum: sum of frontier sizes     */
    uint64_t static_node_sum;     /* reference staticnodeSum                      */
    uint64_t static_work_edges;   /* edges relaxed out of the static region       */
    uint64_t h2d_bytes;           /* dynamic-region payload copied host->device   */
    uint64_t chunks;              /* dynamic-region copies issued                 */
    double total_ms;              /* reference totalProcess                       */
    double pre_ms;                /* reference preProcess (classify/compact/scan) */
    double static_ms;             /* reference staticProcess                      */
    double overload_ms;           /* reference overloadProcess                    */
    double overload_move_ms;      /* reference overloadMoveProcess (copy stream)  */
    double gather_ms;             /* host gather of active dynamic edges          */
    /* Kernel timing from HIP events on the launching stream. */
    double relax_static_kernel_ms;  uint64_t relax_static_launches;  uint64_t relax_static_bytes;
    double relax_dynamic_kernel_ms; uint64_t relax_dynamic_launches; uint64_t relax_dynamic_bytes;
    double frontier_kernel_ms;      uint64_t frontier_launches;      uint64_t frontier_bytes;
    double copy_ms;               /* sum of copy-stream durations (HIP events)    */
    uint64_t checksum;            /* digest of the result array (og_levels_checksum);
                                     multi-source: sum over the sources' digests    */
} og_run_stats;

/* Replaces bfs_opt (BFS/oneGraph_BFS/CalculateOpt.h:9-211),
 * runGPUbfs_subway (BFS/sbw_BFS/main.cpp:323-506) and runGPUBFS
 * (BFS/usm_BFS/main.cpp:120-217) for one source, using the approach given
 * to og_prepare.  levels (V entries) may be NULL.  stats may be NULL. */
og_status og_bfs(og_graph *g, uint32_t source, uint32_t *levels, og_run_stats *stats);

/* Replaces sssp_opt (SSSP/oneGraph/CalculateOpt.h:12-192),
 * runGPUbfs_subway4 (SSSP/subway/subway.cpp:314-647) and runGPUsssp3
 * (SSSP/usm/sssp_usm.cpp:140-284).  Needs an OG_KIND_WEIGHTED graph. */
og_status og_sssp(og_graph *g, uint32_t source, uint32_t *dist, og_run_stats *stats);

/* Replaces cc_opt (CC/onegraph/CalculateOpt.h:6-144), the Subway CC main
 * (CC/sbw/sbw.cpp:305-490) and the USM CC main (CC/usm/usm.cpp:127-257). */
og_status og_cc(og_graph *g, uint32_t *labels, og_run_stats *stats);

/* -------------------------------------------------------------- tuning */

/* Shapes of the static-region relax and the SSSP round order.  The reference
 * has one shape: one work-item per vertex, grid 1024*108
 * (BFS/oneGraph_BFS/globals.h:16-17). */
typedef struct og_tuning {
    uint32_t t1;     /* bin 0: 1 <= degree <= t1, lpv0 lanes per vertex          */
    uint32_t lpv0;   /* 1, 2, 4, 8, 16, 32 or 64                                  */
    uint32_t t2;     /* bin 1: t1 < degree <= t2, one wave per vertex; above: hubs */
    uint32_t tile;   /* hub tile (edges per workgroup, staged in LDS), 64..4096   */
    uint32_t delta;  /* SSSP near-far bucket width: a round relaxes only vertices
                        whose distance is below the current threshold; 0 = every
                        improved vertex every round (the reference's order)     */
    uint32_t defer;  /* SSSP / CC: 1 = a round with static-region work holds the
                        dynamic-region vertices back, so the link rows are read
                        once the in-HBM part has settled, in denser rounds; 0 =
                        both regions every round (the reference's order);
                        SSSP with delta > 0: 2 = lagged, a dynamic-region
                        vertex relaxes only after the static vertices of its
                        near-far bucket have settled, so its link row is
                        read about once                                          */
    uint32_t dyn_grid; /* workgroups of the dynamic-region relax, 0 = 8192; fewer
                        leave CUs to the static relax it overlaps with           */
} og_tuning;

og_status og_get_tuning(const og_graph *g, og_tuning *t);
/* Installs a tuning on a prepared graph (OG_EINVAL if out of range). */
og_status og_set_tuning(og_graph *g, const og_tuning *t);

#define OG_TUNE_BFS 0u
#define OG_TUNE_SSSP 1u
#define OG_TUNE_CC 2u
#define OG_TUNE_MSBFS 3u   /* og_msbfs on one batch of 64 sources drawn uniformly
                              (SplitMix64 seeded with `source`) among the
                              vertices with out-edges */

/* Autotuner: times `algo` from `source` on the prepared graph under
 * candidate tunings (coordinate descent: defer for SSSP / CC with a dynamic
 * region, delta for SSSP, then t1 with lpv0, t2, tile, and dyn_grid with a
 * dynamic region; the median of `runs` timed runs each), installs the fastest and
 * returns it with its time.  Results do not depend on the tuning.  The
 * workgroup size stays 256 (four 64-lane waves); the grid shapes tuned are
 * the lanes per vertex of bin 0 and the dynamic relax's workgroups. */
og_status og_autotune(og_graph *g, uint32_t algo, uint32_t source, uint32_t runs,
                      og_tuning *best, double *best_ms);

/* Multi-source BFS (no reference counterpart; the reference runs one source
 * per bfs_opt iteration, BFS/oneGraph_BFS/CalculateOpt.h:55-209).  Sources
 * go in batches of 64, each batch one frontier traversal carrying a u64
 * mask per vertex, on the approach given to og_prepare.  Outputs (each may
 * be NULL): levels[i*V + v] in the og_bfs convention, checksums[i] =
 * og_levels_checksum of source i's levels, reached[i] = vertices reached. */
og_status og_msbfs(og_graph *g, const uint32_t *sources, uint32_t num_sources,
                   uint32_t *levels, uint64_t *checksums, uint64_t *reached,
                   og_run_stats *stats);

/* Host-link probe (no reference counterpart: a measurement helper for the
 * roofline of the zero-copy loader).  Allocates `bytes` of pinned host
 * memory and times, over 3 repetitions each: shader reads of it through the
 * device mapping, contiguous (*contig_gbs); the same reads over rows of 1-19
 * words scattered with about a fifth of the array taken (*scattered_gbs,
 * useful bytes per second: the access pattern of a sparse dynamic region);
 * and hipMemcpy host-to-device of the whole buffer (*memcpy_gbs). */
og_status og_probe_host_link(uint64_t bytes, double *contig_gbs, double *scattered_gbs, double *memcpy_gbs);

/* Allocation-free BFS (no counterpart in the reference): og_bfs over a CSR
 * that already sits in device memory, for callers that cannot allocate at
 * run time (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  row_ptr (V + 1 entries, row_ptr[V] = E), cols (E)
 * and levels (V, written in the og_bfs convention) are device pointers;
 * stream is a hipStream_t (NULL: the null stream).  The call allocates
 * nothing: its flags, bin lists, hub tiles and counters live in one static
 * device array of about 10.4 MiB per GPU, held until the process ends,
 * which limits it to V <= 2^21 and E <= 2^26 (OG_EINVAL beyond).  Calls
 * that overlap on one GPU share that array and race; it synchronises the
 * stream once per 4 rounds.  stats may be NULL (kernel timings stay 0). */
og_status og_bfs_device_noalloc(const uint64_t *row_ptr, const uint32_t *cols, uint64_t num_vertices,
                                uint64_t num_edges, uint32_t source, uint32_t *levels, void *stream,
                                og_run_stats *stats);

/* Random-gather probe (no reference counterpart: a measurement helper for
 * the relax kernels, whose per-edge target reads are random 4-byte
 * accesses).  Times 3 launches of 134 M random 4-byte reads each
 * (SplitMix64-drawn indices, 8 in flight per lane) from a device table of
 * table_bytes:
 * *gather_gbs = useful bytes per second / 1e9, *mgathers_per_s = million
 * reads per second. */
og_status og_probe_hbm_gather(uint64_t table_bytes, double *gather_gbs, double *mgathers_per_s);

/* Digest of a BFS/SSSP result array: sum mod 2^64 of mix64(v << 32 | x[v])
 * over entries x[v] != OG_UNREACHED (mix64 = SplitMix64 finaliser).  The
 * device computes the same digest into og_run_stats.checksum. */
og_status og_levels_checksum(const uint32_t *values, uint64_t num_vertices, uint64_t *checksum);

typedef struct og_pr_params {
    double damping;          /* 0.85                                   */
    double tolerance;        /* stop when L1(new - old) < tolerance     */
    uint32_t max_iterations; /* 100                                    */
    uint32_t fp64;           /* 0: FP32 ranks with Kahan pull sums; 1: FP64 */
} og_pr_params;

typedef struct og_pr_stats {
    og_run_stats run;
    uint32_t iterations;
    double final_l1;         /* L1 residual of the last iteration (FP64 reduction) */
} og_pr_stats;

/* Replaces pr_opt (PR/OneGraph/CalculateOpt.h:13-185), pr_calQ of Subway
 * (PR/sbw/sbw.cpp:159-283) and USM (PR/usm/pr.cpp:189-255), with the
 * normalised PageRank definition: r = (1-d)/n + d*(sum_in r[u]/outdeg[u] +
 * dangling/n).  Needs an OG_KIND_PULL graph.  Exactly one of ranks_f32 /
 * ranks_f64 is written, chosen by params->fp64 (the other may be NULL). */
og_status og_pagerank(og_graph *g, const og_pr_params *params, float *ranks_f32,
                      double *ranks_f64, og_pr_stats *stats);

/* --------------------------------------------------------------- serving */

/* Query server over device-resident graphs (no reference counterpart: the
 * reference answers one query per process, BFS/oneGraph_BFS/main.cpp:3-14).
 * BFS queries are batched continuously: each graph has a BFS engine with 64
 * query slots (one bit of per-vertex masks each); queries that arrive while
 * others run join at the next round and share its kernels.  SSSP queries are
 * batched the same way: each weighted graph has an SSSP engine with 16 query
 * slots whose distances lie side by side (one 64-byte line per vertex); each
 * slot runs near-far rounds with its own threshold (light/heavy rounds on a
 * graph prepared with split_weight) and a round's edge reads serve every
 * slot active at the source.  Results equal og_bfs / og_sssp on the same graph and
 * source. */

#define OG_SERVER_MAX_GRAPHS 8
#define OG_QUERY_BFS 0u
#define OG_QUERY_SSSP 1u
#define OG_QUERY_KEEP_VALUES 1u   /* submit flag: keep the result array for og_server_take_values */

typedef struct og_server og_server;

typedef struct og_server_params {
    uint64_t state_budget_bytes;  /* HBM for per-query state (else
                                     OG_EBUDGET): per graph the BFS engine's
                                     masks and 4 kept-levels buffers (40 V
                                     bytes); per weighted graph with
                                     OG_SERVER_SSSP_BATCH or
                                     OG_SERVER_NO_SSSP_LANES the batched SSSP
                                     engine's 16 distance columns, light /
                                     heavy pending masks, active list and work
                                     prefix (92 V + 8 bytes) and chunk table
                                     (4 (16 E / 1024 + 2) bytes); then, unless
                                     OG_SERVER_NO_SSSP_LANES, per weighted
                                     graph as many single-source SSSP lanes
                                     (values, frontier flags, bin lists and
                                     degree prefix: about 17-18 V bytes each)
                                     as the rest holds, up to sssp_lanes, at
                                     least one;
                                     0 = a quarter of free HBM                 */
    uint32_t max_pending;         /* queued queries before submit returns OG_EBUSY; 0 = 2^20 */
    uint32_t flags;               /* OG_SERVER_* bits                       */
    uint32_t deadline_ms;         /* a query not admitted into a slot within
                                     this many ms of its submit completes with
                                     OG_ETIMEDOUT instead (0: no deadline);
                                     admitted queries always run to the end */
    uint32_t sssp_delta;          /* bucket width of the SSSP slots' and
                                     lanes' near-far rounds; 0 = the graph's
                                     tuning delta.
                                     On graphs prepared with split_weight the
                                     slots run light/heavy rounds (wide
                                     buckets pay: 256 with weights in [1, 63]),
                                     else they relax whole rows (16-64)       */
    uint32_t sssp_lanes;          /* single-source SSSP lanes per weighted
                                     graph, 1..8 (0: 2), as far as the budget
                                     holds them                                */
    uint32_t sssp_running_max;    /* SSSP lanes running queries at once over
                                     all graphs (0: no cap): fewer concurrent
                                     SSSP traversals leave more of the GPU to
                                     the BFS engines under load                */
} og_server_params;

/* SSSP queries of a weighted graph run on its single-source lanes
 * (og_sssp's device-driven rounds on each lane's own state), one query per
 * lane at a time.  OG_SERVER_SSSP_BATCH adds the batched 16-slot engine: it
 * admits queries while no lane can take them (all busy, or sssp_running_max
 * reached) and it already runs queries or 8 or more wait.  OG_SERVER_NO_SSSP_LANES runs every SSSP query in the
 * batched engine (no lanes).  The engine's state is allocated and charged
 * only when one of the two flags asks for it. */
#define OG_SERVER_NO_SSSP_LANES 0x1u
#define OG_SERVER_SSSP_BATCH 0x2u
#define OG_PATH_BATCH 0u
#define OG_PATH_LANE 1u

/* Testing only: the server's SSSP engines stop at their first round with
 * OG_EHIP, as after a device error, without touching the GPU (the failure
 * path: held queries fail, later SSSP submits return the error, BFS engines
 * keep serving). */
#define OG_SERVER_TEST_FAIL_SSSP 0x80000000u

typedef struct og_query_result {
    uint64_t id;
    uint32_t graph, kind, source, status;   /* status: og_status of the query */
    uint64_t reached, traversed_edges, checksum;  /* checksum = og_levels_checksum of the result */
    double queue_ms;              /* submit -> admitted into a slot      */
    double latency_ms;            /* submit -> result ready              */
    uint32_t rounds;              /* engine rounds the query took part in */
    uint32_t path;                /* what ran it: OG_PATH_BATCH (a slot of a
                                     batch engine) or OG_PATH_LANE (an SSSP
                                     single-source lane)                     */
} og_query_result;

typedef struct og_server_info {
    uint32_t num_graphs;
    uint32_t bfs_slots[OG_SERVER_MAX_GRAPHS];   /* BFS queries in flight at once per graph */
    uint32_t sssp_slots[OG_SERVER_MAX_GRAPHS];  /* SSSP queries in flight at once per graph
                                                   (batched slots + single-source lanes)   */
    uint64_t state_bytes;         /* HBM held for per-query state */
    uint64_t submitted, completed, rejected, pending, in_flight, rounds;
} og_server_info;

/* Graphs must be prepared with the whole graph in device memory (og_prepare
 * with a budget holding every edge and advisek 1, or OG_USM); they stay
 * reserved (OG_EBUSY for other calls) until og_server_destroy.  Slot counts
 * follow params->state_budget_bytes (OG_EBUDGET when a graph gets none). */
og_status og_server_create(og_graph *const *graphs, uint32_t num_graphs,
                           const og_server_params *params, og_server **out);
og_status og_server_submit(og_server *srv, uint32_t graph, uint32_t kind, uint32_t source,
                           uint32_t flags, uint64_t *id);
/* Returns up to max_results finished queries; waits up to timeout_ms for the
 * first one (0: no wait, < 0: wait forever). */
og_status og_server_poll(og_server *srv, og_query_result *results, uint32_t max_results,
                         uint32_t *count, int32_t timeout_ms);
/* Copies the V result values of a finished OG_QUERY_KEEP_VALUES query. */
og_status og_server_take_values(og_server *srv, uint64_t id, uint32_t *values);
og_status og_server_get_info(og_server *srv, og_server_info *info);
/* Stops the engines (queries still queued or running are dropped). */
og_status og_server_destroy(og_server *srv);

#ifdef __cplusplus
}
#endif

#endif /* ONEGRAPH_H */
