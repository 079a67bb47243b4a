// C++ host layer above the C ABI, mirroring the reference's driver functions
// (same names and argument meaning):
//   bfs_opt(path, sourceNode, adviseRate)   BFS/oneGraph_BFS/CalculateOpt.h:9
//   sssp_opt(path, sourceNode, adviseRate)  SSSP/oneGraph/CalculateOpt.h:12
//   cc_opt(path, adviseRate)                CC/onegraph/CalculateOpt.h:6
//   pr_opt(path, adviseRate)                PR/OneGraph/CalculateOpt.h:13
// Differences: they return an og_status instead of exiting, honour the
// source (the BFS reference replaces it with findbignode, CalculateOpt.h:
// 26-32,60, kept behind Options::bignode_runs), take the approach and the
// explicit edge budget from Options, and can write the result array.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "onegraph.h"

namespace onegraph {

struct Options {
    og_approach approach = OG_ASCETIC;
    uint64_t edge_budget_bytes = 0;  // 0: reference default budget
    double budget_fraction = 0;      // > 0: budget = fraction * CSR edge bytes
    uint32_t bignode_runs = 0;       // > 0: sources from findbignode, like the reference BFS
    uint32_t repeat = 1;
    bool json = false;
    bool pr_fp64 = false;
    double pr_tolerance = 1e-7;
    uint32_t pr_max_iterations = 100;
    std::string out_path;            // optional binary dump of the result array
    std::string generate;            // "rmat:SCALE:EF:SEED[:WSEED]" or "er:N:AVG:SEED" instead of -f
    bool text_input = false;         // the input path is a text edge list / MatrixMarket file
    uint32_t edge_flags = 0;         // OG_EDGES_* for text input
    uint64_t weight_seed = 1;        // text input without weights, weighted algorithm: seeded [1, 63]
    std::string check_path;          // result array to compare with (same layout as out_path)
    double check_tol = 1e-6;         // PageRank: max-abs tolerance of the check
    // engine options with no reference counterpart
    uint32_t plan_flags = 0;         // OG_PLAN_* (zero-copy loader, direction-optimising BFS, ...)
    uint32_t split_weight = 0;       // og_plan_params::split_weight (SSSP light/heavy rows)
    int64_t delta = -1;              // >= 0: og_tuning::delta (SSSP near-far bucket width)
    bool autotune = false;           // og_autotune from the (first) source before the runs
};

// Result-array file (--out / --check): V little-endian values, u32 levels /
// distances / labels (UINT32_MAX unreached), f32 or f64 ranks.  Compares
// `got` with the file: exact for integers, max-abs <= tol for ranks.
// OG_OK on a match; OG_EFORMAT (with a message on stderr naming the first
// differing vertex) on a mismatch; OG_EIO when the file is unreadable.
og_status check_result(const std::string &path, const void *got, uint64_t n, int bytes_per_value, bool is_float,
                       double tol);

// Loads `path` as `kind` (the reference reader, GraphMeta::readDataFromFile,
// BFS/oneGraph_BFS/GraphMeta.h:125-145), or a text edge list with
// opt.text_input, or builds opt.generate instead; prints the reference's
// "readDataFromFile" timing line.  The caller frees *out with og_graph_free.
og_status open_graph(const std::string &path, og_graph_kind kind, const Options &opt, og_graph **out);

// Prints the reference's labelled lines (TimeRecord "<name> time is X ms",
// nodeSum, staticnodeSum, "move overload size") and optionally one JSON line.
void print_report(const char *algo, const og_run_stats &st, const og_plan_info &plan, bool json, FILE *f = stdout);

og_status bfs_opt(const std::string &path, uint32_t sourceNode, float adviseRate, const Options &opt = Options());
og_status sssp_opt(const std::string &path, uint32_t sourceNode, double adviseRate, const Options &opt = Options());
og_status cc_opt(const std::string &path, float adviseRate, const Options &opt = Options());
og_status pr_opt(const std::string &path, double adviseRate, const Options &opt = Options());

}  // namespace onegraph
