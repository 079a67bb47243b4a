// onegraph -- drop-in command line for the reference programs.
//
//   onegraph --type bfs|sssp|cc|pr -f <csr file> [-r <source>] [--advisek K]
//            [--approach ascetic|subway|usm] [--edge-budget BYTES | --budget-fraction F]
//            [--bignode-runs N] [--repeat N] [--json] [--out FILE]
//            [--fp64] [--pr-tol T] [--pr-iters N] [--generate SPEC [--save FILE]]
//            [--text [--text-weights] [--symmetrize] [--one-based] ...] [--check FILE]
//
// Accepts every spelling the reference programs use: --type / -type
// (BFS/oneGraph_BFS/ArgumentParser.h:48, SSSP/oneGraph/ArgumentParser.h),
// -f, -r, --method / -method (ignored, as in the reference), and --advisek /
// --adviseK / -adviseK (README.md:31, ArgumentParser.h:55).  Unlike the
// reference parser (pairs from argv[1] in steps of 2, exit(0) on error,
// ArgumentParser.h:40-73) unknown flags and missing values are errors with
// exit status 2; library failures exit with status 1.  The drivers it calls
// (onegraph::bfs_opt ... pr_opt) live in libonegraph.so (drivers.cpp).
#include <chrono>
#include <cinttypes>
#include <cstdlib>
#include <cstring>
#include <memory>

#include <string>
#include <vector>

#include "onegraph_host.hpp"

static void usage(FILE *f) {
    fprintf(f,
            "usage: onegraph --type bfs|sssp|cc|pr -f FILE [-r SOURCE] [--advisek K]\n"
            "                [--approach ascetic|subway|usm] [--edge-budget BYTES | --budget-fraction F]\n"
            "                [--bignode-runs N] [--repeat N] [--json] [--out FILE]\n"
            "                [--fp64] [--pr-tol T] [--pr-iters N]\n"
            "                [--generate rmat:SCALE:EF:SEED[:WSEED] | er:N:AVG:SEED] [--save FILE]\n"
            "                [--text [--text-weights] [--symmetrize] [--keep-self-loops] [--one-based]\n"
            "                 [--weight-seed S]] [--check FILE [--check-tol T]]\n"
            "                [--loader copy|zerocopy] [--delta W] [--split-weight W] [--direction-opt]\n"
            "                [--autotune]\n"
            "  --loader zerocopy: the dynamic region is read in place over the host link;\n"
            "  --delta W: SSSP near-far bucket width; --split-weight W: SSSP rows store\n"
            "  edges lighter than W first (delta-stepping light/heavy); --direction-opt:\n"
            "  bottom-up BFS rounds (whole graph in HBM); --autotune: tune bins, delta and\n"
            "  grid shapes from the first source before the runs.\n"
            "  --text: -f is a text edge list (\"u v [w]\" lines, #/%% comments) or a\n"
            "  MatrixMarket file; with --save FILE it is converted to the reference CSR\n"
            "  format of --type (bfs/cc plain, sssp weighted, pr pull) and written.\n"
            "  --check FILE compares the result array with FILE (the --out layout);\n"
            "  a mismatch exits with status 3.\n"
            "       onegraph --compare A B [--dtype u32|f32|f64] [--check-tol T]\n"
            "  compares two result files without a GPU (status 3 on a mismatch).\n");
}

// --compare A B: two result files of the --out layout, on the host.
static int compare_files(const std::string &a, const std::string &b, const std::string &dtype, double tol) {
    const int bytes = dtype == "f64" ? 8 : 4;
    const bool fl = dtype != "u32";
    if (dtype != "u32" && dtype != "f32" && dtype != "f64") {
        fprintf(stderr, "onegraph: --dtype must be u32, f32 or f64\n");
        return 2;
    }
    FILE *f = fopen(a.c_str(), "rb");
    if (!f) {
        fprintf(stderr, "onegraph: cannot read %s\n", a.c_str());
        return 1;
    }
    std::vector<unsigned char> buf;
    unsigned char chunk[1 << 16];
    size_t n;
    while ((n = fread(chunk, 1, sizeof chunk, f)) > 0) buf.insert(buf.end(), chunk, chunk + n);
    fclose(f);
    if (buf.size() % bytes) {
        fprintf(stderr, "onegraph: %s is not a whole number of %d-byte values\n", a.c_str(), bytes);
        return 1;
    }
    const og_status st = onegraph::check_result(b, buf.data(), buf.size() / bytes, bytes, fl, tol);
    if (st == OG_OK) printf("check: %zu values match\n", buf.size() / bytes);
    return st == OG_OK ? 0 : st == OG_EFORMAT ? 3 : 1;
}

int main(int argc, char **argv) {
    std::string algo, input, save, cmp_a, cmp_b, dtype = "u32";
    long long source = 0;
    double advisek = 0.0;
    onegraph::Options opt;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto need = [&](const char *flag) -> const char * {
            if (i + 1 >= argc) {
                fprintf(stderr, "onegraph: %s needs a value\n", flag);
                usage(stderr);
                exit(2);
            }
            return argv[++i];
        };
        if (a == "-h" || a == "--help") {
            usage(stdout);
            return 0;
        } else if (a == "--type" || a == "-type") {
            algo = need("--type");
        } else if (a == "-f" || a == "--input") {
            input = need("-f");
        } else if (a == "-r" || a == "--source") {
            source = atoll(need("-r"));
        } else if (a == "--advisek" || a == "--adviseK" || a == "-adviseK" || a == "-advisek") {
            advisek = atof(need("--advisek"));
        } else if (a == "--method" || a == "-method") {
            (void)need("--method");
        } else if (a == "--approach") {
            std::string v = need("--approach");
            if (v == "ascetic" || v == "onegraph") opt.approach = OG_ASCETIC;
            else if (v == "subway") opt.approach = OG_SUBWAY;
            else if (v == "usm") opt.approach = OG_USM;
            else {
                fprintf(stderr, "onegraph: unknown approach %s\n", v.c_str());
                return 2;
            }
        } else if (a == "--edge-budget") {
            opt.edge_budget_bytes = strtoull(need("--edge-budget"), nullptr, 10);
        } else if (a == "--budget-fraction") {
            opt.budget_fraction = atof(need("--budget-fraction"));
        } else if (a == "--bignode-runs") {
            opt.bignode_runs = (uint32_t)atoi(need("--bignode-runs"));
        } else if (a == "--repeat") {
            opt.repeat = (uint32_t)std::max(1, atoi(need("--repeat")));
        } else if (a == "--json") {
            opt.json = true;
        } else if (a == "--out") {
            opt.out_path = need("--out");
        } else if (a == "--fp64") {
            opt.pr_fp64 = true;
        } else if (a == "--pr-tol") {
            opt.pr_tolerance = atof(need("--pr-tol"));
        } else if (a == "--pr-iters") {
            opt.pr_max_iterations = (uint32_t)atoi(need("--pr-iters"));
        } else if (a == "--generate") {
            opt.generate = need("--generate");
        } else if (a == "--save") {
            save = need("--save");
        } else if (a == "--text" || a == "--convert") {
            opt.text_input = true;
        } else if (a == "--text-weights") {
            opt.edge_flags |= OG_EDGES_WEIGHTED;
        } else if (a == "--symmetrize") {
            opt.edge_flags |= OG_EDGES_SYMMETRIZE;
        } else if (a == "--keep-self-loops") {
            opt.edge_flags |= OG_EDGES_KEEP_SELF_LOOPS;
        } else if (a == "--one-based") {
            opt.edge_flags |= OG_EDGES_ONE_BASED;
        } else if (a == "--weight-seed") {
            opt.weight_seed = strtoull(need("--weight-seed"), nullptr, 10);
        } else if (a == "--check") {
            opt.check_path = need("--check");
        } else if (a == "--check-tol") {
            opt.check_tol = atof(need("--check-tol"));
        } else if (a == "--compare") {
            cmp_a = need("--compare");
            cmp_b = need("--compare");
        } else if (a == "--loader") {
            const std::string v = need("--loader");
            if (v == "zerocopy") opt.plan_flags |= OG_PLAN_ZERO_COPY;
            else if (v != "copy") {
                fprintf(stderr, "onegraph: --loader takes copy or zerocopy\n");
                return 2;
            }
        } else if (a == "--direction-opt") {
            opt.plan_flags |= OG_PLAN_DIRECTION_OPT;
        } else if (a == "--delta") {
            opt.delta = atoll(need("--delta"));
        } else if (a == "--split-weight") {
            opt.split_weight = (uint32_t)atoi(need("--split-weight"));
        } else if (a == "--autotune") {
            opt.autotune = true;
        } else if (a == "--dtype") {
            dtype = need("--dtype");
        } else {
            fprintf(stderr, "onegraph: unknown argument %s\n", a.c_str());
            usage(stderr);
            return 2;
        }
    }
    if (!cmp_a.empty()) return compare_files(cmp_a, cmp_b, dtype, opt.check_tol);
    if (algo.empty()) algo = "bfs";  // reference default (BFS/oneGraph_BFS/main.cpp:5-7)
    if (input.empty() && opt.generate.empty()) {
        fprintf(stderr, "onegraph: -f FILE (or --generate SPEC) is required\n");
        usage(stderr);
        return 2;
    }
    if (source < 0 || source > 0xFFFFFFFELL) {
        fprintf(stderr, "onegraph: source out of range\n");
        return 2;
    }
    printf("arguments.algo %s\n", algo.c_str());
    if (!save.empty()) {
        og_graph *g = nullptr;
        og_graph_kind kind = algo == "sssp" ? OG_KIND_WEIGHTED : algo == "pr" ? OG_KIND_PULL : OG_KIND_PLAIN;
        onegraph::Options gen = opt;
        og_status st = onegraph::open_graph(input, kind, gen, &g);
        if (st != OG_OK) return 1;
        st = og_graph_save(g, save.c_str());
        og_graph_free(g);
        if (st != OG_OK) {
            fprintf(stderr, "onegraph: save failed: %s\n", og_last_error());
            return 1;
        }
        printf("saved %s\n", save.c_str());
        return 0;
    }
    og_status st;
    if (algo == "bfs") st = onegraph::bfs_opt(input, (uint32_t)source, (float)advisek, opt);
    else if (algo == "sssp") st = onegraph::sssp_opt(input, (uint32_t)source, advisek, opt);
    else if (algo == "cc") st = onegraph::cc_opt(input, (float)advisek, opt);
    else if (algo == "pr") st = onegraph::pr_opt(input, advisek, opt);
    else {
        fprintf(stderr, "onegraph: unknown --type %s\n", algo.c_str());
        return 2;
    }
    if (st == OG_EFORMAT && !opt.check_path.empty()) return 3;  // result check failed
    return st == OG_OK ? 0 : 1;
}
