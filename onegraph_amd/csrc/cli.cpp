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
// exit status 2; library failures exit with status 1.
#include <chrono>
#include <cinttypes>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "onegraph_host.hpp"

namespace onegraph {

namespace {

struct GraphPtr {
    og_graph *g = nullptr;
    ~GraphPtr() {
        if (g) og_graph_free(g);
    }
};

og_status report_error(og_status st, const char *what) {
    if (st != OG_OK) fprintf(stderr, "onegraph: %s failed: %s\n", what, og_last_error());
    return st;
}

// Loads `path` (or builds the --generate graph) as `kind`.
og_status open_graph(const std::string &path, og_graph_kind kind, const Options &opt, og_graph **out) {
    auto t0 = std::chrono::steady_clock::now();
    og_status st;
    if (!opt.generate.empty()) {
        char fam[16] = {0};
        unsigned long long a = 0, b = 0, c = 0, d = 0;
        int n = sscanf(opt.generate.c_str(), "%15[a-z]:%llu:%llu:%llu:%llu", fam, &a, &b, &c, &d);
        og_graph *g = nullptr;
        if (n >= 4 && !strcmp(fam, "rmat")) {
            st = og_graph_generate_rmat((uint32_t)a, (uint32_t)b, 0.57, 0.19, 0.19, c,
                                        kind == OG_KIND_WEIGHTED ? (n >= 5 ? d : c + 1000) : 0, &g);
        } else if (n >= 4 && !strcmp(fam, "er")) {
            st = og_graph_generate_er(a, (uint32_t)b, c, &g);
            if (st == OG_OK && kind == OG_KIND_WEIGHTED) {
                og_graph *w = nullptr;
                st = og_graph_add_weights(g, c + 1000, &w);
                og_graph_free(g);
                g = w;
            }
        } else {
            fprintf(stderr, "onegraph: bad --generate spec '%s'\n", opt.generate.c_str());
            return OG_EINVAL;
        }
        if (st == OG_OK && kind == OG_KIND_PULL) {
            og_graph *p = nullptr;
            st = og_graph_to_pull(g, &p);
            og_graph_free(g);
            g = p;
        }
        if (st != OG_OK) return report_error(st, "generate");
        *out = g;
    } else if (opt.text_input) {
        // text edge list: weights from the file's third column when asked
        // for, else seeded [1, 63] weights for the weighted algorithm
        const bool text_w = (opt.edge_flags & OG_EDGES_WEIGHTED) != 0;
        og_graph *g = nullptr;
        st = og_graph_load_edgelist(path.c_str(), kind == OG_KIND_WEIGHTED ? opt.edge_flags
                                                                              : opt.edge_flags & ~OG_EDGES_WEIGHTED,
                                    &g);
        if (st == OG_OK && kind == OG_KIND_WEIGHTED && !text_w) {
            og_graph *w = nullptr;
            st = og_graph_add_weights(g, opt.weight_seed, &w);
            og_graph_free(g);
            g = w;
        }
        if (st == OG_OK && kind == OG_KIND_PULL) {
            og_graph *p = nullptr;
            st = og_graph_to_pull(g, &p);
            og_graph_free(g);
            g = p;
        }
        if (st != OG_OK) return report_error(st, "convert");
        *out = g;
    } else {
        st = og_graph_load(path.c_str(), kind, out);
        if (st != OG_OK) return report_error(st, "load");
    }
    double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    og_graph_info info;
    og_graph_get_info(*out, &info);
    printf("readDataFromFile %.0f ms\n", ms);
    printf("vertexArrSize = %" PRIu64 " edgeArrSize = %" PRIu64 "\n", info.num_vertices, info.num_edges);
    return OG_OK;
}

og_status prepare(og_graph *g, float advisek, uint32_t param_size, const Options &opt, og_plan_info *plan) {
    og_graph_info info;
    og_graph_get_info(g, &info);
    og_plan_params p{};
    p.advisek = advisek;
    p.param_size = param_size;
    p.edge_budget_bytes = opt.edge_budget_bytes;
    p.flags = opt.plan_flags;
    p.split_weight = info.kind == OG_KIND_WEIGHTED ? opt.split_weight : 0;
    if (opt.budget_fraction > 0) {
        const uint64_t eb = info.kind == OG_KIND_WEIGHTED ? 8 : 4;
        p.edge_budget_bytes = (uint64_t)(opt.budget_fraction * (double)(info.num_edges * eb));
        if (!p.edge_budget_bytes) p.edge_budget_bytes = eb;
    }
    og_status st = og_prepare(g, opt.approach, &p, plan);
    if (st != OG_OK) return report_error(st, "prepare");
    if (opt.delta >= 0 && info.kind != OG_KIND_PULL) {
        og_tuning t;
        og_get_tuning(g, &t);
        t.delta = (uint32_t)opt.delta;
        st = og_set_tuning(g, &t);
        if (st != OG_OK) return report_error(st, "tuning");
    }
    printf("approach %s: total_gpu_size %" PRIu64 " max_partition_size %" PRIu64 " partOverloadSize %" PRIu64
           " overloadSize %" PRIu64 " adviseK %f\n",
           opt.approach == OG_ASCETIC ? "ascetic" : opt.approach == OG_SUBWAY ? "subway" : "usm",
           plan->total_budget_edges, plan->static_edges, plan->staging_edges, plan->overload_edges,
           (double)plan->advisek_used);
    return OG_OK;
}

og_status dump(const Options &opt, const void *data, size_t bytes) {
    if (opt.out_path.empty()) return OG_OK;
    FILE *f = fopen(opt.out_path.c_str(), "wb");
    if (!f || fwrite(data, 1, bytes, f) != bytes) {
        if (f) fclose(f);
        fprintf(stderr, "onegraph: cannot write %s\n", opt.out_path.c_str());
        return OG_EIO;
    }
    fclose(f);
    return OG_OK;
}

// --out and --check of a result array.
og_status finish(const Options &opt, const void *data, uint64_t n, int bytes, bool is_float) {
    og_status st = dump(opt, data, (size_t)n * bytes);
    if (st != OG_OK || opt.check_path.empty()) return st;
    st = check_result(opt.check_path, data, n, bytes, is_float, opt.check_tol);
    if (st == OG_OK) printf("check: %" PRIu64 " values match %s\n", n, opt.check_path.c_str());
    return st;
}

using RunFn = og_status (*)(og_graph *, uint32_t, uint32_t *, og_run_stats *);

og_status bfs_fn(og_graph *g, uint32_t s, uint32_t *o, og_run_stats *st) { return og_bfs(g, s, o, st); }
og_status sssp_fn(og_graph *g, uint32_t s, uint32_t *o, og_run_stats *st) { return og_sssp(g, s, o, st); }

og_status run_sources(const char *algo, og_graph *g, uint32_t source, const Options &opt, RunFn fn,
                      const og_plan_info &plan) {
    og_graph_info info;
    og_graph_get_info(g, &info);
    std::vector<uint32_t> sources;
    if (opt.bignode_runs) {
        uint32_t last = 0;  // reference: src[i] = findbignode(previous), starting at 0
        for (uint32_t i = 0; i < opt.bignode_runs; ++i) {
            og_graph_find_big_node(g, last, &last);
            sources.push_back(last);
        }
    } else {
        sources.push_back(source);
    }
    std::vector<uint32_t> out(info.num_vertices);
    if (opt.autotune) {
        og_tuning best;
        double ms = 0;
        const uint32_t which = fn == sssp_fn ? OG_TUNE_SSSP : OG_TUNE_BFS;
        og_status s = og_autotune(g, which, sources[0], 3, &best, &ms);
        if (s != OG_OK) return report_error(s, "autotune");
        printf("autotune: t1 %u lpv0 %u t2 %u tile %u delta %u defer %u dyn_grid %u (%.3f ms)\n", best.t1, best.lpv0,
               best.t2, best.tile, best.delta, best.defer, best.dyn_grid, ms);
    }
    for (size_t t = 0; t < sources.size(); ++t) {
        printf("----------testIndex = %zu-----------\nsourceNode %u\n", t, sources[t]);
        for (uint32_t r = 0; r < opt.repeat; ++r) {
            og_run_stats st;
            og_status s = fn(g, sources[t], out.data(), &st);
            if (s != OG_OK) return report_error(s, algo);
            print_report(algo, st, plan, opt.json);
        }
    }
    return finish(opt, out.data(), out.size(), 4, false);
}

}  // namespace

og_status check_result(const std::string &path, const void *got, uint64_t n, int bytes_per_value, bool is_float,
                       double tol) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) {
        fprintf(stderr, "onegraph: cannot read %s\n", path.c_str());
        return OG_EIO;
    }
    std::vector<unsigned char> want((size_t)n * bytes_per_value);
    const size_t rd = want.empty() ? 0 : fread(want.data(), 1, want.size(), f);
    const bool longer = fgetc(f) != EOF;
    fclose(f);
    if (rd != want.size() || longer) {
        fprintf(stderr, "check: %s holds %s than %" PRIu64 " values of %d bytes\n", path.c_str(),
                longer ? "more" : "fewer", n, bytes_per_value);
        return OG_EFORMAT;
    }
    uint64_t bad = 0, first = 0;
    double worst = 0;
    for (uint64_t v = 0; v < n; ++v) {
        bool diff;
        if (is_float) {
            double a, b;
            if (bytes_per_value == 8) {
                a = ((const double *)got)[v];
                b = ((const double *)want.data())[v];
            } else {
                a = ((const float *)got)[v];
                b = ((const float *)want.data())[v];
            }
            const double e = a > b ? a - b : b - a;
            worst = e > worst ? e : worst;
            diff = !(e <= tol);
        } else {
            diff = ((const uint32_t *)got)[v] != ((const uint32_t *)want.data())[v];
        }
        if (diff && bad++ == 0) first = v;
    }
    if (!bad) {
        if (is_float) printf("check: max-abs %.3e <= %.1e\n", worst, tol);
        return OG_OK;
    }
    if (is_float)
        fprintf(stderr, "check: %" PRIu64 " of %" PRIu64 " values differ by more than %.1e (max-abs %.3e), first at vertex %" PRIu64 "\n",
                bad, n, tol, worst, first);
    else
        fprintf(stderr, "check: %" PRIu64 " of %" PRIu64 " values differ, first at vertex %" PRIu64 ": got %u, file %u\n",
                bad, n, first, ((const uint32_t *)got)[first], ((const uint32_t *)want.data())[first]);
    return OG_EFORMAT;
}

void print_report(const char *algo, const og_run_stats &st, const og_plan_info &plan, bool json, FILE *f) {
    fprintf(f, "totalProcess time is %.3f ms\n", st.total_ms);
    fprintf(f, "preProcess time is %.3f ms\n", st.pre_ms);
    fprintf(f, "staticProcess time is %.3f ms\n", st.static_ms);
    fprintf(f, "overloadProcess time is %.3f ms\n", st.overload_ms);
    fprintf(f, "overloadMoveProcess time is %.3f ms\n", st.overload_move_ms);
    fprintf(f, "nodeSum : %" PRIu64 "\n", st.node_sum);
    fprintf(f, "move overload size : %" PRIu64 " MB \n", st.h2d_bytes >> 20);
    fprintf(f, "staticnodeSum : %" PRIu64 "\n", st.static_node_sum);
    fprintf(f, "nodesum: %" PRIu64 "\n", st.reached);
    const double gteps = st.total_ms > 0 ? (double)st.traversed_edges / (st.total_ms * 1e-3) / 1e9 : 0.0;
    fprintf(f, "iterations %u, GTEPS %.4f\n", st.iterations, gteps);
    if (json) {
        fprintf(f,
                "{\"algo\":\"%s\",\"approach\":%u,\"iterations\":%u,\"total_ms\":%.4f,\"gteps\":%.5f,"
                "\"reached\":%" PRIu64 ",\"traversed_edges\":%" PRIu64 ",\"work_edges\":%" PRIu64
                ",\"updates\":%" PRIu64 ",\"node_sum\":%" PRIu64 ",\"h2d_bytes\":%" PRIu64 ",\"chunks\":%" PRIu64
                ",\"copy_ms\":%.4f,\"gather_ms\":%.4f,\"relax_static_kernel_ms\":%.4f,"
                "\"relax_dynamic_kernel_ms\":%.4f,\"frontier_kernel_ms\":%.4f,\"static_edges\":%" PRIu64
                ",\"staging_edges\":%" PRIu64 ",\"device_bytes\":%" PRIu64 "}\n",
                algo, st.approach, st.iterations, st.total_ms, gteps, st.reached, st.traversed_edges, st.work_edges,
                st.updates, st.node_sum, st.h2d_bytes, st.chunks, st.copy_ms, st.gather_ms, st.relax_static_kernel_ms,
                st.relax_dynamic_kernel_ms, st.frontier_kernel_ms, plan.static_edges, plan.staging_edges,
                plan.device_bytes);
    }
}

og_status bfs_opt(const std::string &path, uint32_t sourceNode, float adviseRate, const Options &opt) {
    printf("======bfs_opt=======\n");
    GraphPtr g;
    og_plan_info plan;
    og_status st = open_graph(path, OG_KIND_PLAIN, opt, &g.g);
    if (st == OG_OK) st = prepare(g.g, adviseRate, 17, opt, &plan);
    if (st == OG_OK) st = run_sources("bfs", g.g, sourceNode, opt, bfs_fn, plan);
    return st;
}

og_status sssp_opt(const std::string &path, uint32_t sourceNode, double adviseRate, const Options &opt) {
    printf("======sssp_opt=======\n");
    GraphPtr g;
    og_plan_info plan;
    og_status st = open_graph(path, OG_KIND_WEIGHTED, opt, &g.g);
    if (st == OG_OK) st = prepare(g.g, (float)adviseRate, 17, opt, &plan);
    if (st == OG_OK) st = run_sources("sssp", g.g, sourceNode, opt, sssp_fn, plan);
    return st;
}

og_status cc_opt(const std::string &path, float adviseRate, const Options &opt) {
    printf("======cc_opt=======\n");
    GraphPtr g;
    og_plan_info plan;
    og_status st = open_graph(path, OG_KIND_PLAIN, opt, &g.g);
    if (st == OG_OK) st = prepare(g.g, adviseRate, 17, opt, &plan);
    if (st != OG_OK) return st;
    og_graph_info info;
    og_graph_get_info(g.g, &info);
    std::vector<uint32_t> labels(info.num_vertices);
    for (uint32_t r = 0; r < opt.repeat; ++r) {
        og_run_stats rs;
        st = og_cc(g.g, labels.data(), &rs);
        if (st != OG_OK) return report_error(st, "cc");
        print_report("cc", rs, plan, opt.json);
    }
    return finish(opt, labels.data(), labels.size(), 4, false);
}

og_status pr_opt(const std::string &path, double adviseRate, const Options &opt) {
    printf("======pr_opt=======\n");
    GraphPtr g;
    og_plan_info plan;
    og_status st = open_graph(path, OG_KIND_PULL, opt, &g.g);
    if (st == OG_OK) st = prepare(g.g, (float)adviseRate, 19, opt, &plan);
    if (st != OG_OK) return st;
    og_graph_info info;
    og_graph_get_info(g.g, &info);
    og_pr_params p{0.85, opt.pr_tolerance, opt.pr_max_iterations, opt.pr_fp64 ? 1u : 0u};
    std::vector<float> r32(opt.pr_fp64 ? 0 : info.num_vertices);
    std::vector<double> r64(opt.pr_fp64 ? info.num_vertices : 0);
    for (uint32_t r = 0; r < opt.repeat; ++r) {
        og_pr_stats ps;
        st = og_pagerank(g.g, &p, opt.pr_fp64 ? nullptr : r32.data(), opt.pr_fp64 ? r64.data() : nullptr, &ps);
        if (st != OG_OK) return report_error(st, "pagerank");
        print_report("pr", ps.run, plan, opt.json);
        printf("pagerank iterations %u final L1 %.3e (%s)\n", ps.iterations, ps.final_l1, opt.pr_fp64 ? "fp64" : "fp32");
    }
    return opt.pr_fp64 ? finish(opt, r64.data(), r64.size(), 8, true) : finish(opt, r32.data(), r32.size(), 4, true);
}

}  // namespace onegraph

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
