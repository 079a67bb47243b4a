// Driver functions of the C++ host layer (declared in onegraph_host.hpp):
// bfs_opt / sssp_opt / cc_opt / pr_opt mirror the reference's drivers
// (BFS/oneGraph_BFS/CalculateOpt.h:9, SSSP/oneGraph/CalculateOpt.h:12,
// CC/onegraph/CalculateOpt.h:6, PR/OneGraph/CalculateOpt.h:13) on top of the
// C ABI, with the result checker and the reference's labelled report.  They
// are part of libonegraph.so, so C++ callers link them from the library; the
// onegraph command line (cli.cpp) is one such caller.
#include <chrono>
#include <cinttypes>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

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
