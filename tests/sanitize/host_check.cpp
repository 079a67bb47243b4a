// Host-side checks of libonegraph's CPU code under AddressSanitizer and
// UndefinedBehaviorSanitizer (built with g++ by tests/test_sanitize.py from
// onegraph_amd/csrc/host_graph.cpp and drivers.cpp; no device code).  Without
// a GPU the device entry points the drivers call are stand-ins below that
// report OG_EHIP, as the library does on a host without a GPU, so the drivers'
// host paths (graph loading, the result checker, error reporting) run too.
//
// Exercises: the CSR reader and writer (reference layouts) on good,
// truncated and corrupt files; og_graph_from_csr / from_edges / the text
// edge-list and MatrixMarket readers; the R-MAT and ER generators, the pull
// layout and weights; findbignode; the static-region planner; the drivers.
// Exit status 0 when every check holds; the sanitizers abort otherwise.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "og_internal.h"
#include "onegraph.h"
#include "onegraph_host.hpp"

// ---- stand-ins for the device side (engine.hip, pagerank.hip)
namespace og {
void device_state_free(DeviceState *) {}
}  // namespace og
extern "C" {
og_status og_prepare(og_graph *, og_approach, const og_plan_params *, og_plan_info *) {
    return og::fail(OG_EHIP, "no HIP device is visible");
}
og_status og_bfs(og_graph *, uint32_t, uint32_t *, og_run_stats *) { return og::fail(OG_ESTATE, "not prepared"); }
og_status og_sssp(og_graph *, uint32_t, uint32_t *, og_run_stats *) { return og::fail(OG_ESTATE, "not prepared"); }
og_status og_cc(og_graph *, uint32_t *, og_run_stats *) { return og::fail(OG_ESTATE, "not prepared"); }
og_status og_pagerank(og_graph *, const og_pr_params *, float *, double *, og_pr_stats *) {
    return og::fail(OG_ESTATE, "not prepared");
}
og_status og_autotune(og_graph *, uint32_t, uint32_t, uint32_t, og_tuning *, double *) {
    return og::fail(OG_ESTATE, "not prepared");
}
og_status og_get_tuning(const og_graph *, og_tuning *) { return og::fail(OG_ESTATE, "not prepared"); }
og_status og_set_tuning(og_graph *, const og_tuning *) { return og::fail(OG_ESTATE, "not prepared"); }
}

static int failures = 0;
#define CHECK(c)                                                      \
    do {                                                              \
        if (!(c)) {                                                   \
            std::fprintf(stderr, "CHECK failed at %d: %s\n", __LINE__, #c); \
            ++failures;                                               \
        }                                                             \
    } while (0)

static std::vector<uint64_t> ptr_of(og_graph *g, uint64_t V) {
    std::vector<uint64_t> p(V);
    og_graph_export(g, p.data(), nullptr, nullptr, nullptr);
    return p;
}

int main(int argc, char **argv) {
    const std::string dir = argc > 1 ? argv[1] : "/tmp";
    og_graph_info gi;

    // generators, save / load round trip in each layout
    og_graph *w = nullptr, *p = nullptr, *er = nullptr, *pl = nullptr;
    CHECK(og_graph_generate_rmat(10, 8, 0.57, 0.19, 0.19, 11, 12, &w) == OG_OK);
    CHECK(og_graph_get_info(w, &gi) == OG_OK && gi.kind == OG_KIND_WEIGHTED && gi.num_vertices == 1024);
    const uint64_t V = gi.num_vertices, E = gi.num_edges;
    std::vector<uint64_t> ptr(V);
    std::vector<uint32_t> to(E), wt(E);
    CHECK(og_graph_export(w, ptr.data(), to.data(), wt.data(), nullptr) == OG_OK);
    const std::string fw = dir + "/w.bwcsr";
    CHECK(og_graph_save(w, fw.c_str()) == OG_OK);
    og_graph *w2 = nullptr;
    CHECK(og_graph_load(fw.c_str(), OG_KIND_WEIGHTED, &w2) == OG_OK);
    std::vector<uint32_t> to2(E), wt2(E);
    CHECK(og_graph_export(w2, nullptr, to2.data(), wt2.data(), nullptr) == OG_OK);
    CHECK(to == to2 && wt == wt2 && ptr_of(w2, V) == ptr);
    CHECK(og_graph_generate_rmat(9, 4, 0.57, 0.19, 0.19, 3, 0, &p) == OG_OK);
    CHECK(og_graph_generate_er(3000, 6, 5, &er) == OG_OK);
    CHECK(og_graph_to_pull(er, &pl) == OG_OK);
    const std::string fp = dir + "/pl.prcsr";
    CHECK(og_graph_save(pl, fp.c_str()) == OG_OK);
    og_graph *pl2 = nullptr;
    CHECK(og_graph_load(fp.c_str(), OG_KIND_PULL, &pl2) == OG_OK);
    og_graph *pw = nullptr;
    CHECK(og_graph_add_weights(p, 9, &pw) == OG_OK);
    uint32_t big = 0;
    CHECK(og_graph_find_big_node(w, 0, &big) == OG_OK);

    // truncated and corrupt files
    {
        FILE *f = std::fopen(fw.c_str(), "rb");
        std::vector<unsigned char> buf(1 << 20);
        const size_t n = std::fread(buf.data(), 1, buf.size(), f);
        std::fclose(f);
        const std::string ft = dir + "/trunc.bwcsr";
        f = std::fopen(ft.c_str(), "wb");
        std::fwrite(buf.data(), 1, n / 2, f);
        std::fclose(f);
        og_graph *x = nullptr;
        CHECK(og_graph_load(ft.c_str(), OG_KIND_WEIGHTED, &x) != OG_OK && x == nullptr);
        CHECK(std::strlen(og_last_error()) > 0);
        buf[0] = 0xFF;  // V far too large
        f = std::fopen(ft.c_str(), "wb");
        std::fwrite(buf.data(), 1, n, f);
        std::fclose(f);
        CHECK(og_graph_load(ft.c_str(), OG_KIND_WEIGHTED, &x) != OG_OK);
        CHECK(og_graph_load((dir + "/missing").c_str(), OG_KIND_PLAIN, &x) == OG_EIO);
    }

    // from_csr edge cases
    {
        og_graph *x = nullptr;
        const uint64_t p1[1] = {0};
        CHECK(og_graph_from_csr(1, 0, p1, nullptr, nullptr, nullptr, OG_KIND_PLAIN, &x) == OG_OK);
        CHECK(og_graph_free(x) == OG_OK);
        const uint64_t pb[3] = {0, 5, 2};
        const uint32_t cb[4] = {1, 2, 0, 1};
        CHECK(og_graph_from_csr(3, 4, pb, cb, nullptr, nullptr, OG_KIND_PLAIN, &x) == OG_EFORMAT);
        const uint64_t pc[3] = {0, 2, 3};
        const uint32_t cc[4] = {1, 2, 7, 1};  // id 7 >= V
        CHECK(og_graph_from_csr(3, 4, pc, cc, nullptr, nullptr, OG_KIND_PLAIN, &x) == OG_EFORMAT);
    }

    // edge lists: arrays, text, MatrixMarket, a bad line
    {
        const uint32_t s[6] = {0, 1, 1, 2, 2, 3}, d[6] = {1, 2, 2, 2, 0, 0}, ww[6] = {5, 3, 1, 9, 2, 4};
        og_graph *x = nullptr;
        CHECK(og_graph_from_edges(0, 6, s, d, ww, OG_EDGES_SYMMETRIZE, &x) == OG_OK);
        CHECK(og_graph_get_info(x, &gi) == OG_OK && gi.num_vertices == 4);
        og_graph_free(x);
        const std::string ft = dir + "/e.txt", fm = dir + "/e.mtx", fb = dir + "/bad.txt";
        FILE *f = std::fopen(ft.c_str(), "w");
        std::fputs("# comment\n0 1 3\n1 2 4\n2,0,1\n% other\n3\t1\t2\n", f);
        std::fclose(f);
        CHECK(og_graph_load_edgelist(ft.c_str(), OG_EDGES_WEIGHTED, &x) == OG_OK);
        og_graph_free(x);
        f = std::fopen(fm.c_str(), "w");
        std::fputs("%%MatrixMarket matrix coordinate pattern symmetric\n4 4 3\n2 1\n3 2\n4 1\n", f);
        std::fclose(f);
        CHECK(og_graph_load_edgelist(fm.c_str(), 0, &x) == OG_OK);
        og_graph_free(x);
        f = std::fopen(fb.c_str(), "w");
        std::fputs("0 1\n1 x\n", f);
        std::fclose(f);
        CHECK(og_graph_load_edgelist(fb.c_str(), 0, &x) == OG_EFORMAT);
    }

    // planner: budgets from nothing to everything, advisek 0 (auto) .. 1
    for (uint64_t b : {(uint64_t)0, (uint64_t)1, E * 2, E * 8 / 4, E * 8, E * 64})
        for (float k : {0.0f, 0.25f, 0.8f, 1.0f}) {
            og_plan_params pp{};
            pp.edge_budget_bytes = b;
            pp.advisek = k;
            pp.param_size = 17;
            og_plan_info info{};
            const og_status st = og_plan_compute(V, E, ptr.data(), 8, &pp, &info);
            CHECK(st == OG_OK || st == OG_EBUDGET);
            if (st == OG_OK) CHECK(info.static_edges + info.staging_edges <= info.total_budget_edges + 1);
        }

    // drivers: graph loading and their error path without a device
    {
        const std::string fplain = dir + "/p.bcsr";
        CHECK(og_graph_save(p, fplain.c_str()) == OG_OK);
        onegraph::Options opt;
        opt.budget_fraction = 0.25;
        CHECK(onegraph::bfs_opt(fplain, 0, 0.8f, opt) == OG_EHIP);
        CHECK(onegraph::sssp_opt(fw, 0, 0.8, opt) == OG_EHIP);
        CHECK(onegraph::cc_opt(fplain, 0.8f, opt) == OG_EHIP);
        CHECK(onegraph::pr_opt(fp, 0.8, opt) == OG_EHIP);
        CHECK(onegraph::bfs_opt(dir + "/missing", 0, 0.8f, opt) == OG_EIO);
        // the result checker: equal, different and unreadable files
        std::vector<uint32_t> vals = {1, 2, 0xFFFFFFFFu, 4};
        const std::string fr = dir + "/r.bin";
        FILE *f = std::fopen(fr.c_str(), "wb");
        std::fwrite(vals.data(), 4, vals.size(), f);
        std::fclose(f);
        CHECK(onegraph::check_result(fr, vals.data(), vals.size(), 4, false, 0) == OG_OK);
        vals[1] = 3;
        CHECK(onegraph::check_result(fr, vals.data(), vals.size(), 4, false, 0) == OG_EFORMAT);
        CHECK(onegraph::check_result(dir + "/missing", vals.data(), vals.size(), 4, false, 0) == OG_EIO);
    }

    for (og_graph *g : {w, w2, p, er, pl, pl2, pw}) CHECK(og_graph_free(g) == OG_OK);
    CHECK(og_graph_free(nullptr) == OG_OK);
    std::printf("host checks: %d failures\n", failures);
    return failures ? 1 : 0;
}
