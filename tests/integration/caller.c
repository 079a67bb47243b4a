/* A plain C caller of the C ABI (include/onegraph.h), built by
 * tests/test_integration.py with gcc against libonegraph.so:
 *
 *   caller plan FILE BUDGET_BYTES ADVISEK   -- og_graph_load + og_plan_compute
 *                                             (host only), prints the plan
 *   caller sssp FILE SOURCE OUT            -- og_prepare (Ascetic, 25 % budget,
 *                                             advisek 0.8) + og_sssp, writes the
 *                                             distances to OUT (needs a GPU)
 *
 * It replaces sssp_opt (SSSP/oneGraph/CalculateOpt.h:12) the way a C program
 * would.  Exit status: 0 on success, else the og_status of the failing call. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "onegraph.h"

static int fail(const char *what, og_status st) {
    fprintf(stderr, "%s: status %d: %s\n", what, (int)st, og_last_error());
    return (int)st;
}

int main(int argc, char **argv) {
    if (argc < 5) {
        fprintf(stderr, "usage: caller plan FILE BUDGET ADVISEK | caller sssp FILE SOURCE OUT\n");
        return 100;
    }
    og_graph *g = NULL;
    og_status st = og_graph_load(argv[2], OG_KIND_WEIGHTED, &g);
    if (st != OG_OK) return fail("og_graph_load", st);
    og_graph_info gi;
    og_graph_get_info(g, &gi);
    int rc = 0;
    if (strcmp(argv[1], "plan") == 0) {
        uint64_t *ptr = (uint64_t *)malloc(gi.num_vertices * sizeof(uint64_t));
        og_graph_export(g, ptr, NULL, NULL, NULL);
        og_plan_params p = {.edge_budget_bytes = strtoull(argv[3], NULL, 10), .advisek = (float)atof(argv[4])};
        og_plan_info info;
        st = og_plan_compute(gi.num_vertices, gi.num_edges, ptr, 8, &p, &info);
        if (st != OG_OK) {
            rc = fail("og_plan_compute", st);
        } else {
            printf("total %llu static %llu staging %llu vsplit %llu\n", (unsigned long long)info.total_budget_edges,
                   (unsigned long long)info.static_edges, (unsigned long long)info.staging_edges,
                   (unsigned long long)info.vsplit);
        }
        free(ptr);
    } else {
        og_plan_params p = {.edge_budget_bytes = gi.num_edges * 8 / 4, .advisek = 0.8f};
        st = og_prepare(g, OG_ASCETIC, &p, NULL);
        if (st != OG_OK) {
            rc = fail("og_prepare", st);
        } else {
            uint32_t *dist = (uint32_t *)malloc(gi.num_vertices * sizeof(uint32_t));
            og_run_stats rs;
            st = og_sssp(g, (uint32_t)strtoul(argv[3], NULL, 10), dist, &rs);
            if (st != OG_OK) {
                rc = fail("og_sssp", st);
            } else {
                FILE *f = fopen(argv[4], "wb");
                fwrite(dist, sizeof(uint32_t), gi.num_vertices, f);
                fclose(f);
                printf("reached %llu\n", (unsigned long long)rs.reached);
            }
            free(dist);
        }
    }
    og_graph_free(g);
    return rc;
}
