/* A plain C caller of the C ABI (include/onegraph.h), built by
 * tests/test_integration.py with gcc against libonegraph.so:
 *
 *   caller plan FILE BUDGET_BYTES ADVISEK   -- og_graph_load + og_plan_compute
 *                                             (host only), prints the plan
 *   caller sssp FILE SOURCE OUT            -- og_prepare (Ascetic, 25 % budget,
 *                                             advisek 0.8) + og_sssp, writes the
 *                                             distances to OUT (needs a GPU)
 *   caller serve FILE SOURCE X             -- the graph whole on the device, one
 *                                             og_sssp and one og_bfs from SOURCE,
 *                                             then a server (two SSSP lanes) that
 *                                             answers the same two queries; prints
 *                                             "ok" when the checksums agree (needs a
 *                                             GPU)
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
        fprintf(stderr, "usage: caller plan FILE BUDGET ADVISEK | caller sssp FILE SOURCE OUT | caller serve FILE SOURCE X\n");
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
    } else if (strcmp(argv[1], "serve") == 0) {
        og_plan_params p = {.edge_budget_bytes = gi.num_edges * 8, .advisek = 1.0f};
        const uint32_t src = (uint32_t)strtoul(argv[3], NULL, 10);
        og_run_stats rs_sssp, rs_bfs;
        og_server *srv = NULL;
        if ((st = og_prepare(g, OG_ASCETIC, &p, NULL)) != OG_OK) {
            rc = fail("og_prepare", st);
        } else if ((st = og_sssp(g, src, NULL, &rs_sssp)) != OG_OK || (st = og_bfs(g, src, NULL, &rs_bfs)) != OG_OK) {
            rc = fail("og_sssp / og_bfs", st);
        } else {
            og_server_params sp = {.state_budget_bytes = 0, .sssp_lanes = 2, .sssp_running_max = 0};
            if ((st = og_server_create(&g, 1, &sp, &srv)) != OG_OK) return fail("og_server_create", st);
            uint64_t id_sssp = 0, id_bfs = 0;
            if ((st = og_server_submit(srv, 0, OG_QUERY_SSSP, src, 0, &id_sssp)) != OG_OK ||
                (st = og_server_submit(srv, 0, OG_QUERY_BFS, src, 0, &id_bfs)) != OG_OK) {
                rc = fail("og_server_submit", st);
            } else {
                og_query_result res[2];
                uint32_t got = 0, n = 0, ok = 0;
                while (got < 2 && (st = og_server_poll(srv, res + got, 2 - got, &n, 10000)) == OG_OK && n) got += n;
                for (uint32_t k = 0; k < got; ++k) {
                    if (res[k].status != OG_OK) continue;
                    if (res[k].id == id_sssp && res[k].kind == OG_QUERY_SSSP && res[k].path == OG_PATH_LANE &&
                        res[k].checksum == rs_sssp.checksum)
                        ok++;
                    if (res[k].id == id_bfs && res[k].kind == OG_QUERY_BFS && res[k].checksum == rs_bfs.checksum) ok++;
                }
                og_server_info info;
                og_server_get_info(srv, &info);
                printf("%s sssp_slots %u\n", ok == 2 ? "ok" : "mismatch", info.sssp_slots[0]);
                rc = ok == 2 ? 0 : 101;
            }
            og_server_destroy(srv);
        }
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
