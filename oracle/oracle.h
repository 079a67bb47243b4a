/* oracle.h -- CPU restatements used as the checker.  TEST INFRASTRUCTURE
 * ONLY: never linked into libonegraph.  See oracle.c for the reference
 * file:line each function follows and how it is pinned. */
#ifndef OG_ORACLE_H
#define OG_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

void og_oracle_bfs(uint64_t V, uint64_t E, const uint64_t *ptr, const uint32_t *cols, uint32_t src, uint32_t *value,
                   uint64_t *node_sum, uint32_t *iterations);
void og_oracle_sssp(uint64_t V, uint64_t E, const uint64_t *ptr, const uint32_t *to, const uint32_t *w, uint32_t src,
                    uint32_t *value);
uint64_t og_oracle_cc(uint64_t V, uint64_t E, const uint64_t *ptr, const uint32_t *cols, uint32_t *value);
void og_oracle_pagerank(uint64_t V, uint64_t E, const uint32_t *outdeg, const uint64_t *ptr, const uint32_t *in_cols,
                        double damping, double tol, uint32_t max_it, double *rank, uint32_t *iterations,
                        double *final_l1);
void og_oracle_pagerank_par(uint64_t V, uint64_t E, const uint32_t *outdeg, const uint64_t *ptr,
                            const uint32_t *in_cols, double damping, double tol, uint32_t max_it, double *rank,
                            uint32_t *iterations, double *final_l1, int threads);
/* Seconds spent in the compute loop of the last oracle call (this thread's
 * process; allocation and initialisation excluded). */
double og_oracle_loop_seconds(void);

#ifdef __cplusplus
}
#endif
#endif
