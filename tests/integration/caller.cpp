// A C++ caller of the host layer (include/onegraph_host.hpp), built by
// tests/test_integration.py with g++ against libonegraph.so.  The call is
// the reference BFS main's (BFS/oneGraph_BFS/main.cpp:3-14, bfs_opt of
// CalculateOpt.h:9) with an explicit device edge budget:
//
//   caller_cpp FILE SOURCE ADVISEK OUT
//
// Exit status: 0 on success, else the og_status of the failing call.
#include <cstdio>
#include <cstdlib>

#include "onegraph_host.hpp"

int main(int argc, char **argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: caller_cpp FILE SOURCE ADVISEK OUT\n");
        return 100;
    }
    onegraph::Options opt;
    opt.budget_fraction = 0.25;  // explicit device edge budget: a quarter of the CSR edge bytes
    opt.out_path = argv[4];
    const og_status st = onegraph::bfs_opt(argv[1], (uint32_t)std::strtoul(argv[2], nullptr, 10),
                                           (float)std::atof(argv[3]), opt);
    if (st != OG_OK) std::fprintf(stderr, "%s\n", og_last_error());
    return (int)st;
}
