// Internal declarations shared by the host (C++) and device (HIP) sources of
// libonegraph.  Nothing here crosses the C ABI.
#pragma once

#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <functional>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/onegraph.h"

namespace og {

// ----------------------------------------------------------------- errors

// Records a message for og_last_error() and returns the status unchanged.
og_status fail(og_status st, const char *fmt, ...);
void clear_error();

struct Error {  // thrown inside the library only, converted at the ABI
    og_status st;
    std::string msg;
};
[[noreturn]] void raise(og_status st, const char *fmt, ...);

// Runs fn and converts every escape (Error, bad_alloc, anything) to a status.
og_status guard(const std::function<og_status()> &fn);

// ------------------------------------------------------------ host memory

// Pinned (page-locked, device-mapped) host memory when a HIP device exists,
// plain malloc otherwise (pure-CPU use: generators, planner, tests).
void *host_alloc(size_t bytes, bool *pinned);
void host_free(void *p, bool pinned);
bool hip_available();

// ------------------------------------------------------------ thread pool

// Runs fn(begin, end, worker) over [0, n) in chunks of `grain`, on up to
// `threads` threads (0 = hardware concurrency).  Dynamic chunk scheduling.
void parallel_for(uint64_t n, uint64_t grain, unsigned threads,
                  const std::function<void(uint64_t, uint64_t, unsigned)> &fn);
unsigned default_threads();

// Persistent worker pool used by the dynamic-region gather, so that the hot
// loop does not create threads per chunk.
class WorkerPool {
  public:
    explicit WorkerPool(unsigned n);
    ~WorkerPool();
    unsigned size() const { return (unsigned)workers_.size() + 1; }
    // Runs fn(t) for t in [0, size()) with the caller as worker 0; returns
    // when all are done.
    void run(const std::function<void(unsigned)> &fn);

  private:
    void loop(unsigned id);
    std::vector<std::thread *> workers_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(unsigned)> *job_ = nullptr;
    uint64_t generation_ = 0;
    unsigned pending_ = 0;
    bool stop_ = false;
};

// ---------------------------------------------------------------- RNG

// SplitMix64 finaliser (Steele, Lea, Flood 2014), used as a counter-based
// generator: draw(key, i) = mix64(key + (i + 1) * golden).
static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static inline uint64_t rng_key(uint64_t seed) { return mix64(seed ^ 0x6A09E667F3BCC909ull); }
static inline uint64_t rng_draw(uint64_t key, uint64_t ctr) {
    return mix64(key + (ctr + 1) * 0x9E3779B97F4A7C15ull);
}

// ------------------------------------------------------------ device side

struct DeviceState;   // defined in engine.hip
void device_state_free(DeviceState *d);

}  // namespace og

// The graph object behind the opaque og_graph handle.
struct og_graph {
    uint64_t V = 0, E = 0;
    og_graph_kind kind = OG_KIND_PLAIN;
    uint64_t *ptr = nullptr;        // V + 1 entries (sentinel ptr[V] = E)
    uint32_t *cols = nullptr;       // E (plain / pull); also the "to" ids of weighted
    uint32_t *wedges = nullptr;     // 2E, AoS {to, w} (weighted only)
    uint32_t *out_degree = nullptr; // V (pull only)
    uint32_t max_degree = 0;
    bool pinned_ptr = false, pinned_cols = false, pinned_wedges = false, pinned_outdeg = false;
    og::DeviceState *dev = nullptr;
    std::atomic<int> busy{0};

    uint32_t edge_bytes() const { return kind == OG_KIND_WEIGHTED ? 8u : 4u; }
    uint32_t degree(uint64_t v) const { return (uint32_t)(ptr[v + 1] - ptr[v]); }
    const void *edge_data() const {
        return kind == OG_KIND_WEIGHTED ? (const void *)wedges : (const void *)cols;
    }
};

namespace og {
// Allocates the arrays of an empty graph (host side).
void graph_alloc(og_graph *g, uint64_t V, uint64_t E, og_graph_kind kind);
void graph_release_host(og_graph *g);
void graph_finish(og_graph *g);  // computes max_degree
}  // namespace og
