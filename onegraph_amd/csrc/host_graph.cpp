// Host side of libonegraph: error plumbing, pinned memory, thread pool,
// reference-format CSR I/O, synthetic generators and the static-region
// planner.  Device work lives in engine.hip / pagerank.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cctype>
#include <cstring>
#include <string>
#include <vector>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <memory>

#include "og_internal.h"

namespace og {

// ================================================================ errors

static thread_local std::string t_last_error;

static std::string vformat(const char *fmt, va_list ap) {
    char buf[1024];
    vsnprintf(buf, sizeof buf, fmt, ap);
    return std::string(buf);
}

og_status fail(og_status st, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    t_last_error = vformat(fmt, ap);
    va_end(ap);
    return st;
}

void clear_error() { t_last_error.clear(); }

void raise(og_status st, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::string m = vformat(fmt, ap);
    va_end(ap);
    throw Error{st, m};
}

og_status guard(const std::function<og_status()> &fn) {
    try {
        return fn();
    } catch (const Error &e) {
        return fail(e.st, "%s", e.msg.c_str());
    } catch (const std::bad_alloc &) {
        return fail(OG_ENOMEM_HOST, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(OG_EINVAL, "internal error: %s", e.what());
    } catch (...) {
        return fail(OG_EINVAL, "internal error");
    }
}

// ============================================================ host memory

bool hip_available() {
    static int cached = -1;
    if (cached < 0) {
        int n = 0;
        cached = (hipGetDeviceCount(&n) == hipSuccess && n > 0) ? 1 : 0;
        (void)hipGetLastError();
    }
    return cached == 1;
}

void *host_alloc(size_t bytes, bool *pinned) {
    if (bytes == 0) bytes = 8;
    if (hip_available()) {
        void *p = nullptr;
        if (hipHostMalloc(&p, bytes, hipHostMallocDefault) == hipSuccess && p) {
            *pinned = true;
            return p;
        }
        (void)hipGetLastError();
    }
    void *p = std::malloc(bytes);
    if (!p) raise(OG_ENOMEM_HOST, "host allocation of %zu bytes failed", bytes);
    *pinned = false;
    return p;
}

void host_free(void *p, bool pinned) {
    if (!p) return;
    if (pinned)
        (void)hipHostFree(p);
    else
        std::free(p);
}

// ============================================================ thread pool

unsigned default_threads() {
    unsigned n = std::thread::hardware_concurrency();
    if (const char *e = std::getenv("OMP_NUM_THREADS")) {
        int v = std::atoi(e);
        if (v > 0) n = (unsigned)v;
    }
    if (n == 0) n = 1;
    return std::min(n, 64u);
}

void parallel_for(uint64_t n, uint64_t grain, unsigned threads,
                  const std::function<void(uint64_t, uint64_t, unsigned)> &fn) {
    if (n == 0) return;
    if (grain == 0) grain = 1;
    if (threads == 0) threads = default_threads();
    uint64_t chunks = (n + grain - 1) / grain;
    if (threads > chunks) threads = (unsigned)chunks;
    if (threads <= 1) {
        fn(0, n, 0);
        return;
    }
    std::atomic<uint64_t> next{0};
    std::exception_ptr err = nullptr;
    std::mutex err_mu;
    auto body = [&](unsigned t) {
        try {
            for (;;) {
                uint64_t c = next.fetch_add(1, std::memory_order_relaxed);
                if (c >= chunks) break;
                uint64_t b = c * grain, e = std::min(n, b + grain);
                fn(b, e, t);
            }
        } catch (...) {
            std::lock_guard<std::mutex> lk(err_mu);
            if (!err) err = std::current_exception();
        }
    };
    std::vector<std::thread> ts;
    ts.reserve(threads - 1);
    for (unsigned t = 1; t < threads; ++t) ts.emplace_back(body, t);
    body(0);
    for (auto &t : ts) t.join();
    if (err) std::rethrow_exception(err);
}

WorkerPool::WorkerPool(unsigned n) {
    if (n == 0) n = default_threads();
    for (unsigned i = 1; i < n; ++i) workers_.push_back(new std::thread(&WorkerPool::loop, this, i));
}

WorkerPool::~WorkerPool() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    for (auto *t : workers_) {
        t->join();
        delete t;
    }
}

void WorkerPool::loop(unsigned id) {
    uint64_t seen = 0;
    for (;;) {
        const std::function<void(unsigned)> *job;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return stop_ || generation_ != seen; });
            if (stop_) return;
            seen = generation_;
            job = job_;
        }
        (*job)(id);
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (--pending_ == 0) done_cv_.notify_one();
        }
    }
}

void WorkerPool::run(const std::function<void(unsigned)> &fn) {
    if (workers_.empty()) {
        fn(0);
        return;
    }
    {
        std::lock_guard<std::mutex> lk(mu_);
        job_ = &fn;
        pending_ = (unsigned)workers_.size();
        ++generation_;
    }
    cv_.notify_all();
    fn(0);
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [&] { return pending_ == 0; });
}

// ================================================================ graphs

void graph_alloc(og_graph *g, uint64_t V, uint64_t E, og_graph_kind kind) {
    g->V = V;
    g->E = E;
    g->kind = kind;
    g->ptr = (uint64_t *)host_alloc((V + 1) * sizeof(uint64_t), &g->pinned_ptr);
    if (kind == OG_KIND_WEIGHTED)
        g->wedges = (uint32_t *)host_alloc(2 * E * sizeof(uint32_t), &g->pinned_wedges);
    else
        g->cols = (uint32_t *)host_alloc(E * sizeof(uint32_t), &g->pinned_cols);
    if (kind == OG_KIND_PULL)
        g->out_degree = (uint32_t *)host_alloc(V * sizeof(uint32_t), &g->pinned_outdeg);
}

void graph_release_host(og_graph *g) {
    host_free(g->ptr, g->pinned_ptr);
    host_free(g->cols, g->pinned_cols);
    host_free(g->wedges, g->pinned_wedges);
    host_free(g->out_degree, g->pinned_outdeg);
    g->ptr = nullptr;
    g->cols = g->wedges = g->out_degree = nullptr;
}

void graph_finish(og_graph *g) {
    uint32_t md = 0;
    for (uint64_t v = 0; v < g->V; ++v) md = std::max<uint32_t>(md, (uint32_t)(g->ptr[v + 1] - g->ptr[v]));
    g->max_degree = md;
}

// Checks offsets (monotone, within E) and column ids (< V) in parallel.
static void validate(const og_graph *g) {
    if (g->V == 0) raise(OG_EFORMAT, "graph has no vertices");
    if (g->V > 0xFFFFFFFEull) raise(OG_EFORMAT, "V=%llu exceeds the u32 vertex-id range", (unsigned long long)g->V);
    for (uint64_t v = 0; v < g->V; ++v)
        if (g->ptr[v] > g->ptr[v + 1])
            raise(OG_EFORMAT, "row offsets decrease at vertex %llu", (unsigned long long)v);
    if (g->ptr[g->V] != g->E) raise(OG_EFORMAT, "last row overruns the edge array");
    std::atomic<uint64_t> bad{UINT64_MAX};
    const uint64_t V = g->V;
    parallel_for(g->E, 1 << 22, 0, [&](uint64_t b, uint64_t e, unsigned) {
        if (g->kind == OG_KIND_WEIGHTED) {
            for (uint64_t i = b; i < e; ++i)
                if (g->wedges[2 * i] >= V) { bad.store(i); return; }
        } else {
            for (uint64_t i = b; i < e; ++i)
                if (g->cols[i] >= V) { bad.store(i); return; }
        }
    });
    if (bad.load() != UINT64_MAX)
        raise(OG_EFORMAT, "edge %llu names a vertex >= V", (unsigned long long)bad.load());
}

}  // namespace og

using namespace og;

// ============================================================ C ABI: misc

extern "C" const char *og_last_error(void) { return t_last_error.c_str(); }

extern "C" const char *og_version(void) { return "onegraph-mi355x 0.1 (gfx950)"; }

extern "C" og_status og_device_count(int *count) {
    if (!count) return fail(OG_EINVAL, "count is NULL");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        n = 0;
    }
    *count = n;
    return OG_OK;
}

// ============================================================ C ABI: I/O

namespace {

struct File {
    FILE *f = nullptr;
    ~File() {
        if (f) fclose(f);
    }
};

void read_exact(FILE *f, void *dst, uint64_t bytes, const char *what) {
    uint8_t *p = (uint8_t *)dst;
    while (bytes) {
        size_t step = (size_t)std::min<uint64_t>(bytes, 1ull << 28);
        size_t got = fread(p, 1, step, f);
        if (got != step) raise(OG_EIO, "file truncated while reading %s", what);
        p += step;
        bytes -= step;
    }
}

void write_exact(FILE *f, const void *src, uint64_t bytes) {
    const uint8_t *p = (const uint8_t *)src;
    while (bytes) {
        size_t step = (size_t)std::min<uint64_t>(bytes, 1ull << 28);
        if (fwrite(p, 1, step, f) != step) raise(OG_EIO, "write failed");
        p += step;
        bytes -= step;
    }
}

}  // namespace

// Byte layout follows GraphMeta::readDataFromFile (BFS/oneGraph_BFS/
// GraphMeta.h:125-145): u64 V, u64 E, [u32 outDegree[V] for PR], u64 ptr[V],
// Edge[E].  The reference reads the 8-byte V into a 4-byte field (:130); here
// V is read as u64 and range-checked.
extern "C" og_status og_graph_load(const char *path, og_graph_kind kind, og_graph **out) {
    clear_error();
    if (!path || !out) return fail(OG_EINVAL, "path/out is NULL");
    if (kind != OG_KIND_PLAIN && kind != OG_KIND_WEIGHTED && kind != OG_KIND_PULL)
        return fail(OG_EINVAL, "unknown graph kind %d", (int)kind);
    *out = nullptr;
    return guard([&]() -> og_status {
        File file;
        file.f = fopen(path, "rb");
        if (!file.f) raise(OG_EIO, "cannot open %s", path);
        fseeko(file.f, 0, SEEK_END);
        uint64_t size = (uint64_t)ftello(file.f);
        fseeko(file.f, 0, SEEK_SET);
        uint64_t hdr[2];
        if (size < 16) raise(OG_EIO, "file %s shorter than the 16-byte header", path);
        read_exact(file.f, hdr, 16, "header");
        uint64_t V = hdr[0], E = hdr[1];
        if (V == 0 || V > 0xFFFFFFFEull) raise(OG_EFORMAT, "header vertex count %llu out of range", (unsigned long long)V);
        uint32_t eb = kind == OG_KIND_WEIGHTED ? 8 : 4;
        if (E > (size / eb) + 1) raise(OG_EIO, "header edge count %llu larger than the file", (unsigned long long)E);
        uint64_t need = 16 + (kind == OG_KIND_PULL ? 4 * V : 0) + 8 * V + E * eb;
        if (size < need) raise(OG_EIO, "file truncated: %llu bytes, layout needs %llu", (unsigned long long)size, (unsigned long long)need);
        if (size > need) raise(OG_EFORMAT, "file has %llu trailing bytes", (unsigned long long)(size - need));
        std::unique_ptr<og_graph> g(new og_graph);
        graph_alloc(g.get(), V, E, kind);
        try {
            if (kind == OG_KIND_PULL) read_exact(file.f, g->out_degree, 4 * V, "outDegree");
            read_exact(file.f, g->ptr, 8 * V, "row offsets");
            g->ptr[V] = E;
            if (kind == OG_KIND_WEIGHTED)
                read_exact(file.f, g->wedges, 8 * E, "edges");
            else
                read_exact(file.f, g->cols, 4 * E, "edges");
            validate(g.get());
            graph_finish(g.get());
        } catch (...) {
            graph_release_host(g.get());
            throw;
        }
        *out = g.release();
        return OG_OK;
    });
}

extern "C" og_status og_graph_save(const og_graph *g, const char *path) {
    clear_error();
    if (!g || !path) return fail(OG_EINVAL, "graph/path is NULL");
    return guard([&]() -> og_status {
        File file;
        file.f = fopen(path, "wb");
        if (!file.f) raise(OG_EIO, "cannot create %s", path);
        uint64_t hdr[2] = {g->V, g->E};
        write_exact(file.f, hdr, 16);
        if (g->kind == OG_KIND_PULL) write_exact(file.f, g->out_degree, 4 * g->V);
        write_exact(file.f, g->ptr, 8 * g->V);
        if (g->kind == OG_KIND_WEIGHTED)
            write_exact(file.f, g->wedges, 8 * g->E);
        else
            write_exact(file.f, g->cols, 4 * g->E);
        if (fflush(file.f) != 0) raise(OG_EIO, "flush failed for %s", path);
        return OG_OK;
    });
}

extern "C" og_status og_graph_from_csr(uint64_t V, uint64_t E, const uint64_t *ptr, const uint32_t *cols,
                                       const uint32_t *weights, const uint32_t *out_degree,
                                       og_graph_kind kind, og_graph **out) {
    clear_error();
    if (!out || !ptr || (E && !cols)) return fail(OG_EINVAL, "null array");
    if (kind == OG_KIND_WEIGHTED && E && !weights) return fail(OG_EINVAL, "weighted graph needs weights");
    if (kind == OG_KIND_PULL && !out_degree) return fail(OG_EINVAL, "pull graph needs out_degree");
    if (kind != OG_KIND_PLAIN && kind != OG_KIND_WEIGHTED && kind != OG_KIND_PULL)
        return fail(OG_EINVAL, "unknown graph kind %d", (int)kind);
    if (V == 0 || V > 0xFFFFFFFEull) return fail(OG_EFORMAT, "vertex count out of range");
    *out = nullptr;
    return guard([&]() -> og_status {
        std::unique_ptr<og_graph> g(new og_graph);
        graph_alloc(g.get(), V, E, kind);
        try {
            std::memcpy(g->ptr, ptr, 8 * V);
            g->ptr[V] = E;
            if (kind == OG_KIND_WEIGHTED) {
                parallel_for(E, 1 << 22, 0, [&](uint64_t b, uint64_t e, unsigned) {
                    for (uint64_t i = b; i < e; ++i) {
                        g->wedges[2 * i] = cols[i];
                        g->wedges[2 * i + 1] = weights[i];
                    }
                });
            } else if (E) {
                std::memcpy(g->cols, cols, 4 * E);
            }
            if (kind == OG_KIND_PULL) std::memcpy(g->out_degree, out_degree, 4 * V);
            validate(g.get());
            graph_finish(g.get());
        } catch (...) {
            graph_release_host(g.get());
            throw;
        }
        *out = g.release();
        return OG_OK;
    });
}

extern "C" og_status og_graph_get_info(const og_graph *g, og_graph_info *info) {
    if (!g || !info) return fail(OG_EINVAL, "graph/info is NULL");
    info->num_vertices = g->V;
    info->num_edges = g->E;
    info->kind = (uint32_t)g->kind;
    info->max_degree = g->max_degree;
    return OG_OK;
}

extern "C" og_status og_graph_export(const og_graph *g, uint64_t *ptr, uint32_t *cols, uint32_t *weights,
                                     uint32_t *out_degree) {
    if (!g) return fail(OG_EINVAL, "graph is NULL");
    if (ptr) std::memcpy(ptr, g->ptr, 8 * g->V);
    if (g->kind == OG_KIND_WEIGHTED) {
        if (cols || weights) {
            for (uint64_t i = 0; i < g->E; ++i) {
                if (cols) cols[i] = g->wedges[2 * i];
                if (weights) weights[i] = g->wedges[2 * i + 1];
            }
        }
    } else {
        if (cols && g->E) std::memcpy(cols, g->cols, 4 * g->E);
        if (weights) return fail(OG_EUNSUPPORTED, "graph has no weights");
    }
    if (out_degree) {
        if (g->kind != OG_KIND_PULL) return fail(OG_EUNSUPPORTED, "graph has no outDegree array");
        std::memcpy(out_degree, g->out_degree, 4 * g->V);
    }
    return OG_OK;
}

// findbignode (BFS/oneGraph_BFS/GraphMeta.h:149-165): scans i in
// (last, V-1) for degree > 3000, returning `last` when none is found.
extern "C" og_status og_graph_find_big_node(const og_graph *g, uint32_t last, uint32_t *out) {
    if (!g || !out) return fail(OG_EINVAL, "graph/out is NULL");
    for (uint64_t i = (uint64_t)last + 1; i + 1 < g->V; ++i) {
        if (g->ptr[i + 1] - g->ptr[i] > 3000) {
            *out = (uint32_t)i;
            return OG_OK;
        }
    }
    *out = last;
    return OG_OK;
}

extern "C" og_status og_graph_free(og_graph *g) {
    clear_error();
    if (!g) return OG_OK;
    // a graph reserved by a running call or by an og_server keeps its
    // device buffers: freeing them under live kernels would be a device
    // use-after-free, so the call is refused and the graph left alone
    int expect = 0;
    if (!g->busy.compare_exchange_strong(expect, 1))
        return fail(OG_EBUSY, "graph is in use (a running call or an og_server holds it); destroy the server first");
    if (g->dev) device_state_free(g->dev);
    g->dev = nullptr;
    graph_release_host(g);
    delete g;
    return OG_OK;
}

// ======================================================= CSR construction

namespace {

// Builds a deduplicated, row-sorted CSR from M candidate arcs produced by
// gen(i, &u, &v) (false = drop).  symmetric adds v->u for each u->v.
// Two generation passes (count, scatter) keep memory at one edge array.
template <class Gen>
void build_csr(og_graph *g, uint64_t V, uint64_t M, bool symmetric, Gen gen) {
    const unsigned T = default_threads();
    std::vector<uint32_t> cnt(V + 1, 0);
    uint32_t *c = cnt.data();
    const uint64_t grain = 1 << 20;
    parallel_for(M, grain, T, [&](uint64_t b, uint64_t e, unsigned) {
        for (uint64_t i = b; i < e; ++i) {
            uint32_t u, v;
            if (!gen(i, u, v)) continue;
            __atomic_fetch_add(&c[u], 1u, __ATOMIC_RELAXED);
            if (symmetric) __atomic_fetch_add(&c[v], 1u, __ATOMIC_RELAXED);
        }
    });
    std::vector<uint64_t> off(V + 1);
    uint64_t s = 0;
    for (uint64_t v = 0; v < V; ++v) {
        off[v] = s;
        s += cnt[v];
    }
    off[V] = s;
    std::vector<uint64_t> cur(off.begin(), off.end() - 1);
    uint64_t *cu = cur.data();
    std::vector<uint32_t> tmp(s);
    uint32_t *t = tmp.data();
    parallel_for(M, grain, T, [&](uint64_t b, uint64_t e, unsigned) {
        for (uint64_t i = b; i < e; ++i) {
            uint32_t u, v;
            if (!gen(i, u, v)) continue;
            t[__atomic_fetch_add(&cu[u], 1ull, __ATOMIC_RELAXED)] = v;
            if (symmetric) t[__atomic_fetch_add(&cu[v], 1ull, __ATOMIC_RELAXED)] = u;
        }
    });
    cur.clear();
    cur.shrink_to_fit();
    // Sort + unique every row; cnt becomes the deduplicated degree.
    parallel_for(V, 4096, T, [&](uint64_t b, uint64_t e, unsigned) {
        for (uint64_t v = b; v < e; ++v) {
            uint32_t *r0 = t + off[v], *r1 = t + off[v + 1];
            std::sort(r0, r1);
            c[v] = (uint32_t)(std::unique(r0, r1) - r0);
        }
    });
    uint64_t E = 0;
    for (uint64_t v = 0; v < V; ++v) E += cnt[v];
    graph_alloc(g, V, E, OG_KIND_PLAIN);
    uint64_t acc = 0;
    for (uint64_t v = 0; v < V; ++v) {
        g->ptr[v] = acc;
        acc += cnt[v];
    }
    g->ptr[V] = acc;
    parallel_for(V, 16384, T, [&](uint64_t b, uint64_t e, unsigned) {
        for (uint64_t v = b; v < e; ++v)
            if (cnt[v]) std::memcpy(g->cols + g->ptr[v], t + off[v], 4ull * cnt[v]);
    });
    graph_finish(g);
}

inline uint32_t weight_of(uint64_t wkey, uint64_t i) {
    return 1u + (uint32_t)((((rng_draw(wkey, i) >> 32) * 63ull)) >> 32);
}

}  // namespace

// R-MAT (Chakrabarti, Zhan, Faloutsos 2004) in the Graph500 quadrant order:
// a = (src bit 0, dst bit 0), b = (0, 1), c = (1, 0), d = (1, 1), most
// significant bit first, one 32-bit uniform per level (two levels per
// 64-bit draw), no noise, no vertex permutation.
extern "C" og_status og_graph_generate_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                                            uint64_t seed, uint64_t weight_seed, og_graph **out) {
    clear_error();
    if (!out) return fail(OG_EINVAL, "out is NULL");
    if (scale == 0 || scale > 31) return fail(OG_EINVAL, "scale must be in [1, 31]");
    if (edge_factor == 0) return fail(OG_EINVAL, "edge_factor must be > 0");
    if (!(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0)) return fail(OG_EINVAL, "bad R-MAT probabilities");
    *out = nullptr;
    return guard([&]() -> og_status {
        const uint64_t V = 1ull << scale;
        const uint64_t M = (uint64_t)edge_factor << scale;
        const uint64_t key = rng_key(seed);
        const uint32_t TA = (uint32_t)std::floor(a * 4294967296.0);
        const uint32_t TAB = (uint32_t)std::floor((a + b) * 4294967296.0);
        const double abc = a + b + c;
        const uint64_t TABC64 = (uint64_t)std::floor(abc * 4294967296.0);
        const uint32_t W = (scale + 1) / 2;
        auto gen = [&](uint64_t i, uint32_t &u, uint32_t &v) -> bool {
            uint32_t s = 0, d = 0;
            for (uint32_t l = 0; l < scale; l += 2) {
                uint64_t r = rng_draw(key, i * W + l / 2);
                for (uint32_t h = 0; h < 2 && l + h < scale; ++h) {
                    uint32_t x = (uint32_t)(r >> (32 * h));
                    uint32_t sb, db;
                    if (x < TA) { sb = 0; db = 0; }
                    else if (x < TAB) { sb = 0; db = 1; }
                    else if ((uint64_t)x < TABC64) { sb = 1; db = 0; }
                    else { sb = 1; db = 1; }
                    s = (s << 1) | sb;
                    d = (d << 1) | db;
                }
            }
            u = s;
            v = d;
            return u != v;
        };
        std::unique_ptr<og_graph> g(new og_graph);
        build_csr(g.get(), V, M, false, gen);
        if (weight_seed) {
            og_graph *w = nullptr;
            og_status st = og_graph_add_weights(g.get(), weight_seed, &w);
            graph_release_host(g.get());
            if (st != OG_OK) return st;
            *out = w;
            return OG_OK;
        }
        *out = g.release();
        return OG_OK;
    });
}

// G(n, m) with m = n*avg/2 candidate pairs (u, v) drawn uniformly; self-pairs
// and repeats dropped; every pair stored in both directions.
extern "C" og_status og_graph_generate_er(uint64_t n, uint32_t avg_degree, uint64_t seed, og_graph **out) {
    clear_error();
    if (!out) return fail(OG_EINVAL, "out is NULL");
    if (n < 2 || n > 0xFFFFFFFEull) return fail(OG_EINVAL, "vertex count out of range");
    if (avg_degree == 0) return fail(OG_EINVAL, "avg_degree must be > 0");
    *out = nullptr;
    return guard([&]() -> og_status {
        const uint64_t M = n * avg_degree / 2;
        const uint64_t key = rng_key(seed);
        auto gen = [&](uint64_t i, uint32_t &u, uint32_t &v) -> bool {
            u = (uint32_t)(((unsigned __int128)rng_draw(key, 2 * i) * n) >> 64);
            v = (uint32_t)(((unsigned __int128)rng_draw(key, 2 * i + 1) * n) >> 64);
            return u != v;
        };
        std::unique_ptr<og_graph> g(new og_graph);
        build_csr(g.get(), n, M, true, gen);
        *out = g.release();
        return OG_OK;
    });
}

// ===================================================== edge-list conversion

namespace {

// Weighted counterpart of build_csr: arcs carry (v << 32 | w) so that a row
// sort orders by target, then weight, and a unique-by-target keeps the
// smallest weight of repeated arcs.
template <class Gen>
void build_wcsr(og_graph *g, uint64_t V, uint64_t M, bool symmetric, Gen gen) {
    const unsigned T = default_threads();
    std::vector<uint32_t> cnt(V + 1, 0);
    uint32_t *c = cnt.data();
    const uint64_t grain = 1 << 20;
    parallel_for(M, grain, T, [&](uint64_t b, uint64_t e, unsigned) {
        for (uint64_t i = b; i < e; ++i) {
            uint32_t u, v, w;
            if (!gen(i, u, v, w)) continue;
            __atomic_fetch_add(&c[u], 1u, __ATOMIC_RELAXED);
            if (symmetric && u != v) __atomic_fetch_add(&c[v], 1u, __ATOMIC_RELAXED);
        }
    });
    std::vector<uint64_t> off(V + 1);
    uint64_t s = 0;
    for (uint64_t v = 0; v < V; ++v) {
        off[v] = s;
        s += cnt[v];
    }
    off[V] = s;
    std::vector<uint64_t> cur(off.begin(), off.end() - 1);
    uint64_t *cu = cur.data();
    std::vector<uint64_t> tmp(s);
    uint64_t *t = tmp.data();
    parallel_for(M, grain, T, [&](uint64_t b, uint64_t e, unsigned) {
        for (uint64_t i = b; i < e; ++i) {
            uint32_t u, v, w;
            if (!gen(i, u, v, w)) continue;
            t[__atomic_fetch_add(&cu[u], 1ull, __ATOMIC_RELAXED)] = ((uint64_t)v << 32) | w;
            if (symmetric && u != v) t[__atomic_fetch_add(&cu[v], 1ull, __ATOMIC_RELAXED)] = ((uint64_t)u << 32) | w;
        }
    });
    cur.clear();
    cur.shrink_to_fit();
    parallel_for(V, 4096, T, [&](uint64_t b, uint64_t e, unsigned) {
        for (uint64_t v = b; v < e; ++v) {
            uint64_t *r0 = t + off[v], *r1 = t + off[v + 1];
            std::sort(r0, r1);
            c[v] = (uint32_t)(std::unique(r0, r1, [](uint64_t x, uint64_t y) { return (x >> 32) == (y >> 32); }) - r0);
        }
    });
    uint64_t E = 0;
    for (uint64_t v = 0; v < V; ++v) E += cnt[v];
    graph_alloc(g, V, E, OG_KIND_WEIGHTED);
    uint64_t acc = 0;
    for (uint64_t v = 0; v < V; ++v) {
        g->ptr[v] = acc;
        acc += cnt[v];
    }
    g->ptr[V] = acc;
    parallel_for(V, 16384, T, [&](uint64_t b, uint64_t e, unsigned) {
        for (uint64_t v = b; v < e; ++v)
            for (uint32_t k = 0; k < cnt[v]; ++k) {
                const uint64_t x = t[off[v] + k];
                g->wedges[2 * (g->ptr[v] + k)] = (uint32_t)(x >> 32);
                g->wedges[2 * (g->ptr[v] + k) + 1] = (uint32_t)x;
            }
    });
    graph_finish(g);
}

og_graph *edges_to_graph(uint64_t V, uint64_t M, const uint32_t *src, const uint32_t *dst, const uint32_t *wt,
                         uint32_t flags) {
    if (V == 0) {
        std::atomic<uint64_t> mx{0};
        parallel_for(M, 1 << 20, 0, [&](uint64_t b, uint64_t e, unsigned) {
            uint64_t m = 0;
            for (uint64_t i = b; i < e; ++i) m = std::max<uint64_t>(m, std::max(src[i], dst[i]));
            uint64_t c = mx.load();
            while (m > c && !mx.compare_exchange_weak(c, m)) {
            }
        });
        if (!M) raise(OG_EINVAL, "no arcs and no vertex count");
        V = mx.load() + 1;
    } else {
        std::atomic<uint64_t> bad{~0ull};
        parallel_for(M, 1 << 20, 0, [&](uint64_t b, uint64_t e, unsigned) {
            for (uint64_t i = b; i < e; ++i)
                if (src[i] >= V || dst[i] >= V) {
                    uint64_t c = bad.load();
                    while (i < c && !bad.compare_exchange_weak(c, i)) {
                    }
                    break;
                }
        });
        if (bad.load() != ~0ull)
            raise(OG_EINVAL, "arc %llu (%u -> %u) names a vertex >= %llu", (unsigned long long)bad.load(),
                  src[bad.load()], dst[bad.load()], (unsigned long long)V);
    }
    if (V > 0xFFFFFFFFull) raise(OG_EINVAL, "more than 2^32 - 1 vertices");
    const bool sym = flags & OG_EDGES_SYMMETRIZE, loops = flags & OG_EDGES_KEEP_SELF_LOOPS;
    std::unique_ptr<og_graph> g(new og_graph);
    if (wt) {
        build_wcsr(g.get(), V, M, sym, [&](uint64_t i, uint32_t &u, uint32_t &v, uint32_t &w) {
            u = src[i];
            v = dst[i];
            w = wt[i];
            return loops || u != v;
        });
    } else {
        build_csr(g.get(), V, M, sym, [&](uint64_t i, uint32_t &u, uint32_t &v) {
            u = src[i];
            v = dst[i];
            return loops || u != v;
        });
    }
    return g.release();
}

// Parses an unsigned decimal at p (after optional blanks); false if none.
inline bool parse_u64(const char *&p, const char *end, uint64_t &x) {
    while (p < end && (*p == ' ' || *p == '\t' || *p == ',')) ++p;
    if (p >= end || *p < '0' || *p > '9') return false;
    uint64_t v = 0;
    while (p < end && *p >= '0' && *p <= '9') {
        if (v > (~0ull - 9) / 10) return false;
        v = v * 10 + (uint64_t)(*p - '0');
        ++p;
    }
    if (p < end && (*p == '.' || *p == 'e' || *p == 'E')) return false;  // not an integer
    x = v;
    return true;
}

}  // namespace

extern "C" og_status og_graph_from_edges(uint64_t num_vertices, uint64_t num_edges, const uint32_t *src,
                                         const uint32_t *dst, const uint32_t *weights, uint32_t flags,
                                         og_graph **out) {
    clear_error();
    if (!out) return fail(OG_EINVAL, "out is NULL");
    if (num_edges && (!src || !dst)) return fail(OG_EINVAL, "src/dst is NULL");
    *out = nullptr;
    return guard([&]() -> og_status {
        *out = edges_to_graph(num_vertices, num_edges, src, dst, weights, flags);
        return OG_OK;
    });
}

extern "C" og_status og_graph_load_edgelist(const char *path, uint32_t flags, og_graph **out) {
    clear_error();
    if (!path || !out) return fail(OG_EINVAL, "path/out is NULL");
    *out = nullptr;
    return guard([&]() -> og_status {
        FILE *f = fopen(path, "rb");
        if (!f) raise(OG_EIO, "cannot open %s", path);
        std::unique_ptr<FILE, int (*)(FILE *)> fc(f, fclose);
        if (fseeko(f, 0, SEEK_END) != 0) raise(OG_EIO, "cannot seek %s", path);
        const uint64_t size = (uint64_t)ftello(f);
        fseeko(f, 0, SEEK_SET);
        std::vector<char> buf(size + 1);
        if (size && fread(buf.data(), 1, size, f) != size) raise(OG_EIO, "short read on %s", path);
        buf[size] = '\n';
        const char *base = buf.data(), *end = base + size;
        auto line_of = [&](const char *at) { return 1 + (uint64_t)std::count(base, at, '\n'); };
        auto next_line = [&](const char *p) {
            const char *q = (const char *)memchr(p, '\n', (size_t)(end - p));
            return q ? q + 1 : end;
        };
        // MatrixMarket: banner, comments, then "rows cols entries".
        uint64_t V = 0;
        const char *data = base;
        bool weighted = flags & OG_EDGES_WEIGHTED;
        if (size >= 14 && !memcmp(base, "%%MatrixMarket", 14)) {
            const char *eol = next_line(base);
            std::string banner(base, eol);
            for (auto &ch : banner) ch = (char)tolower((unsigned char)ch);
            if (banner.find("coordinate") == std::string::npos)
                raise(OG_EFORMAT, "%s: only MatrixMarket coordinate files hold a graph", path);
            if (banner.find("symmetric") != std::string::npos) flags |= OG_EDGES_SYMMETRIZE;
            if (banner.find("pattern") != std::string::npos && weighted)
                raise(OG_EFORMAT, "%s: a pattern MatrixMarket file has no weights", path);
            flags |= OG_EDGES_ONE_BASED;
            const char *p = eol;
            while (p < end && (*p == '%' || *p == '\n' || *p == '\r')) p = next_line(p);
            uint64_t rows = 0, cols = 0, nnz = 0;
            if (!parse_u64(p, end, rows) || !parse_u64(p, end, cols) || !parse_u64(p, end, nnz))
                raise(OG_EFORMAT, "%s:%llu: bad MatrixMarket size line", path, (unsigned long long)line_of(p));
            V = std::max(rows, cols);
            data = next_line(p);
        }
        const uint64_t one = (flags & OG_EDGES_ONE_BASED) ? 1 : 0;
        // Line-aligned chunks, one per thread; each parses into its own arrays.
        const unsigned T = default_threads();
        std::vector<const char *> cut(T + 1);
        cut[0] = data;
        for (unsigned t = 1; t < T; ++t) {
            const char *p = data + (uint64_t)(end - data) * t / T;
            cut[t] = std::max(cut[t - 1], p > data ? next_line(p - 1) : data);
        }
        cut[T] = end;
        struct Part {
            std::vector<uint32_t> u, v, w;
            const char *err = nullptr;
            const char *why = "";
        };
        std::vector<Part> part(T);
        parallel_for(T, 1, T, [&](uint64_t b, uint64_t e, unsigned) {
            for (uint64_t t = b; t < e; ++t) {
                Part &pt = part[t];
                for (const char *p = cut[t]; p < cut[t + 1];) {
                    const char *eol = next_line(p);
                    const char *q = p;
                    while (q < eol && (*q == ' ' || *q == '\t')) ++q;
                    if (q >= eol || *q == '\n' || *q == '\r' || *q == '#' || *q == '%') {
                        p = eol;
                        continue;
                    }
                    uint64_t a = 0, c = 0, w = 1;
                    if (!parse_u64(q, eol, a) || !parse_u64(q, eol, c)) {
                        pt.err = p;
                        pt.why = "expected two vertex ids";
                        break;
                    }
                    if (weighted && !parse_u64(q, eol, w)) {
                        pt.err = p;
                        pt.why = "expected an integer weight";
                        break;
                    }
                    if (a < one || c < one || a - one >= 0xFFFFFFFFull || c - one >= 0xFFFFFFFFull ||
                        w > 0xFFFFFFFFull) {
                        pt.err = p;
                        pt.why = "vertex id or weight out of range";
                        break;
                    }
                    pt.u.push_back((uint32_t)(a - one));
                    pt.v.push_back((uint32_t)(c - one));
                    if (weighted) pt.w.push_back((uint32_t)w);
                    p = eol;
                }
            }
        });
        for (auto &pt : part)
            if (pt.err) raise(OG_EFORMAT, "%s:%llu: %s", path, (unsigned long long)line_of(pt.err), pt.why);
        uint64_t M = 0;
        for (auto &pt : part) M += pt.u.size();
        std::vector<uint32_t> su(M), sv(M), sw(weighted ? M : 0);
        uint64_t at = 0;
        for (auto &pt : part) {
            std::copy(pt.u.begin(), pt.u.end(), su.begin() + at);
            std::copy(pt.v.begin(), pt.v.end(), sv.begin() + at);
            if (weighted) std::copy(pt.w.begin(), pt.w.end(), sw.begin() + at);
            at += pt.u.size();
            std::vector<uint32_t>().swap(pt.u);
            std::vector<uint32_t>().swap(pt.v);
            std::vector<uint32_t>().swap(pt.w);
        }
        *out = edges_to_graph(V, M, su.data(), sv.data(), weighted ? sw.data() : nullptr, flags);
        return OG_OK;
    });
}

// Weight of the edge at CSR position i: 1 + floor(hi32(draw(i)) * 63 / 2^32).
extern "C" og_status og_graph_add_weights(const og_graph *src, uint64_t weight_seed, og_graph **out) {
    clear_error();
    if (!src || !out) return fail(OG_EINVAL, "graph/out is NULL");
    if (src->kind != OG_KIND_PLAIN) return fail(OG_EUNSUPPORTED, "weights can be added to a plain graph only");
    *out = nullptr;
    return guard([&]() -> og_status {
        std::unique_ptr<og_graph> g(new og_graph);
        graph_alloc(g.get(), src->V, src->E, OG_KIND_WEIGHTED);
        std::memcpy(g->ptr, src->ptr, 8 * (src->V + 1));
        const uint64_t wkey = rng_key(weight_seed);
        parallel_for(src->E, 1 << 20, 0, [&](uint64_t b, uint64_t e, unsigned) {
            for (uint64_t i = b; i < e; ++i) {
                g->wedges[2 * i] = src->cols[i];
                g->wedges[2 * i + 1] = weight_of(wkey, i);
            }
        });
        g->max_degree = src->max_degree;
        *out = g.release();
        return OG_OK;
    });
}

// Pull layout for PageRank (the input read by PR/cpu/Graph.h): outDegree of
// every vertex, then the in-edge CSR with in-lists sorted by source id.
extern "C" og_status og_graph_to_pull(const og_graph *src, og_graph **out) {
    clear_error();
    if (!src || !out) return fail(OG_EINVAL, "graph/out is NULL");
    if (src->kind == OG_KIND_PULL) return fail(OG_EUNSUPPORTED, "graph is already in pull layout");
    *out = nullptr;
    return guard([&]() -> og_status {
        const uint64_t V = src->V, E = src->E;
        const bool w = src->kind == OG_KIND_WEIGHTED;
        auto dst_of = [&](uint64_t i) { return w ? src->wedges[2 * i] : src->cols[i]; };
        std::unique_ptr<og_graph> g(new og_graph);
        graph_alloc(g.get(), V, E, OG_KIND_PULL);
        std::vector<uint32_t> indeg(V, 0);
        uint32_t *ind = indeg.data();
        parallel_for(E, 1 << 20, 0, [&](uint64_t b, uint64_t e, unsigned) {
            for (uint64_t i = b; i < e; ++i) __atomic_fetch_add(&ind[dst_of(i)], 1u, __ATOMIC_RELAXED);
        });
        uint64_t s = 0;
        for (uint64_t v = 0; v < V; ++v) {
            g->ptr[v] = s;
            s += indeg[v];
            g->out_degree[v] = (uint32_t)(src->ptr[v + 1] - src->ptr[v]);
        }
        g->ptr[V] = s;
        std::vector<uint64_t> cur(g->ptr, g->ptr + V);
        uint64_t *cu = cur.data();
        parallel_for(V, 4096, 0, [&](uint64_t b, uint64_t e, unsigned) {
            for (uint64_t u = b; u < e; ++u)
                for (uint64_t i = src->ptr[u]; i < src->ptr[u + 1]; ++i)
                    g->cols[__atomic_fetch_add(&cu[dst_of(i)], 1ull, __ATOMIC_RELAXED)] = (uint32_t)u;
        });
        parallel_for(V, 4096, 0, [&](uint64_t b, uint64_t e, unsigned) {
            for (uint64_t v = b; v < e; ++v) std::sort(g->cols + g->ptr[v], g->cols + g->ptr[v + 1]);
        });
        graph_finish(g.get());
        *out = g.release();
        return OG_OK;
    });
}

// ================================================================ planner

// Restates GraphMeta::getMaxPartitionSize (BFS/oneGraph_BFS/GraphMeta.h:
// 299-360) with the 12 GiB constant (:305) replaced by an explicit budget:
//   total      = budget / sizeof(Edge)                           (:309)
//   adviseK    = (10 - E/total)/9, <0 -> 0.5, >1 -> 1.0; user > 0 wins (:311-324)
//   static     = (float)adviseK * total, clamped to E             (:326-329)
//   isInStatic = ptr[i] < static && ptr[i] + deg[i] - 1 < static  (:345-353)
//   staging    = total - static                                   (:357)
//   overload   = E - sum(deg of static vertices)                  (:358)
extern "C" og_status og_plan_compute(uint64_t V, uint64_t E, const uint64_t *ptr, uint32_t edge_bytes,
                                     const og_plan_params *params, og_plan_info *info) {
    clear_error();
    if (!ptr || !params || !info) return fail(OG_EINVAL, "null argument");
    if (edge_bytes != 4 && edge_bytes != 8) return fail(OG_EINVAL, "edge_bytes must be 4 or 8");
    if (V == 0) return fail(OG_EINVAL, "graph has no vertices");
    std::memset(info, 0, sizeof *info);
    uint64_t total;
    if (params->edge_budget_bytes) {
        total = params->edge_budget_bytes / edge_bytes;
    } else {
        const uint64_t avail = 12ull << 30;
        const uint64_t reduce = (uint64_t)(params->param_size ? params->param_size : 17) * 4ull * V;
        if (reduce >= avail) return fail(OG_EBUDGET, "vertex arrays exceed the default 12 GiB budget");
        total = (avail - reduce) / edge_bytes;
    }
    if (total == 0) return fail(OG_EBUDGET, "edge budget holds no edge");
    float k = (10.0f - (float)E / (float)total) / 9.0f;
    if (k < 0) k = 0.5f;
    if (k > 1) k = 1.0f;
    if (params->advisek > 0) k = params->advisek;
    // The reference forms this product in float (GraphMeta.h:326), which drops
    // up to ulp(total) edges once total > 2^24; double keeps advisek = 1 with a
    // budget of E edges fully static.
    uint64_t stat = (uint64_t)((double)k * (double)total);
    if (stat > E) stat = E;
    if (stat > total) stat = total;
    // ptr has V entries in the reference layout; the last row ends at E
    // (GraphMeta.h:222).
    auto row_end = [&](uint64_t i) { return i + 1 < V ? ptr[i + 1] : E; };
    uint64_t nstatic = 0, edges_in_static = 0;
    for (uint64_t i = 0; i < V; ++i) {
        uint64_t deg = row_end(i) - ptr[i];
        if (ptr[i] < stat && ptr[i] + deg - 1 < stat) {
            ++nstatic;
            edges_in_static += deg;
        }
    }
    // vsplit: first vertex whose row ends beyond the static prefix.
    uint64_t lo = 0, hi = V;
    while (lo < hi) {
        uint64_t mid = (lo + hi) / 2;
        if (row_end(mid) <= stat) lo = mid + 1; else hi = mid;
    }
    info->total_budget_edges = total;
    info->static_edges = stat;
    info->staging_edges = total - stat;
    info->overload_edges = E - edges_in_static;
    info->static_vertices = nstatic;
    info->vsplit = lo;
    info->advisek_used = k;
    return OG_OK;
}
