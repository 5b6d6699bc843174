// Deterministic synthetic graphs: R-MAT and Erdos-Renyi, built on the host.
//
// The samples come from the shared sampler (gen_sampler.hpp), so generation
// is reproducible regardless of the thread count and the sample list is
// regenerated instead of stored between the degree-count and the scatter
// pass.  Self-loops and duplicates are removed and every adjacency list is
// sorted.  gen_device.hip builds the same CSR on the GPU.
#include <algorithm>

#include "gen_sampler.hpp"
#include "internal.hpp"

namespace asc {

int make_sampler(const asc_gen_params* p, Sampler& s, uint64_t& M) {
    if (!p) return ASC_ERR_INVALID;
    s = Sampler{};
    s.type = p->type;
    s.key = mix64(p->seed ^ 0x5A17EC0DEull);
    if (p->type == 0) {
        if (p->scale < 1 || p->scale > 31) return ASC_ERR_RANGE;
        s.scale = p->scale;
        s.V = uint64_t(1) << p->scale;
        M = p->num_samples ? p->num_samples : (uint64_t)p->edge_factor * s.V;
        double a = p->a, b = p->b, c = p->c;
        if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return ASC_ERR_INVALID;
        auto thr = [](double x) -> uint32_t {
            double t = x * 4294967296.0;
            return t >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)t;
        };
        s.ta = thr(a);
        s.tab = thr(a + b);
        s.tabc = thr(a + b + c);
    } else if (p->type == 1) {
        if (p->num_vertices == 0 || p->num_vertices >= 0xFFFFFFFEull) return ASC_ERR_RANGE;
        s.V = p->num_vertices;
        M = p->num_samples;
    } else {
        return ASC_ERR_INVALID;
    }
    if (p->weighted && (p->wmin == 0 || p->wmax < p->wmin)) return ASC_ERR_INVALID;
    return ASC_OK;
}

namespace {

// Exclusive scan of cnt[0..n) into off[0..n], in parallel.
void parallel_scan(ThreadPool& pool, const uint32_t* cnt, uint64_t* off, uint64_t n) {
    int nt = pool.size();
    std::vector<uint64_t> part(nt + 1, 0);
    uint64_t chunk = (n + nt - 1) / nt;
    pool.run([&](int t, int) {
        uint64_t b = std::min(n, (uint64_t)t * chunk), e = std::min(n, b + chunk), s = 0;
        for (uint64_t i = b; i < e; i++) s += cnt[i];
        part[t + 1] = s;
    });
    for (int t = 0; t < nt; t++) part[t + 1] += part[t];
    pool.run([&](int t, int) {
        uint64_t b = std::min(n, (uint64_t)t * chunk), e = std::min(n, b + chunk), s = part[t];
        for (uint64_t i = b; i < e; i++) {
            off[i] = s;
            s += cnt[i];
        }
    });
    off[n] = part[nt];
}

// Builds a deduplicated sorted CSR from an edge source enumerated twice.
// emit(i, cb) calls cb(u, v) for each directed edge produced by sample i.
template <class Emit>
int build_csr(ThreadPool& pool, uint64_t V, uint64_t M, Emit emit, HostBuf<uint64_t>& off_out,
              HostBuf<uint32_t>& dst_out, uint64_t& E_out) {
    HostBuf<uint32_t> cnt;
    ASC_TRY(cnt.alloc(V));
    pool.parallel_for(V, [&](uint64_t b, uint64_t e) { std::memset(cnt.p + b, 0, 4 * (e - b)); });
    pool.parallel_for(M, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++)
            emit(i, [&](uint32_t u, uint32_t) { __atomic_fetch_add(&cnt[u], 1u, __ATOMIC_RELAXED); });
    });
    HostBuf<uint64_t> off0;
    ASC_TRY(off0.alloc(V + 1));
    parallel_scan(pool, cnt.p, off0.p, V);
    const uint64_t E0 = off0[V];
    HostBuf<uint32_t> raw;
    ASC_TRY(raw.alloc(E0 + 1));
    // cursor = off0 copy (as u64 in cnt-sized chunks: reuse a u64 array)
    HostBuf<uint64_t> cur;
    ASC_TRY(cur.alloc(V));
    pool.parallel_for(V, [&](uint64_t b, uint64_t e) { std::memcpy(cur.p + b, off0.p + b, 8 * (e - b)); });
    pool.parallel_for(M, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++)
            emit(i, [&](uint32_t u, uint32_t v) {
                uint64_t pos = __atomic_fetch_add(&cur[u], 1ull, __ATOMIC_RELAXED);
                raw[pos] = v;
            });
    });
    cur.release();
    // sort + unique each adjacency; new degree into cnt
    std::atomic<uint64_t> next{0};
    const uint64_t grain = 4096;
    pool.run([&](int, int) {
        for (;;) {
            uint64_t b = next.fetch_add(grain);
            if (b >= V) break;
            uint64_t e = std::min(V, b + grain);
            for (uint64_t u = b; u < e; u++) {
                uint32_t* s = raw.p + off0[u];
                uint32_t* t = raw.p + off0[u + 1];
                std::sort(s, t);
                cnt[u] = (uint32_t)(std::unique(s, t) - s);
            }
        }
    });
    ASC_TRY(off_out.alloc(V + 1));
    parallel_scan(pool, cnt.p, off_out.p, V);
    E_out = off_out[V];
    ASC_TRY(dst_out.alloc(E_out + 1));
    pool.parallel_for(V, [&](uint64_t b, uint64_t e) {
        for (uint64_t u = b; u < e; u++)
            std::memcpy(dst_out.p + off_out[u], raw.p + off0[u], 4 * (size_t)cnt[u]);
    });
    return ASC_OK;
}

}  // namespace

int generate_graph(const asc_gen_params* p, HostCSR& out) {
    Sampler s;
    uint64_t M = 0;
    ASC_TRY(make_sampler(p, s, M));
    const uint64_t V = s.V;
    ThreadPool local(p->threads > 0 ? p->threads : 1);
    ThreadPool& pool = p->threads > 0 ? local : host_pool();

    const bool sym = p->symmetrize != 0;
    auto emit = [&](uint64_t i, auto&& cb) {
        uint32_t u, v;
        s.sample(i, u, v);
        if (u == v) return;
        cb(u, v);
        if (sym) cb(v, u);
    };
    HostBuf<uint64_t> off;
    HostBuf<uint32_t> dst;
    uint64_t E = 0;
    ASC_TRY(build_csr(pool, V, M, emit, off, dst, E));

    out.V = V;
    out.E = E;
    if (p->csc) {
        // PageRank input: in-CSR + out-degrees.  For symmetric graphs the
        // in-CSR equals the out-CSR; otherwise transpose.
        out.kind = ASC_KIND_CSC;
        ASC_TRY(out.outdeg.alloc(V));
        pool.parallel_for(V, [&](uint64_t b, uint64_t e) {
            for (uint64_t u = b; u < e; u++) out.outdeg[u] = (uint32_t)(off[u + 1] - off[u]);
        });
        if (sym) {
            out.off = std::move(off);
            out.edges = std::move(dst);
        } else {
            auto emit_t = [&](uint64_t u, auto&& cb) {
                for (uint64_t k = off[u]; k < off[u + 1]; k++) cb(dst[k], (uint32_t)u);
            };
            HostBuf<uint64_t> ioff;
            HostBuf<uint32_t> isrc;
            uint64_t E2 = 0;
            ASC_TRY(build_csr(pool, V, V, emit_t, ioff, isrc, E2));
            out.off = std::move(ioff);
            out.edges = std::move(isrc);
        }
    } else if (p->weighted) {
        out.kind = ASC_KIND_WCSR;
        ASC_TRY(out.edges.alloc(2 * E + 1));
        const uint32_t span = p->wmax - p->wmin + 1;
        const uint64_t wkey = weight_key(p->seed);
        pool.parallel_for(V, [&](uint64_t b, uint64_t e) {
            for (uint64_t u = b; u < e; u++)
                for (uint64_t k = off[u]; k < off[u + 1]; k++) {
                    out.edges[2 * k] = dst[k];
                    out.edges[2 * k + 1] = pair_weight(wkey, (uint32_t)u, dst[k], p->wmin, span);
                }
        });
        out.off = std::move(off);
    } else {
        out.kind = ASC_KIND_CSR;
        out.off = std::move(off);
        out.edges = std::move(dst);
    }
    return ASC_OK;
}

}  // namespace asc

extern "C" {

int asc_generate(const asc_gen_params* p, asc_host_csr* out) {
    if (!p || !out) return ASC_ERR_INVALID;
    std::memset(out, 0, sizeof(*out));
    asc::HostCSR h;
    ASC_TRY(asc::generate_graph(p, h));
    out->num_vertices = h.V;
    out->num_edges = h.E;
    out->kind = h.kind;
    out->off = h.off.p;
    out->edges = h.edges.p;
    out->out_degree = h.outdeg.p;
    // ownership moves to the caller's struct
    h.off.p = nullptr;
    h.edges.p = nullptr;
    h.outdeg.p = nullptr;
    return ASC_OK;
}

void asc_host_csr_free(asc_host_csr* h) {
    if (!h) return;
    std::free(h->off);
    std::free(h->edges);
    std::free(h->out_degree);
    std::memset(h, 0, sizeof(*h));
}

void asc_host_csr_view(const asc_host_csr* h, asc_csr_view* v) {
    if (!h || !v) return;
    std::memset(v, 0, sizeof(*v));
    v->num_vertices = h->num_vertices;
    v->num_edges = h->num_edges;
    v->kind = h->kind;
    v->has_sentinel = 1;
    v->off64 = h->off;
    v->edges = h->edges;
    v->out_degree = h->out_degree;
}

}  // extern "C"
