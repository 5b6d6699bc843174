// Edge-list conversion: text ("u v [w]" per line, '#' / '%' comment lines),
// Matrix Market (1-based, size line skipped) and the reference's binary
// ".toB" pairs {u32 src, u32 dst} into a host CSR that asc_write_graph
// stores as .bcsr / .bwcsr / .bcsc.
//
// Reference: converter/converter.cpp:22-420 (convertTxtToByte, then
// convertByteToBCSR / BCSC / BWCSR).  Its defects are not reproduced: V is
// one more than the largest id on either side (the reference takes the
// largest source only, converter.cpp:126) or the count the caller gives, so
// trailing isolated vertices survive; offsets of trailing zero-degree
// vertices are exact (converter.cpp:137-138); generated weights lie in
// [wmin, wmax] with wmin >= 1 by default (the reference draws rand() % 20,
// zero included, converter.cpp:308, 396) and are equal in both directions.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstring>

#include "gen_sampler.hpp"
#include "internal.hpp"

namespace asc {
namespace {

struct Parsed {
    std::vector<uint32_t> u, v, w;   // w empty when no line had a weight
    int weight_lines = 0, plain_lines = 0;
    bool bad = false;
};

inline bool is_sep(char c) { return c == ' ' || c == '\t' || c == ',' || c == '\r'; }

// Parses the lines that start inside [b, e) (a line starting before b is
// its predecessor's); fields past the third are ignored.
void parse_range(const char* text, size_t n, size_t b, size_t e, Parsed& out) {
    size_t i = b;
    if (b > 0 && text[b - 1] != '\n') {
        while (i < n && text[i] != '\n') i++;
        i++;
    }
    while (i < e && i < n) {
        size_t j = i;
        while (j < n && text[j] != '\n') j++;
        size_t k = i;
        while (k < j && is_sep(text[k])) k++;
        if (k < j && text[k] != '#' && text[k] != '%') {
            uint64_t f[3];
            int nf = 0;
            while (k < j && nf < 3) {
                if (text[k] < '0' || text[k] > '9') {
                    out.bad = true;
                    break;
                }
                uint64_t x = 0;
                int digits = 0;
                while (k < j && text[k] >= '0' && text[k] <= '9') {
                    x = x * 10 + (uint64_t)(text[k++] - '0');
                    digits++;
                }
                if (digits > 10) x = ~0ull;   // beyond 32 bits (no 64-bit wrap)
                f[nf++] = x;
                while (k < j && is_sep(text[k])) k++;
            }
            if (nf < 2 || f[0] > 0xFFFFFFFFull || f[1] > 0xFFFFFFFFull || (nf == 3 && f[2] > 0xFFFFFFFFull)) {
                out.bad = true;
            } else {
                out.u.push_back((uint32_t)f[0]);
                out.v.push_back((uint32_t)f[1]);
                if (nf == 3) {
                    out.w.resize(out.u.size() - 1, 0);
                    out.w.push_back((uint32_t)f[2]);
                    out.weight_lines++;
                } else {
                    out.plain_lines++;
                }
            }
        }
        i = j + 1;
    }
}

struct MappedFile {
    const char* p = nullptr;
    size_t n = 0;
    int fd = -1;
    ~MappedFile() {
        if (p && n) munmap((void*)p, n);
        if (fd >= 0) close(fd);
    }
    int open_ro(const char* path) {
        fd = ::open(path, O_RDONLY);
        if (fd < 0) {
            set_last_error("cannot open %s", path);
            return ASC_ERR_IO;
        }
        struct stat st;
        if (fstat(fd, &st) != 0) return ASC_ERR_IO;
        n = (size_t)st.st_size;
        if (n == 0) return ASC_OK;
        void* m = mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
        if (m == MAP_FAILED) {
            set_last_error("cannot map %s", path);
            return ASC_ERR_IO;
        }
        p = (const char*)m;
        return ASC_OK;
    }
};

void scan_u64(ThreadPool& pool, const uint32_t* cnt, uint64_t* off, uint64_t n) {
    const int nt = pool.size();
    std::vector<uint64_t> part(nt + 1, 0);
    const uint64_t chunk = (n + nt - 1) / nt;
    pool.run([&](int t, int) {
        uint64_t b = std::min<uint64_t>(n, t * chunk), e = std::min<uint64_t>(n, b + chunk), s = 0;
        for (uint64_t i = b; i < e; i++) s += cnt[i];
        part[t + 1] = s;
    });
    for (int t = 0; t < nt; t++) part[t + 1] += part[t];
    pool.run([&](int t, int) {
        uint64_t b = std::min<uint64_t>(n, t * chunk), e = std::min<uint64_t>(n, b + chunk), s = part[t];
        for (uint64_t i = b; i < e; i++) {
            off[i] = s;
            s += cnt[i];
        }
    });
    off[n] = part[nt];
}

// CSR of the (src, dst, w) triples grouped by `key_side` (0: src, the
// out-CSR; 1: dst, the in-CSR listing sources), adjacency sorted by the
// other endpoint, then weight; with dedup only the first (lightest) copy of
// each pair stays.
int group_csr(ThreadPool& pool, uint64_t V, const std::vector<uint32_t>& su, const std::vector<uint32_t>& sv,
              const std::vector<uint32_t>* sw, bool sym, bool dedup, bool self_loops, int key_side,
              HostBuf<uint64_t>& off_out, HostBuf<uint64_t>& keys_out, uint64_t& E_out) {
    const uint64_t M = su.size();
    HostBuf<uint32_t> cnt;
    ASC_TRY(cnt.alloc(V + 1));
    std::memset(cnt.p, 0, 4 * (V + 1));
    auto each = [&](uint64_t i, auto&& cb) {
        const uint32_t a = su[i], b = sv[i], w = sw ? (*sw)[i] : 0u;
        if (a == b && !self_loops) return;
        cb(a, b, w);
        if (sym && a != b) cb(b, a, w);
    };
    pool.parallel_for(M, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++)
            each(i, [&](uint32_t x, uint32_t y, uint32_t) {
                __atomic_fetch_add(&cnt[key_side ? y : x], 1u, __ATOMIC_RELAXED);
            });
    });
    HostBuf<uint64_t> off0;
    ASC_TRY(off0.alloc(V + 1));
    scan_u64(pool, cnt.p, off0.p, V);
    const uint64_t E0 = off0[V];
    HostBuf<uint64_t> raw, cur;
    ASC_TRY(raw.alloc(E0 + 1));
    ASC_TRY(cur.alloc(V + 1));
    std::memcpy(cur.p, off0.p, 8 * (V + 1));
    pool.parallel_for(M, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++)
            each(i, [&](uint32_t x, uint32_t y, uint32_t w) {
                const uint32_t k = key_side ? y : x, other = key_side ? x : y;
                const uint64_t pos = __atomic_fetch_add(&cur[k], 1ull, __ATOMIC_RELAXED);
                raw[pos] = ((uint64_t)other << 32) | w;
            });
    });
    cur.release();
    std::atomic<uint64_t> next{0};
    pool.run([&](int, int) {
        for (;;) {
            const uint64_t b = next.fetch_add(4096);
            if (b >= V) break;
            for (uint64_t x = b; x < std::min(V, b + 4096); x++) {
                uint64_t* s = raw.p + off0[x];
                uint64_t* t = raw.p + off0[x + 1];
                std::sort(s, t);
                if (dedup) {
                    uint64_t* o = s;
                    for (uint64_t* q = s; q < t; q++)
                        if (o == s || (o[-1] >> 32) != (*q >> 32)) *o++ = *q;
                    cnt[x] = (uint32_t)(o - s);
                } else {
                    cnt[x] = (uint32_t)(t - s);
                }
            }
        }
    });
    ASC_TRY(off_out.alloc(V + 1));
    scan_u64(pool, cnt.p, off_out.p, V);
    E_out = off_out[V];
    ASC_TRY(keys_out.alloc(E_out + 1));
    pool.parallel_for(V, [&](uint64_t b, uint64_t e) {
        for (uint64_t x = b; x < e; x++) std::memcpy(keys_out.p + off_out[x], raw.p + off0[x], 8 * (size_t)cnt[x]);
    });
    return ASC_OK;
}

}  // namespace

int read_edge_list(const char* path, const asc_edge_list_opts* o, HostCSR& out) {
    if (!path || !o) return ASC_ERR_INVALID;
    if (o->weighted && o->csc) return ASC_ERR_INVALID;
    if (o->weighted && (o->wmax < o->wmin)) return ASC_ERR_INVALID;
    ThreadPool local(o->threads > 0 ? o->threads : 1);
    ThreadPool& pool = o->threads > 0 ? local : host_pool();
    MappedFile f;
    ASC_TRY(f.open_ro(path));
    std::vector<uint32_t> su, sv, sw;
    bool one_based = o->one_based != 0;
    const size_t plen = std::strlen(path);
    if (plen >= 4 && std::strcmp(path + plen - 4, ".toB") == 0) {
        // binary pairs {u32 src, u32 dst}
        if (f.n % 8) {
            set_last_error("%s: %zu bytes is not a whole number of u32 pairs", path, f.n);
            return ASC_ERR_FORMAT;
        }
        const uint64_t M = f.n / 8;
        su.resize(M);
        sv.resize(M);
        const uint32_t* w = (const uint32_t*)f.p;
        pool.parallel_for(M, [&](uint64_t b, uint64_t e) {
            for (uint64_t i = b; i < e; i++) {
                su[i] = w[2 * i];
                sv[i] = w[2 * i + 1];
            }
        });
    } else {
        const bool mtx = f.n >= 14 && std::memcmp(f.p, "%%MatrixMarket", 14) == 0;
        size_t data_start = 0;
        if (mtx) {
            // 1-based ids; the first line that is not a comment gives the sizes
            one_based = true;
            size_t i = 0;
            for (;;) {
                size_t j = i;
                while (j < f.n && f.p[j] != '\n') j++;
                size_t k = i;
                while (k < j && is_sep(f.p[k])) k++;
                i = j + 1;
                if ((k < j && f.p[k] != '%') || j >= f.n) break;
            }
            data_start = std::min(i, f.n);
        }
        const int nt = pool.size() * 4;
        std::vector<Parsed> parts(nt);
        const size_t chunk = (f.n + nt - 1) / nt;
        std::atomic<int> next{0};
        pool.run([&](int, int) {
            for (;;) {
                const int t = next.fetch_add(1);
                if (t >= nt) break;
                const size_t b = std::max(data_start, std::min(f.n, (size_t)t * chunk));
                const size_t e = std::min(f.n, (size_t)t * chunk + chunk);
                if (b < e) parse_range(f.p, f.n, b, e, parts[t]);
            }
        });
        int wl = 0, pl = 0;
        size_t M = 0;
        for (auto& p : parts) {
            if (p.bad) {
                set_last_error("%s: a line is not \"u v [w]\" with ids and weights below 2^32", path);
                return ASC_ERR_FORMAT;
            }
            wl += p.weight_lines;
            pl += p.plain_lines;
            M += p.u.size();
        }
        if (wl && pl) {
            set_last_error("%s: some lines carry a weight and some do not", path);
            return ASC_ERR_FORMAT;
        }
        su.reserve(M);
        sv.reserve(M);
        if (wl) sw.reserve(M);
        for (auto& p : parts) {
            su.insert(su.end(), p.u.begin(), p.u.end());
            sv.insert(sv.end(), p.v.begin(), p.v.end());
            if (wl) sw.insert(sw.end(), p.w.begin(), p.w.end());
            std::vector<uint32_t>().swap(p.u);
            std::vector<uint32_t>().swap(p.v);
            std::vector<uint32_t>().swap(p.w);
        }
    }
    const uint64_t M = su.size();
    // ids: 1-based input is shifted; V from the ids unless given
    std::atomic<bool> zero_id{false};
    std::atomic<uint64_t> maxid{0};
    pool.parallel_for(M, [&](uint64_t b, uint64_t e) {
        uint64_t mx = 0;
        bool z = false;
        for (uint64_t i = b; i < e; i++) {
            if (one_based) {
                if (su[i] == 0 || sv[i] == 0) z = true;
                else su[i]--, sv[i]--;
            }
            mx = std::max<uint64_t>(mx, std::max(su[i], sv[i]));
        }
        if (z) zero_id = true;
        uint64_t cur = maxid.load();
        while (mx > cur && !maxid.compare_exchange_weak(cur, mx)) {
        }
    });
    if (zero_id) {
        set_last_error("%s: id 0 in a 1-based edge list", path);
        return ASC_ERR_FORMAT;
    }
    uint64_t V = M ? maxid.load() + 1 : 0;
    if (o->num_vertices) {
        if (o->num_vertices < V) {
            set_last_error("%s: id %llu >= the %llu vertices given", path, (unsigned long long)(V - 1),
                           (unsigned long long)o->num_vertices);
            return ASC_ERR_RANGE;
        }
        V = o->num_vertices;
    }
    if (V >= 0xFFFFFFFFull) return ASC_ERR_RANGE;
    const bool sym = o->symmetrize != 0, dedup = o->dedup != 0, loops = o->keep_self_loops != 0;
    HostBuf<uint64_t> off, keys;
    uint64_t E = 0;
    out.V = V;
    if (o->csc) {
        // in-CSR (sources of each vertex) plus out-degrees for PageRank
        ASC_TRY(group_csr(pool, V, su, sv, nullptr, sym, dedup, loops, 1, off, keys, E));
        out.kind = ASC_KIND_CSC;
        ASC_TRY(out.outdeg.alloc(V + 1));
        std::memset(out.outdeg.p, 0, 4 * (V + 1));
        ASC_TRY(out.edges.alloc(E + 1));
        pool.parallel_for(E, [&](uint64_t b, uint64_t e) {
            for (uint64_t k = b; k < e; k++) {
                out.edges[k] = (uint32_t)(keys[k] >> 32);
                __atomic_fetch_add(&out.outdeg[out.edges[k]], 1u, __ATOMIC_RELAXED);
            }
        });
    } else {
        ASC_TRY(group_csr(pool, V, su, sv, sw.empty() ? nullptr : &sw, sym, dedup, loops, 0, off, keys, E));
        if (o->weighted) {
            out.kind = ASC_KIND_WCSR;
            ASC_TRY(out.edges.alloc(2 * E + 1));
            const bool given = !sw.empty();
            const uint32_t span = o->wmax - o->wmin + 1;
            const uint64_t wkey = weight_key(o->seed);
            pool.parallel_for(V, [&](uint64_t b, uint64_t e) {
                for (uint64_t x = b; x < e; x++)
                    for (uint64_t k = off[x]; k < off[x + 1]; k++) {
                        const uint32_t d = (uint32_t)(keys[k] >> 32);
                        out.edges[2 * k] = d;
                        out.edges[2 * k + 1] = given ? (uint32_t)keys[k] : pair_weight(wkey, (uint32_t)x, d, o->wmin, span);
                    }
            });
        } else {
            out.kind = ASC_KIND_CSR;
            ASC_TRY(out.edges.alloc(E + 1));
            pool.parallel_for(E, [&](uint64_t b, uint64_t e) {
                for (uint64_t k = b; k < e; k++) out.edges[k] = (uint32_t)(keys[k] >> 32);
            });
        }
    }
    out.E = E;
    out.off = std::move(off);
    out.offset_bytes = 8;
    return ASC_OK;
}

}  // namespace asc

extern "C" {

void asc_edge_list_opts_init(asc_edge_list_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->dedup = 1;
    o->wmin = 1;
    o->wmax = 64;
    o->seed = 1;
}

int asc_read_edge_list(const char* path, const asc_edge_list_opts* o, asc_host_csr* out) {
    if (!path || !o || !out) return ASC_ERR_INVALID;
    std::memset(out, 0, sizeof(*out));
    asc::HostCSR h;
    ASC_TRY(asc::read_edge_list(path, o, h));
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

}  // extern "C"
