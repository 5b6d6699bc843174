// Binary CSR / CSC readers and writers.
//
// Layouts (little-endian), as read by GraphMeta.cuh:151-171 and written by
// converter/converter.cpp:178-181, 296-300, 413-416:
//   BCSR      u64 V, u64 E, u64 off[V], u32 dst[E]
//   BWCSR     u64 V, u64 E, u64 off[V], {u32 dst, u32 w}[E]
//   BCSC      u64 V, u64 E, u32 outdeg[V], u64 inoff[V], u32 src[E]
//   LEGACY32  u32 V, u32 E, u32 off[V], u32 dst[E]          (bfs.cu:294-304)
//   LEGACY32W u32 V, u32 E, u32 off[V], {u32 dst, u32 w}[E] (sssp.cu:176-186)
// No layout stores the off[V] sentinel; the reader appends off[V] = E.
#include <sys/stat.h>

#include <algorithm>

#include "internal.hpp"

namespace asc {

namespace {

struct File {
    FILE* f = nullptr;
    ~File() {
        if (f) fclose(f);
    }
};

int read_exact(FILE* f, void* dst, size_t bytes) {
    char* p = (char*)dst;
    while (bytes) {
        size_t chunk = bytes < (size_t(1) << 30) ? bytes : (size_t(1) << 30);
        size_t got = fread(p, 1, chunk, f);
        if (got != chunk) return ASC_ERR_IO;
        p += got;
        bytes -= got;
    }
    return ASC_OK;
}

int write_exact(FILE* f, const void* src, size_t bytes) {
    const char* p = (const char*)src;
    while (bytes) {
        size_t chunk = bytes < (size_t(1) << 30) ? bytes : (size_t(1) << 30);
        if (fwrite(p, 1, chunk, f) != chunk) return ASC_ERR_IO;
        p += chunk;
        bytes -= chunk;
    }
    return ASC_OK;
}

// Expected file size of each layout for a given header.
uint64_t layout_size(int fmt, uint64_t V, uint64_t E) {
    switch (fmt) {
        case ASC_FMT_BCSR: return 16 + 8 * V + 4 * E;
        case ASC_FMT_BWCSR: return 16 + 8 * V + 8 * E;
        case ASC_FMT_BCSC: return 16 + 12 * V + 4 * E;
        case ASC_FMT_LEGACY32: return 8 + 4 * V + 4 * E;
        case ASC_FMT_LEGACY32W: return 8 + 4 * V + 8 * E;
    }
    return 0;
}

bool header_plausible(uint64_t V, uint64_t E, uint64_t size) {
    return V > 0 && V < (uint64_t(1) << 32) - 2 && E <= size;
}

}  // namespace

int validate_csr(const HostCSR& h) {
    if (h.V == 0 || h.V >= 0xFFFFFFFEull) {
        set_last_error("vertex count %llu out of range [1, 2^32-2)", (unsigned long long)h.V);
        return ASC_ERR_RANGE;
    }
    if (h.off[h.V] != h.E) {
        set_last_error("offset sentinel %llu != E %llu", (unsigned long long)h.off[h.V], (unsigned long long)h.E);
        return ASC_ERR_FORMAT;
    }
    std::atomic<int> bad{0};
    std::atomic<uint64_t> bad_at{0};
    const uint64_t V = h.V;
    host_pool().parallel_for(V, [&](uint64_t b, uint64_t e) {
        for (uint64_t v = b; v < e; v++)
            if (h.off[v] > h.off[v + 1]) {
                bad = 1;
                bad_at = v;
                return;
            }
    });
    if (bad) {
        set_last_error("offsets decrease at vertex %llu", (unsigned long long)bad_at.load());
        return ASC_ERR_FORMAT;
    }
    const int words = h.edge_words();
    host_pool().parallel_for(h.E, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++)
            if (h.edges[i * words] >= V) {
                bad = 2;
                bad_at = i;
                return;
            }
    });
    if (bad) {
        set_last_error("edge %llu has an id >= V", (unsigned long long)bad_at.load());
        return ASC_ERR_FORMAT;
    }
    return ASC_OK;
}

// Reads an open graph file as layout `fmt` (header already parsed).
static int read_graph_as(FILE* f, const char* path, uint64_t size, int fmt, uint64_t V64, uint64_t E64,
                         uint32_t V32, uint32_t E32, HostCSR& out);

int read_graph_file(const char* path, int format, HostCSR& out) {
    if (!path) return ASC_ERR_INVALID;
    struct stat sb;
    if (stat(path, &sb) != 0) {
        set_last_error("cannot stat %s", path);
        return ASC_ERR_IO;
    }
    const uint64_t size = (uint64_t)sb.st_size;
    File fh;
    fh.f = fopen(path, "rb");
    if (!fh.f) {
        set_last_error("cannot open %s", path);
        return ASC_ERR_IO;
    }
    unsigned char hdr[16] = {0};
    size_t got = fread(hdr, 1, 16, fh.f);
    if (got < 8) {
        set_last_error("%s: file shorter than any header", path);
        return ASC_ERR_FORMAT;
    }
    uint64_t V64 = 0, E64 = 0;
    uint32_t V32 = 0, E32 = 0;
    std::memcpy(&V32, hdr, 4);
    std::memcpy(&E32, hdr + 4, 4);
    if (got == 16) {
        std::memcpy(&V64, hdr, 8);
        std::memcpy(&E64, hdr + 8, 8);
    }
    if (format != ASC_FMT_AUTO) return read_graph_as(fh.f, path, size, format, V64, E64, V32, E32, out);
    // Every layout whose size matches the header is tried in this order; the
    // first that reads and validates wins.  Two layouts can match one size
    // (BWCSR and BCSC when V == E), so a size match alone does not decide.
    const int order[] = {ASC_FMT_BCSR, ASC_FMT_BWCSR, ASC_FMT_BCSC, ASC_FMT_LEGACY32, ASC_FMT_LEGACY32W};
    int r = ASC_ERR_FORMAT;
    bool matched = false;
    for (int f : order) {
        const bool wide = f <= ASC_FMT_BCSC;
        const uint64_t V = wide ? V64 : V32, E = wide ? E64 : E32;
        if (wide && got < 16) continue;
        if (!header_plausible(V, E, size) || layout_size(f, V, E) != size) continue;
        matched = true;
        HostCSR h;
        r = read_graph_as(fh.f, path, size, f, V64, E64, V32, E32, h);
        if (r == ASC_OK) {
            out = std::move(h);
            return ASC_OK;
        }
        if (r != ASC_ERR_FORMAT && r != ASC_ERR_RANGE) return r;   // I/O or memory: no other layout helps
    }
    if (!matched) set_last_error("%s: size %llu matches no layout", path, (unsigned long long)size);
    return r;
}

static int read_graph_as(FILE* f, const char* path, uint64_t size, int fmt, uint64_t V64, uint64_t E64,
                         uint32_t V32, uint32_t E32, HostCSR& out) {
    const bool wide = fmt == ASC_FMT_BCSR || fmt == ASC_FMT_BWCSR || fmt == ASC_FMT_BCSC;
    const uint64_t V = wide ? V64 : V32, E = wide ? E64 : E32;
    if (fmt < ASC_FMT_BCSR || fmt > ASC_FMT_LEGACY32W) return ASC_ERR_INVALID;
    if (!header_plausible(V, E, size) || layout_size(fmt, V, E) != size) {
        set_last_error("%s: size %llu does not match header V=%llu E=%llu", path, (unsigned long long)size,
                       (unsigned long long)V, (unsigned long long)E);
        return ASC_ERR_FORMAT;
    }
    if (fseek(f, wide ? 16 : 8, SEEK_SET) != 0) return ASC_ERR_IO;
    out.V = V;
    out.E = E;
    out.offset_bytes = wide ? 8 : 4;
    out.kind = (fmt == ASC_FMT_BWCSR || fmt == ASC_FMT_LEGACY32W) ? ASC_KIND_WCSR
               : (fmt == ASC_FMT_BCSC) ? ASC_KIND_CSC
                                         : ASC_KIND_CSR;
    if (fmt == ASC_FMT_BCSC) {
        ASC_TRY(out.outdeg.alloc(V));
        ASC_TRY(read_exact(f, out.outdeg.p, 4 * V));
    }
    ASC_TRY(out.off.alloc(V + 1));
    if (wide) {
        ASC_TRY(read_exact(f, out.off.p, 8 * V));
    } else {
        HostBuf<uint32_t> o32;
        ASC_TRY(o32.alloc(V));
        ASC_TRY(read_exact(f, o32.p, 4 * V));
        host_pool().parallel_for(V, [&](uint64_t b, uint64_t e) {
            for (uint64_t i = b; i < e; i++) out.off[i] = o32[i];
        });
    }
    out.off[V] = E;
    ASC_TRY(out.edges.alloc(E * out.edge_words() + 1));
    int r = read_exact(f, out.edges.p, E * out.edge_bytes());
    if (r != ASC_OK) {
        set_last_error("%s: short read of the edge array", path);
        return r;
    }
    return validate_csr(out);
}

int csr_from_view(const asc_csr_view* v, HostCSR& out, bool borrow) {
    if (!v || v->num_vertices == 0 || (!v->off64 && !v->off32) || (v->num_edges && !v->edges)) {
        set_last_error("invalid csr view");
        return ASC_ERR_INVALID;
    }
    if (v->kind < ASC_KIND_CSR || v->kind > ASC_KIND_CSC) return ASC_ERR_INVALID;
    if (v->kind == ASC_KIND_CSC && !v->out_degree) {
        set_last_error("CSC view needs out_degree");
        return ASC_ERR_INVALID;
    }
    const uint64_t V = v->num_vertices, E = v->num_edges;
    out.V = V;
    out.E = E;
    out.kind = v->kind;
    out.offset_bytes = v->off64 ? 8 : 4;
    if (v->has_sentinel) {
        uint64_t s = v->off64 ? v->off64[V] : v->off32[V];
        if (s != E) {
            set_last_error("sentinel offset %llu != E", (unsigned long long)s);
            return ASC_ERR_FORMAT;
        }
    }
    // borrowed: the records and out-degrees (and u64 offsets with their
    // sentinel) stay the caller's; narrower offsets are widened into a copy
    if (borrow && v->off64 && v->has_sentinel) {
        out.off.borrow(const_cast<uint64_t*>(v->off64), V + 1);
    } else {
        ASC_TRY(out.off.alloc(V + 1));
        for (uint64_t i = 0; i < V; i++) out.off[i] = v->off64 ? v->off64[i] : v->off32[i];
        out.off[V] = E;
    }
    if (borrow) {
        if (E) out.edges.borrow(const_cast<uint32_t*>(v->edges), E * out.edge_words());
    } else {
        ASC_TRY(out.edges.alloc(E * out.edge_words() + 1));
        if (E) std::memcpy(out.edges.p, v->edges, E * out.edge_bytes());
    }
    if (v->kind == ASC_KIND_CSC) {
        if (borrow) {
            out.outdeg.borrow(const_cast<uint32_t*>(v->out_degree), V);
        } else {
            ASC_TRY(out.outdeg.alloc(V));
            std::memcpy(out.outdeg.p, v->out_degree, 4 * V);
        }
    }
    return validate_csr(out);
}

int write_graph_file(const char* path, int format, const asc_csr_view* v) {
    if (!path || !v || v->num_vertices == 0 || (!v->off64 && !v->off32)) return ASC_ERR_INVALID;
    const uint64_t V = v->num_vertices, E = v->num_edges;
    const bool wide = format == ASC_FMT_BCSR || format == ASC_FMT_BWCSR || format == ASC_FMT_BCSC;
    const bool weighted = format == ASC_FMT_BWCSR || format == ASC_FMT_LEGACY32W;
    if (format < ASC_FMT_BCSR || format > ASC_FMT_LEGACY32W) return ASC_ERR_INVALID;
    if (weighted != (v->kind == ASC_KIND_WCSR)) {
        set_last_error("format weightedness does not match the graph kind");
        return ASC_ERR_UNSUPPORTED;
    }
    if (format == ASC_FMT_BCSC && (v->kind != ASC_KIND_CSC || !v->out_degree)) {
        set_last_error("BCSC needs an in-CSR with out-degrees");
        return ASC_ERR_UNSUPPORTED;
    }
    if (!wide && (V >= 0xFFFFFFFFull || E >= 0xFFFFFFFFull)) return ASC_ERR_RANGE;
    File fh;
    fh.f = fopen(path, "wb");
    if (!fh.f) {
        set_last_error("cannot create %s", path);
        return ASC_ERR_IO;
    }
    if (wide) {
        uint64_t hdr[2] = {V, E};
        ASC_TRY(write_exact(fh.f, hdr, 16));
    } else {
        uint32_t hdr[2] = {(uint32_t)V, (uint32_t)E};
        ASC_TRY(write_exact(fh.f, hdr, 8));
    }
    if (format == ASC_FMT_BCSC) ASC_TRY(write_exact(fh.f, v->out_degree, 4 * V));
    const uint64_t chunk = 1 << 20;
    if (wide) {
        std::vector<uint64_t> buf;
        for (uint64_t b = 0; b < V; b += chunk) {
            uint64_t e = std::min(V, b + chunk);
            buf.resize(e - b);
            for (uint64_t i = b; i < e; i++) buf[i - b] = v->off64 ? v->off64[i] : v->off32[i];
            ASC_TRY(write_exact(fh.f, buf.data(), 8 * (e - b)));
        }
    } else {
        std::vector<uint32_t> buf;
        for (uint64_t b = 0; b < V; b += chunk) {
            uint64_t e = std::min(V, b + chunk);
            buf.resize(e - b);
            for (uint64_t i = b; i < e; i++) buf[i - b] = (uint32_t)(v->off64 ? v->off64[i] : v->off32[i]);
            ASC_TRY(write_exact(fh.f, buf.data(), 4 * (e - b)));
        }
    }
    ASC_TRY(write_exact(fh.f, v->edges, E * (weighted ? 8 : 4)));
    if (fflush(fh.f) != 0) return ASC_ERR_IO;
    return ASC_OK;
}

}  // namespace asc

extern "C" int asc_write_graph(const char* path, int format, const asc_csr_view* view) {
    return asc::write_graph_file(path, format, view);
}

extern "C" int asc_read_graph(const char* path, int format, asc_host_csr* out) {
    if (!out) return ASC_ERR_INVALID;
    std::memset(out, 0, sizeof(*out));
    asc::HostCSR h;
    ASC_TRY(asc::read_graph_file(path, format, h));
    out->num_vertices = h.V;
    out->num_edges = h.E;
    out->kind = h.kind;
    out->off = h.off.p;
    out->edges = h.edges.p;
    out->out_degree = h.outdeg.p;
    h.off.p = nullptr;  // ownership moves to the caller (asc_host_csr_free)
    h.edges.p = nullptr;
    h.outdeg.p = nullptr;
    return ASC_OK;
}
