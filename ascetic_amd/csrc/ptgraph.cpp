// ptgraph: command-line driver over the C ABI (include/ascetic.h).
//
// Keeps the reference CLI (main.cu:4-22, ArgumentParser.cu:15-48):
//   ptgraph --input <file> --type bfs|cc|sssp|pr --source <n> --adviseK <f> [--method <n>]
// flags are key/value pairs, --type defaults to bfs, --source to 0,
// --adviseK <= 0 selects the automatic static-region rule, --method is
// accepted and ignored.  Additions:
//   --budget <bytes>        device edge budget (replaces --adviseK)
//   --part-mb <n>           on-demand part buffer size
//   --stream gather|zerocopy|hybrid
//   --format auto|bcsr|bwcsr|bcsc|legacy32|legacy32w
//   --values ref|zero       output convention (source = 1 / 0)
//   --pr-mode ref|norm      PageRank: reference active-set FP64 (default) or
//                           normalised fixed iterations (--iters, default 20)
//   --out <file>            write the V result values (raw little-endian)
//   --verbose 0|1|2         per-phase timings in the reference's wording (2: per round)
//   --dense f               dense-chunk fraction (0 = per round, >1 = gather only)
//   --overlap 0|1           resident steps between streamed chunks (SSSP, CC)
//   --bfs td|auto           top-down only, or direction-optimizing when possible
//   --cc lp|auto            label propagation only, or union-find when possible
//   --pr-method auto|pull|blocked|tiled
//   --type gen              write a synthetic graph: --gen rmat|er --scale s
//                           --ef f --vertices n --samples m --seed x
//                           --weighted 0|1 --wmax w --csc 0|1 --out file
//                           --format bcsr|bwcsr|bcsc|legacy32|legacy32w
//   --type info             print the header, layout and planned split
//   --type convert          edge list -> binary graph (converter/converter.cpp):
//                           --input edges.txt|.mtx|.toB --out file --format
//                           bcsr|bwcsr|bcsc|legacy32|legacy32w [--symmetrize 0|1]
//                           [--dedup 0|1] [--self-loops 0|1] [--one-based 0|1]
//                           [--vertices n] [--wmin w --wmax w --seed x]
// Every run ends with one JSON line of statistics.  Errors are reported with
// the library's message and a nonzero exit status; nothing calls exit()
// inside the library.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/ascetic.h"

extern "C" const char* asc_last_error(void);

namespace {

struct Args {
    std::map<std::string, std::string> kv;
    std::string get(const char* k, const char* def = "") const {
        auto it = kv.find(k);
        return it == kv.end() ? def : it->second;
    }
    double num(const char* k, double def) const {
        auto it = kv.find(k);
        return it == kv.end() ? def : atof(it->second.c_str());
    }
    bool has(const char* k) const { return kv.count(k) != 0; }
};

void usage() {
    fprintf(stderr,
            "usage: ptgraph --input <graph file> [--type bfs|cc|sssp|pr|info] [--source N] [--adviseK f]\n"
            "               [--budget bytes] [--part-mb n] [--stream gather|zerocopy|hybrid] [--format fmt]\n"
            "               [--values ref|zero] [--pr-mode ref|norm] [--pr-method auto|pull|blocked|tiled] [--iters n]\n"
            "               [--out file] [--verbose 0|1|2] [--dense f] [--overlap 0|1] [--bfs td|auto] [--cc lp|auto]\n"
            "               [--pack 0|1|2]  (streamed records: 8-byte / packed 4 or 5 bytes / always 5 bytes)\n"
            "       ptgraph --type gen --gen rmat|er [--scale s --ef f | --vertices n --samples m] [--seed x]\n"
            "               [--weighted 0|1] [--wmax w] [--csc 0|1] --format fmt --out file\n"
            "       ptgraph --type convert --input edges.txt|.mtx|.toB --format fmt --out file [--symmetrize 0|1]\n"
            "               [--dedup 0|1] [--self-loops 0|1] [--one-based 0|1] [--vertices n] [--wmin w --wmax w]\n");
}

int fail(int code, const char* what) {
    fprintf(stderr, "ptgraph: %s: %s (%d) %s\n", what, asc_strerror(code), code, asc_last_error());
    return 1;
}

int parse_format(const std::string& f) {
    if (f.empty() || f == "auto") return ASC_FMT_AUTO;
    if (f == "bcsr") return ASC_FMT_BCSR;
    if (f == "bwcsr") return ASC_FMT_BWCSR;
    if (f == "bcsc") return ASC_FMT_BCSC;
    if (f == "legacy32") return ASC_FMT_LEGACY32;
    if (f == "legacy32w") return ASC_FMT_LEGACY32W;
    return -1;
}

int write_raw(const std::string& path, const void* p, size_t bytes) {
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) return ASC_ERR_IO;
    size_t w = fwrite(p, 1, bytes, f);
    int r = fclose(f);
    return (w == bytes && r == 0) ? ASC_OK : ASC_ERR_IO;
}

void print_stats_json(const char* algo, const asc_graph_info& gi, const asc_stats& st, double gteps) {
    printf("{\"algo\": \"%s\", \"V\": %llu, \"E\": %llu, \"static_edges\": %llu, \"ondemand_edges\": %llu, "
           "\"iterations\": %u, \"node_sum\": %llu, \"edges_traversed\": %llu, \"streamed_bytes\": %llu, "
           "\"parts\": %u, \"reached\": %llu, \"total_ms\": %.3f, \"static_ms\": %.3f, \"overload_ms\": %.3f, "
           "\"fill_ms\": %.3f, \"gteps\": %.4f, \"host_cores\": %d}\n",
           algo, (unsigned long long)gi.num_vertices, (unsigned long long)gi.num_edges,
           (unsigned long long)gi.static_edges, (unsigned long long)gi.ondemand_edges, st.iterations,
           (unsigned long long)st.node_sum, (unsigned long long)st.edges_traversed,
           (unsigned long long)st.streamed_bytes, st.parts, (unsigned long long)st.reached, st.total_ms,
           st.static_ms, st.overload_ms, st.fill_ms, gteps, st.host_cores);
}

int run_gen(const Args& a) {
    asc_gen_params p;
    asc_gen_params_init(&p);
    std::string kind = a.get("--gen", "rmat");
    p.type = kind == "er" ? 1 : 0;
    p.scale = (int)a.num("--scale", 16);
    p.edge_factor = (int)a.num("--ef", 16);
    p.num_vertices = (uint64_t)a.num("--vertices", 0);
    p.num_samples = (uint64_t)a.num("--samples", 0);
    p.seed = (uint64_t)a.num("--seed", 1);
    p.weighted = (int)a.num("--weighted", 0);
    p.wmin = (uint32_t)a.num("--wmin", 1);
    p.wmax = (uint32_t)a.num("--wmax", 64);
    p.csc = (int)a.num("--csc", 0);
    p.symmetrize = (int)a.num("--symmetrize", 1);
    int fmt = parse_format(a.get("--format", p.csc ? "bcsc" : p.weighted ? "bwcsr" : "bcsr"));
    std::string out = a.get("--out");
    if (fmt <= 0 || out.empty()) {
        usage();
        return 2;
    }
    asc_host_csr h;
    int r = asc_generate(&p, &h);
    if (r) return fail(r, "asc_generate");
    asc_csr_view v;
    asc_host_csr_view(&h, &v);
    r = asc_write_graph(out.c_str(), fmt, &v);
    printf("{\"generated\": \"%s\", \"V\": %llu, \"E\": %llu}\n", out.c_str(), (unsigned long long)h.num_vertices,
           (unsigned long long)h.num_edges);
    asc_host_csr_free(&h);
    return r ? fail(r, "asc_write_graph") : 0;
}

int run_convert(const Args& a) {
    asc_edge_list_opts o;
    asc_edge_list_opts_init(&o);
    const std::string input = a.get("--input"), out = a.get("--out");
    const int fmt = parse_format(a.get("--format", "bcsr"));
    if (input.empty() || out.empty() || fmt <= 0) {
        usage();
        return 2;
    }
    o.weighted = fmt == ASC_FMT_BWCSR || fmt == ASC_FMT_LEGACY32W;
    o.csc = fmt == ASC_FMT_BCSC;
    o.symmetrize = (int)a.num("--symmetrize", 0);
    o.dedup = (int)a.num("--dedup", 1);
    o.keep_self_loops = (int)a.num("--self-loops", 0);
    o.one_based = (int)a.num("--one-based", 0);
    o.num_vertices = (uint64_t)a.num("--vertices", 0);
    o.wmin = (uint32_t)a.num("--wmin", 1);
    o.wmax = (uint32_t)a.num("--wmax", 64);
    o.seed = (uint64_t)a.num("--seed", 1);
    asc_host_csr h;
    int r = asc_read_edge_list(input.c_str(), &o, &h);
    if (r) return fail(r, "asc_read_edge_list");
    asc_csr_view v;
    asc_host_csr_view(&h, &v);
    r = asc_write_graph(out.c_str(), fmt, &v);
    printf("{\"converted\": \"%s\", \"V\": %llu, \"E\": %llu}\n", out.c_str(),
           (unsigned long long)h.num_vertices, (unsigned long long)h.num_edges);
    asc_host_csr_free(&h);
    return r ? fail(r, "asc_write_graph") : 0;
}

}  // namespace

int main(int argc, char** argv) {
    Args a;
    // key/value pairs, as ArgumentParser::Parse walks them (ArgumentParser.cu:17)
    for (int i = 1; i + 1 < argc; i += 2) a.kv[argv[i]] = argv[i + 1];
    if (argc == 2 && (!strcmp(argv[1], "--help") || !strcmp(argv[1], "-h"))) {
        usage();
        return 0;
    }
    std::string type = a.get("--type", "bfs");  // main.cu:7-9
    if (type == "gen") return run_gen(a);
    if (type == "convert") return run_convert(a);
    std::string input = a.get("--input");
    if (input.empty()) {
        usage();
        return 2;
    }
    int fmt = parse_format(a.get("--format", "auto"));
    if (fmt < 0) {
        usage();
        return 2;
    }
    // the algorithm implies the layout where only one serves it: PageRank
    // reads the in-CSR file (GraphMeta.cuh:159-162)
    if (fmt == ASC_FMT_AUTO && type == "pr") fmt = ASC_FMT_BCSC;

    asc_config cfg;
    asc_config_init(&cfg);
    cfg.advise_rate = a.num("--adviseK", 0.0);
    cfg.edge_budget_bytes = (uint64_t)a.num("--budget", 0);
    cfg.part_bytes = (uint64_t)a.num("--part-mb", 0) * (1ull << 20);
    const std::string sm = a.get("--stream", "gather");
    cfg.stream_mode = sm == "zerocopy" ? ASC_STREAM_ZEROCOPY : sm == "hybrid" ? ASC_STREAM_HYBRID : ASC_STREAM_GATHER;
    cfg.dense_threshold = a.num("--dense", 0.0);
    cfg.stream_overlap = (int)a.num("--overlap", 1);
    cfg.stream_pack = (int)a.num("--pack", 1);
    cfg.bfs_direction = a.get("--bfs", "auto") == "td" ? ASC_BFS_TOP_DOWN : ASC_BFS_AUTO;
    cfg.cc_method = a.get("--cc", "auto") == "lp" ? ASC_CC_LABEL_PROP : ASC_CC_AUTO;
    cfg.value_convention = a.get("--values", "ref") == "zero" ? ASC_VALUES_ZERO_BASED : ASC_VALUES_REFERENCE;
    cfg.verbose = (int)a.num("--verbose", 0);
    cfg.profile = 1;
    asc_ctx* ctx = nullptr;
    int r = asc_ctx_create(&cfg, &ctx);
    if (r) return fail(r, "asc_ctx_create");
    asc_graph* g = nullptr;
    r = asc_graph_load(ctx, input.c_str(), fmt, &g);
    if (r) {
        asc_ctx_destroy(ctx);
        return fail(r, "asc_graph_load");
    }
    asc_graph_info gi;
    asc_graph_get_info(g, &gi);
    if (type == "info") {
        printf("{\"V\": %llu, \"E\": %llu, \"kind\": %d, \"offset_bytes\": %d, \"static_edges\": %llu, "
               "\"static_vertices\": %llu, \"ondemand_edges\": %llu, \"part_edges\": %llu, \"max_degree\": %u, "
               "\"max_degree_vertex\": %u}\n",
               (unsigned long long)gi.num_vertices, (unsigned long long)gi.num_edges, gi.kind, gi.offset_bytes,
               (unsigned long long)gi.static_edges, (unsigned long long)gi.static_vertices,
               (unsigned long long)gi.ondemand_edges, (unsigned long long)gi.part_edges, gi.max_degree,
               gi.max_degree_vertex);
        asc_graph_destroy(g);
        asc_ctx_destroy(ctx);
        return 0;
    }
    const uint32_t src = (uint32_t)a.num("--source", 0);  // ArgumentParser.cu:8
    const std::string out = a.get("--out");
    asc_stats st;
    memset(&st, 0, sizeof(st));
    double gteps = 0;
    int rc = 0;
    if (type == "bfs" || type == "sssp" || type == "cc") {
        std::vector<uint32_t> vals(gi.num_vertices);
        if (type == "bfs") r = asc_bfs(g, src, vals.data(), &st);
        else if (type == "sssp") r = asc_sssp(g, src, vals.data(), &st);
        else r = asc_cc(g, vals.data(), &st);
        if (r) {
            rc = fail(r, type.c_str());
        } else {
            const double edges = type == "cc" ? (double)gi.num_edges / 2 : (double)st.reached_degree_sum / 2;
            gteps = st.total_ms > 0 ? edges / (st.total_ms * 1e-3) / 1e9 : 0;
            print_stats_json(type.c_str(), gi, st, gteps);
            if (!out.empty() && write_raw(out, vals.data(), 4 * vals.size())) rc = fail(ASC_ERR_IO, "--out");
        }
    } else if (type == "pr") {
        asc_pr_opts o;
        asc_pr_opts_init(&o);
        const bool norm = a.get("--pr-mode", "ref") == "norm";
        o.mode = norm ? ASC_PR_NORMALISED : ASC_PR_REFERENCE;
        const std::string pm = a.get("--pr-method", "auto");
        o.method = pm == "pull" ? ASC_PR_PULL : pm == "blocked" ? ASC_PR_BLOCKED : pm == "tiled" ? ASC_PR_TILED : ASC_PR_AUTO;
        o.iterations = (int)a.num("--iters", 20);
        if (norm) {
            std::vector<float> rank(gi.num_vertices);
            r = asc_pr(g, &o, rank.data(), nullptr, nullptr, &st);
            if (!r && !out.empty() && write_raw(out, rank.data(), 4 * rank.size())) r = ASC_ERR_IO;
        } else {
            std::vector<double> rank(gi.num_vertices);
            r = asc_pr(g, &o, nullptr, rank.data(), nullptr, &st);
            if (!r && !out.empty() && write_raw(out, rank.data(), 8 * rank.size())) r = ASC_ERR_IO;
        }
        if (r) {
            rc = fail(r, "pr");
        } else {
            gteps = st.total_ms > 0 ? (double)gi.num_edges * st.iterations / (st.total_ms * 1e-3) / 1e9 : 0;
            print_stats_json("pr", gi, st, gteps);
        }
    } else {
        usage();
        rc = 2;
    }
    asc_graph_destroy(g);
    asc_ctx_destroy(ctx);
    return rc;
}
