// Error text, the host thread pool and small C-ABI helpers.
#include <cstdarg>

#include "internal.hpp"

namespace asc {

static thread_local char g_err[512] = {0};

void set_last_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}
const char* last_error() { return g_err; }

int hardware_threads() {
    unsigned n = std::thread::hardware_concurrency();
    return n ? (int)n : 1;
}

ThreadPool::ThreadPool(int n) {
    if (n < 1) n = 1;
    for (int i = 1; i < n; i++) workers_.emplace_back([this, i] { loop(i); });
}

ThreadPool::~ThreadPool() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : workers_) t.join();
}

void ThreadPool::loop(int id) {
    uint64_t seen = 0;
    for (;;) {
        const std::function<void(int, int)>* job;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
            if (stop_) return;
            seen = gen_;
            job = job_;
        }
        (*job)(id, size());
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (--pending_ == 0) done_cv_.notify_all();
        }
    }
}

void ThreadPool::run(const std::function<void(int, int)>& fn) {
    if (workers_.empty()) {
        fn(0, 1);
        return;
    }
    {
        std::lock_guard<std::mutex> lk(mu_);
        job_ = &fn;
        pending_ = (int)workers_.size();
        gen_++;
    }
    cv_.notify_all();
    fn(0, size());
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [&] { return pending_ == 0; });
    job_ = nullptr;
}

void ThreadPool::parallel_for(uint64_t n, const std::function<void(uint64_t, uint64_t)>& fn) {
    if (n == 0) return;
    run([&](int t, int nt) {
        uint64_t chunk = (n + nt - 1) / nt;
        uint64_t b = (uint64_t)t * chunk, e = b + chunk < n ? b + chunk : n;
        if (b < e) fn(b, e);
    });
}

ThreadPool& host_pool() {
    static ThreadPool pool(hardware_threads() < 16 ? hardware_threads() : 16);
    return pool;
}

}  // namespace asc

extern "C" {

const char* asc_strerror(int code) {
    switch (code) {
        case ASC_OK: return "ok";
        case ASC_ERR_INVALID: return "invalid argument";
        case ASC_ERR_IO: return "i/o error";
        case ASC_ERR_FORMAT: return "file does not match the graph format";
        case ASC_ERR_NOMEM: return "out of memory";
        case ASC_ERR_DEVICE: return "HIP runtime error";
        case ASC_ERR_NOGPU: return "no HIP device";
        case ASC_ERR_RANGE: return "size out of range";
        case ASC_ERR_UNSUPPORTED: return "graph kind does not support this operation";
        case ASC_ERR_BUSY: return "busy";
        case ASC_ERR_STATE: return "invalid state";
        default: return "unknown error";
    }
}

int asc_abi_version(void) { return ASC_ABI_VERSION; }

// Text of the last error on the calling thread (not in the public header's
// status list: a diagnostic for the CLI and the Python wrapper).
const char* asc_last_error(void) { return asc::last_error(); }

void asc_config_init(asc_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->device = 0;
    c->advise_rate = 0.0;
    c->edge_budget_bytes = 0;
    c->part_bytes = 0;
    c->device_mem_limit = 0;
    c->gather_threads = 0;
    c->stream_mode = ASC_STREAM_GATHER;
    c->value_convention = ASC_VALUES_REFERENCE;
    c->heavy_threshold = 0;
    c->pr_group = 0;
    c->sssp_delta = 0;
    c->verbose = 0;
    c->profile = 0;
    c->stream_pack = 1;
    c->bfs_direction = ASC_BFS_AUTO;
    c->cc_method = ASC_CC_AUTO;
    c->dense_threshold = 0.0;
    c->stream_overlap = 1;
    c->bfs_persistent = 0;
    c->release_host_edges = 0;
    c->part_buffers = 0;
    c->pad_cfg = 0;
    c->max_pinned_bytes = 0;
}

void asc_pr_opts_init(asc_pr_opts* o) {
    if (!o) return;
    o->mode = ASC_PR_NORMALISED;
    o->damping = 0.85;
    o->iterations = 20;
    o->threshold = 0.001;
    o->max_iterations = 0;
    o->method = ASC_PR_AUTO;
    o->tile_src_bits = 0;
    o->tile_dst_bits = 0;
}

void asc_gen_params_init(asc_gen_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->type = 0;
    p->scale = 10;
    p->edge_factor = 16;
    p->a = 0.57;
    p->b = 0.19;
    p->c = 0.19;
    p->seed = 1;
    p->symmetrize = 1;
    p->weighted = 0;
    p->wmin = 1;
    p->wmax = 64;
    p->offset_bytes = 8;
    p->csc = 0;
    p->threads = 0;
}

void asc_server_opts_init(asc_server_opts* o) {
    if (!o) return;
    o->lanes = 0;  // the graph's tuned batch width (64 until tuned)
    o->state_budget_bytes = 0;
    o->queue_capacity = 4096;
    o->bfs_direction = ASC_BFS_AUTO;
    o->step_graphs = 1;
    o->pad = 0;
}

}  // extern "C"
