// Batched query serving (placeholder until the multi-source engine lands).
#include "internal.hpp"

extern "C" {
int asc_server_create(asc_graph*, const asc_server_opts*, asc_server** out) {
    if (out) *out = nullptr;
    return ASC_ERR_UNSUPPORTED;
}
int asc_server_destroy(asc_server*) { return ASC_ERR_UNSUPPORTED; }
int asc_submit(asc_server*, int, uint32_t, uint64_t*) { return ASC_ERR_UNSUPPORTED; }
int asc_wait(asc_server*, uint64_t, uint32_t*, asc_qstats*) { return ASC_ERR_UNSUPPORTED; }
int asc_server_drain(asc_server*) { return ASC_ERR_UNSUPPORTED; }
}
