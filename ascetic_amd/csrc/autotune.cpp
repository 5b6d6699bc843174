// Autotuner (placeholder until the sweep lands).
#include "internal.hpp"

extern "C" {
int asc_autotune(asc_graph*, double, asc_tuning*) { return ASC_ERR_UNSUPPORTED; }
int asc_tuning_save(const char*, const asc_graph*, const asc_tuning*) { return ASC_ERR_UNSUPPORTED; }
int asc_tuning_load(const char*, asc_graph*, asc_tuning*) { return ASC_ERR_UNSUPPORTED; }
}
