// Entry points whose kernels are not written yet.
#include "smdp_internal.h"
extern "C" {
int smdp_sssp_multi_source(smdp_ctx*, const smdp_csr*, const int32_t*, int32_t, int32_t*, smdp_stats*) { return SMDP_ERR_SOLUTION; }
int smdp_sssp_multi_graph(smdp_ctx*, const smdp_csr*, int32_t, const int32_t*, int32_t* const*, smdp_stats*) { return SMDP_ERR_SOLUTION; }
int smdp_sssp_multi_source_run(smdp_graph*, const int32_t*, int32_t, smdp_stats*) { return SMDP_ERR_SOLUTION; }
int smdp_batch_create(smdp_ctx*, const smdp_csr*, int32_t, const int32_t*, smdp_batch**) { return SMDP_ERR_SOLUTION; }
int smdp_batch_run(smdp_batch*, smdp_stats*) { return SMDP_ERR_SOLUTION; }
int smdp_batch_read(smdp_batch*, int32_t, int32_t*) { return SMDP_ERR_SOLUTION; }
void smdp_batch_destroy(smdp_batch*) {}
int smdp_autotune(smdp_graph*, int, int32_t, const char*, smdp_params*) { return SMDP_ERR_SOLUTION; }
}
