// Entry points whose kernels are not written yet.
#include "smdp_internal.h"
extern "C" {
int smdp_autotune(smdp_graph*, int, int32_t, const char*, smdp_params*) { return SMDP_ERR_SOLUTION; }
}
