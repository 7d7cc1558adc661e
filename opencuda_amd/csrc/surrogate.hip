// surrogate.hip -- learned BF16 surrogate (placeholder entry points).
#include <hip/hip_runtime.h>

#include "bf_internal.h"

extern "C" {
int64_t bf_sur_param_count(void) { return -998; }
int bf_sur_create(uint64_t, bf_sur *) { return -998; }
int bf_sur_destroy(bf_sur) { return -998; }
int bf_sur_get_params(bf_sur, float *) { return -998; }
int bf_sur_set_params(bf_sur, const float *) { return -998; }
int bf_sur_train_step(bf_sur, const unsigned char *, const unsigned char *, const float *, int32_t,
                      int32_t, int32_t, float, float, float, float, float *, void *)
{
    return -998;
}
int bf_sur_forward_backward(bf_sur, const unsigned char *, const unsigned char *, const float *,
                            int32_t, int32_t, int32_t, float *, float *, float *, void *)
{
    return -998;
}
int bf_sur_infer(bf_sur, const bf_image_view *, const bf_image_view *, float, void *) { return -998; }
int bf_sur_filter(bf_sur, const bf_image_view *, const bf_image_view *, bf_lut, int32_t, int32_t,
                  float, int32_t, int32_t *, int32_t *, void *)
{
    return -998;
}
}
