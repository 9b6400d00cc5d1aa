// Temporary stubs.
#include "common.hpp"
extern "C" {
int scir_hip_ragged_plan_create(const int64_t *, const int64_t *, int64_t, int32_t, scir_hip_ragged_plan *) { return scir::fail(SCIR_HIP_ERR_UNSUPPORTED, "ragged not built yet"); }
int scir_hip_ragged_plan_execute(scir_hip_ragged_plan, const float *, const float *, float *, void *) { return scir::fail(SCIR_HIP_ERR_UNSUPPORTED, "ragged not built yet"); }
int scir_hip_ragged_plan_execute_per_signal(scir_hip_ragged_plan, const float *, const float *, float *, void *) { return scir::fail(SCIR_HIP_ERR_UNSUPPORTED, "ragged not built yet"); }
int scir_hip_ragged_plan_destroy(scir_hip_ragged_plan) { return 0; }
int scir_hip_fir_ragged_f32_dev(const float *, const int64_t *, const float *, const int64_t *, int64_t, float *, int32_t, void *) { return scir::fail(SCIR_HIP_ERR_UNSUPPORTED, "ragged not built yet"); }
}
