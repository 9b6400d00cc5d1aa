// Placeholder tuning table (replaced by the autotuner).
#include "common.hpp"
namespace scir {
TuneChoice tune_lookup(int, int64_t) { return TuneChoice{}; }
}
extern "C" int scir_hip_autotune(const char *, int32_t) { return scir::fail(SCIR_HIP_ERR_UNSUPPORTED, "autotune not built yet"); }
extern "C" int scir_hip_autotune_load(const char *) { return scir::fail(SCIR_HIP_ERR_UNSUPPORTED, "autotune not built yet"); }
