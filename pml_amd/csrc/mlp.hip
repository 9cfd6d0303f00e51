// Learned cost model of the batched dispatch (placeholder until the model is
// trained: no model loaded, the dispatcher falls back to the autotuner cache
// and the threshold rule).
#include "dispatch.hpp"

namespace pmlg {

int learned_config(pmlg_ctx, int64_t, int64_t, int64_t, int64_t, int64_t)
{
    return -1;
}

}  // namespace pmlg
