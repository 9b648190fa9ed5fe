// tg_mixed.hip -- batched mixed-precision collision operator (MFMA).
#include "tg_internal.hpp"

namespace tg {

tg_status mixed_prepare(tg_ctx* c) { return TG_OK; }

tg_status mixed_apply_device(tg_ctx* c, const double* d_f, double* d_I, int32_t batch,
                             const double* h_bath, int conserve) {
    return set_error(c, TG_ERR_STATE, "mixed-precision path not built");
}

void mixed_destroy(tg_ctx* c) {}

}  // namespace tg
