// k_push32.hip -- FP32 form (placeholder until the FP32 kernels land).
#include <vector>

#include "pic_internal.h"

namespace pic {

static int fp32_missing(pic_ctx* c) {
    return set_err(c, PIC_ERR_STATE, "the FP32 form is not built yet");
}
int launch_push32(pic_ctx* c, int) { return fp32_missing(c); }
int sort32(pic_ctx* c, int) { return fp32_missing(c); }
int boundary32(pic_ctx* c, int) { return fp32_missing(c); }
int loader32(pic_ctx* c, int, const pic_loader*, int64_t, const int64_t*) { return fp32_missing(c); }
int diag_p32(pic_ctx* c, int, double) { return fp32_missing(c); }

}  // namespace pic

int particles_set32(pic_ctx* c, int, int64_t, const double*, const double*, const double*,
                    const double*, const double*, const double*, const double*, const int64_t*) {
    return pic::set_err(c, PIC_ERR_STATE, "the FP32 form is not built yet");
}
int particles_get32(pic_ctx* c, int, int64_t, std::vector<double>*, std::vector<int64_t>&) {
    return pic::set_err(c, PIC_ERR_STATE, "the FP32 form is not built yet");
}
