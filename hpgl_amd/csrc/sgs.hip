// sgs.hip - sequential Gaussian simulation entry points.
#include <hip/hip_runtime.h>

#include <stdexcept>

#include "../../include/hpgl.h"

extern "C" __attribute__((visibility("default"))) void hpgl_amd_report_error(const char *msg);

extern "C" __attribute__((visibility("default"))) void hpgl_sgs_simulation(hpgl_cont_masked_array_t *,
                                                                           hpgl_sgs_params_t *,
                                                                           hpgl_non_parametric_cdf_t *, double *,
                                                                           hpgl_ubyte_array_t *) {
    hpgl_amd_report_error("hpgl_sgs_simulation: not available in this build");
}

extern "C" __attribute__((visibility("default"))) void hpgl_sgs_lvm_simulation(hpgl_cont_masked_array_t *,
                                                                               hpgl_sgs_params_t *,
                                                                               hpgl_non_parametric_cdf_t *,
                                                                               hpgl_float_array_t *,
                                                                               hpgl_ubyte_array_t *) {
    hpgl_amd_report_error("hpgl_sgs_lvm_simulation: not available in this build");
}
