// Adjoint (reverse-mode) kernels of the path stage. Entry points report FM_EUNSUPPORTED until implemented.
#include <hip/hip_runtime.h>

#include "fm_device.h"

int fm_set_err(int code, const char *msg);

extern "C" {

int fm_cchi_backward(const fm_edge *, int32_t, int32_t, int32_t, int32_t, const double *, const int64_t *, const void *,
                     const double *, double *, void *) {
    return fm_set_err(FM_EUNSUPPORTED, "fm_cchi_backward is not implemented yet");
}

int fm_chi_fine_backward(const fm_edge *, int32_t, int32_t, const void *, const double *, const fm_grid *, const void *,
                         void *, double *, void *) {
    return fm_set_err(FM_EUNSUPPORTED, "fm_chi_fine_backward is not implemented yet");
}

int fm_chi_paths_backward(const fm_edge *, int32_t, const fm_paths *, const fm_grid *, int32_t, int32_t, const int32_t *,
                          const int32_t *, const double *, const void *, int32_t, const int32_t *, const int32_t *,
                          double *, double *, void *) {
    return fm_set_err(FM_EUNSUPPORTED, "fm_chi_paths_backward is not implemented yet");
}

}  // extern "C"
