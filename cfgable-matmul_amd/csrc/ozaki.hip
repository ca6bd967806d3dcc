// ozaki.hip -- split-precision (Ozaki-style) GEMM and rho.
#include "internal.hpp"

namespace cfgmm {

int launch_dgemm_ozaki(hipStream_t, double *, const double *, const double *, size_t, size_t, size_t, size_t, size_t,
                       size_t, bool, bool, int) {
    return fail(CFGMM_EINVAL, "split-precision GEMM is not built yet");
}

int launch_rho_ozaki(hipStream_t, double *, const double *, const double *, const double *, size_t, size_t, size_t,
                     size_t, size_t, const uint8_t *, size_t, size_t, int) {
    return fail(CFGMM_EINVAL, "split-precision rho is not built yet");
}

}  // namespace cfgmm
