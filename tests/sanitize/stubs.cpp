// Stand-ins for the kernel launchers and device-side objects capi.cpp calls, for the host-only
// sanitizer build.  Each records that it was reached and fails: argument validation is meant to
// return before any of them, and without a GPU check_device() does too.
#include <cstdio>

#include "../../cfgable-matmul_amd/csrc/internal.hpp"

int g_stub_calls = 0;

namespace cfgmm {
namespace {
int stub(const char *name) {
    ++g_stub_calls;
    return fail(CFGMM_EHIP, std::string("stub reached: ") + name);
}
}  // namespace

int launch_dgemm(hipStream_t, double *, const double *, const double *, size_t, size_t, size_t, size_t, size_t,
                 size_t, bool, bool) { return stub("launch_dgemm"); }
int launch_dgemm_non0tab(hipStream_t, double *, const double *, const double *, size_t, size_t, size_t, size_t,
                         size_t, size_t, bool, bool, const uint8_t *, size_t, size_t) {
    return stub("launch_dgemm_non0tab");
}
int launch_sgemm_non0tab(hipStream_t, float *, const float *, const float *, size_t, size_t, size_t, size_t, size_t,
                         size_t, bool, bool, const uint8_t *, size_t, size_t) {
    return stub("launch_sgemm_non0tab");
}
int launch_rho(hipStream_t, double *, const double *, const double *, const double *, size_t, size_t, size_t, size_t,
               size_t, const uint8_t *, size_t, size_t) { return stub("launch_rho"); }
int launch_rho_multi(hipStream_t, double *, size_t, const double *, const double *, size_t, size_t, const double *,
                     size_t, size_t, size_t, size_t, size_t, const uint8_t *, size_t, size_t, const RhoScratch *) {
    return stub("launch_rho_multi");
}
int launch_rho_noalloc(hipStream_t, double *, const double *, const double *, const double *, size_t, size_t, size_t,
                       size_t, size_t, const uint8_t *, size_t, size_t) { return stub("launch_rho_noalloc"); }
int launch_dgemm_ozaki(hipStream_t, double *, const double *, const double *, size_t, size_t, size_t, size_t, size_t,
                       size_t, bool, bool, int) { return stub("launch_dgemm_ozaki"); }
int launch_rho_batched(hipStream_t, const cfgmm_rho_problem *, size_t) { return stub("launch_rho_batched"); }
int launch_rho_ozaki(hipStream_t, double *, const double *, const double *, const double *, size_t, size_t, size_t,
                     size_t, size_t, const uint8_t *, size_t, size_t, int) { return stub("launch_rho_ozaki"); }
int launch_vxc(hipStream_t, double *, size_t, const double *, const double *, const double *, size_t, size_t, size_t,
               const uint8_t *, size_t, size_t) { return stub("launch_vxc"); }
int launch_gto(hipStream_t, double *, size_t, uint8_t *, size_t, const cfgmm_gto_shell *, size_t, const double *,
               const double *, size_t, const double *, size_t, size_t, double) { return stub("launch_gto"); }
int resident_create(hipStream_t, const double *, const double *, const double *, size_t, size_t, size_t, size_t,
                    size_t, const uint8_t *, size_t, size_t, int, ResidentRho **) { return stub("resident_create"); }
int resident_eval(hipStream_t, const ResidentRho *, double *, size_t, size_t, size_t) { return stub("resident_eval"); }
int resident_eval_batched(hipStream_t, const ResidentEvalReq *, size_t) { return stub("resident_eval_batched"); }
int resident_destroy(ResidentRho *) { return stub("resident_destroy"); }
size_t resident_bytes(const ResidentRho *) { return 0; }
size_t resident_ngrid(const ResidentRho *) { return 0; }
size_t resident_nset(const ResidentRho *) { return 0; }
size_t resident_blksize(const ResidentRho *) { return 1; }
int resident_device(const ResidentRho *) { return 0; }
int server_create(const cfgmm_server_config *, Server **) { return stub("server_create"); }
int server_destroy(Server *) { return stub("server_destroy"); }
int server_add_molecule(Server *, int, const double *, const double *, const uint8_t *, size_t, size_t, size_t,
                        size_t) { return stub("server_add_molecule"); }
int server_warm(Server *, const int *, size_t) { return stub("server_warm"); }
int server_run(Server *, const cfgmm_request *, size_t, cfgmm_request_result *) { return stub("server_run"); }
int server_stats(const Server *, cfgmm_server_stats *) { return stub("server_stats"); }
}  // namespace cfgmm
