// internal.hpp -- shared host/device helpers of libcfgmm (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>

#include "../../include/cfgmm.h"

namespace cfgmm {

// ---- error reporting ------------------------------------------------------
// Thread-local last-error string returned by cfgmm_last_error().
void set_error(const std::string &msg);
void clear_error();
int fail(int status, const std::string &msg);

// Returns CFGMM_OK or records the HIP error and returns CFGMM_EHIP / CFGMM_EOOM.
int hip_status(hipError_t e, const char *what);

#define CFGMM_HIP_TRY(expr)                                      \
    do {                                                         \
        int _st = ::cfgmm::hip_status((expr), #expr);            \
        if (_st != CFGMM_OK) return _st;                         \
    } while (0)

// ---- device workspace -------------------------------------------------------
// A growable device buffer owned by one stream.  Entry points that need
// scratch memory (split-K partials, compacted AO lists, rho partials) take it
// from the workspace bound to their stream; growing it synchronizes that
// stream first so that no in-flight kernel still reads the old buffer.
struct Workspace {
    std::mutex mu;  // held while this stream's buffers are synchronized, grown or handed out
    void *ptr = nullptr;
    size_t bytes = 0;
    void *host = nullptr;  // pinned host staging for tables copied to the device
    size_t host_bytes = 0;
    hipEvent_t host_done = nullptr;  // recorded after the last copy out of `host`
};
int workspace_get(hipStream_t stream, size_t bytes, void **out);
// The stream's flag word: beside the workspace (not in it), zeroed whenever the workspace is
// allocated, written only with epochs.  Valid after workspace_get on the stream, until the next
// workspace_get that grows it.
int *workspace_flag(hipStream_t stream);
// A new value per call, process-wide, never 0 (wraps after 2^32 calls): a flag word holding the
// current value was written during this call; any other value (0 or an earlier call's) was not.
unsigned workspace_epoch();
// Pinned host memory of the stream, for host tables that a call copies to the device.  It waits
// until the copies enqueued before the last host_staging_release have read it; the caller fills
// it, enqueues its copies, then calls host_staging_release.
int host_staging_get(hipStream_t stream, size_t bytes, void **out);
int host_staging_release(hipStream_t stream);

// ---- tracing -----------------------------------------------------------------
// trace_start returns nullptr unless tracing is on; trace_stop(nullptr) is a no-op.
hipEvent_t trace_start(hipStream_t stream);
void trace_stop(hipStream_t stream, const char *name, hipEvent_t start);
// While one exists on a thread, trace_start returns nullptr there (no events are created).
struct TraceOff {
    TraceOff();
    ~TraceOff();
};

// ---- tunables (cfgmm_tune_set / cfgmm_tune_get) ----------------------------------
int tune_get(int param);

// ---- device properties ------------------------------------------------------
int device_cu_count();  // compute units of the current device (cached per device)

// ---- small math --------------------------------------------------------------
__host__ __device__ inline size_t div_up(size_t a, size_t b) { return (a + b - 1) / b; }

// ---- kernel launchers (one per .hip file) -----------------------------------
// *flag = 0 unless every D_s (s < nset) equals its transpose on [0, nao)^2 (each pair of entries
// bitwise or within 2^-50 of the larger magnitude: rho::tile_pair_differs); the caller
// sets *flag beforehand (nothing is written when all are symmetric)
int launch_dm_symmetry(hipStream_t stream, const double *dms, size_t stride_ao, size_t nao, size_t nset, int *flag);
// Compacts, for each column block b < nblk, the rows x < nrows with
// tab[x*ldtab + b] != 0 into lists[b*nrows + *] (ascending) and counts[b].  With keys, also
// keys[b*maxchunks + c] = the plan bucket (of kPlanBuckets over nrows) of the midpoint of the
// c-th 64-entry chunk of the list, -1 past its end.
int launch_build_lists(hipStream_t stream, const uint8_t *tab, size_t ldtab, size_t nrows, size_t nblk, int *lists,
                       int *counts, int *keys = nullptr, size_t maxchunks = 0);
constexpr int kPlanBuckets = 1024;
// rho[s*stride_grid + g] += sum over the 64-row chunks of g's block of partial[chunk][s][g]
// (partial laid out [chunk][nset][ngrid]; chunk order fixed).
int launch_rho_reduce(hipStream_t stream, double *rhos, const double *partial, const int *counts, size_t stride_grid,
                      size_t ngrid, size_t nset, size_t blksize);
int launch_dgemm(hipStream_t stream, double *c, const double *a, const double *b, size_t m, size_t n, size_t k,
                 size_t lda, size_t ldb, size_t ldc, bool transa, bool transb);
int launch_dgemm_non0tab(hipStream_t stream, double *c, const double *a, const double *b, size_t m, size_t n,
                         size_t k, size_t lda, size_t ldb, size_t ldc, bool transa, bool transb, const uint8_t *tab,
                         size_t ldtab, size_t blksize);
// the same masked GEMM in f32 (v_mfma_f32_16x16x4f32)
int launch_sgemm_non0tab(hipStream_t stream, float *c, const float *a, const float *b, size_t m, size_t n, size_t k,
                         size_t lda, size_t ldb, size_t ldc, bool transa, bool transb, const uint8_t *tab,
                         size_t ldtab, size_t blksize);
int launch_rho(hipStream_t stream, double *rhos, const double *dms, const double *ao_lhs, const double *ao_rhs,
               size_t stride_ao, size_t stride_grid, size_t nao, size_t ngrid, size_t nset, const uint8_t *tab,
               size_t ldtab, size_t blksize);
// Caller-provided scratch of one rho evaluation (instead of the stream's workspace); `epoch`
// tags the symmetry flag the evaluation writes in it.
struct RhoScratch {
    void *ptr;
    size_t bytes;
    unsigned epoch;
    int *flag;  // a word beside ptr .. ptr + bytes holding 0 or earlier calls' epochs
};
int launch_rho_multi(hipStream_t stream, double *rhos, size_t rho_cs, const double *dms, const double *ao_lhs,
                     size_t lhs_cs, size_t ncomp, const double *ao_rhs, size_t stride_ao, size_t stride_grid,
                     size_t nao, size_t ngrid, size_t nset, const uint8_t *tab, size_t ldtab, size_t blksize,
                     const RhoScratch *scratch = nullptr);
// launch_rho with its scratch in the library's fixed device array (cfgmm_aodmao2rho_dev_noalloc)
int launch_rho_noalloc(hipStream_t stream, double *rhos, const double *dms, const double *ao_lhs,
                       const double *ao_rhs, size_t stride_ao, size_t stride_grid, size_t nao, size_t ngrid,
                       size_t nset, const uint8_t *tab, size_t ldtab, size_t blksize);
int launch_dgemm_ozaki(hipStream_t stream, double *c, const double *a, const double *b, size_t m, size_t n, size_t k,
                       size_t lda, size_t ldb, size_t ldc, bool transa, bool transb, int nslices);
int launch_rho_batched(hipStream_t stream, const cfgmm_rho_problem *problems, size_t nprob);
int launch_rho_ozaki(hipStream_t stream, double *rhos, const double *dms, const double *ao_lhs, const double *ao_rhs,
                     size_t stride_ao, size_t stride_grid, size_t nao, size_t ngrid, size_t nset, const uint8_t *tab,
                     size_t ldtab, size_t blksize, int nslices);

// Resident (per-grid-block compacted, int8-sliced) rho operands: rho_resident.hip.
struct ResidentRho;
int resident_create(hipStream_t stream, const double *dms, const double *ao_lhs, const double *ao_rhs,
                    size_t stride_ao, size_t stride_grid, size_t nao, size_t ngrid, size_t nset, const uint8_t *tab,
                    size_t ldtab, size_t blksize, int nslices, ResidentRho **out);
int resident_eval(hipStream_t stream, const ResidentRho *h, double *rhos, size_t stride_rho, size_t ngrid_used,
                  size_t nset_used);
int resident_destroy(ResidentRho *h);
struct ResidentEvalReq {
    const ResidentRho *h;
    double *rhos;
    size_t stride_rho, ngrid, nset;
};
int resident_eval_batched(hipStream_t stream, const ResidentEvalReq *reqs, size_t n);
size_t resident_bytes(const ResidentRho *h);
// Device bytes a handle over this mask (host copy, nao x ldtab) takes at most; `symmetric`: every
// D_s is symmetric and lhs is rhs (the upper-triangle form).
size_t resident_bytes_bound(const uint8_t *tab, size_t ldtab, size_t nao, size_t nblk, size_t nset, size_t blksize,
                            int nslices, bool symmetric);
size_t resident_ngrid(const ResidentRho *h);
size_t resident_nset(const ResidentRho *h);
size_t resident_blksize(const ResidentRho *h);
int resident_device(const ResidentRho *h);

// The AO-grid-AO contraction after rho (vxc.hip).
int launch_vxc(hipStream_t stream, double *V, size_t ldv, const double *ao_lhs, const double *ao_rhs, const double *w,
               size_t stride_grid, size_t nao, size_t ngrid, const uint8_t *tab, size_t ldtab, size_t blksize);
// cfgmm_eval_gto_dev (shells, exps, coefs: host arrays, validated by the caller)
int launch_gto(hipStream_t stream, double *phi, size_t stride_grid, uint8_t *tab, size_t ldtab,
               const cfgmm_gto_shell *shells, size_t nshell, const double *exps, const double *coefs,
               size_t nprim_total, const double *coords, size_t ngrid, size_t blksize, double cutoff);

// Serving helpers (serve_kernels.hip) and the request loop (server.cpp).
int launch_fill_uniform(hipStream_t stream, double *x, size_t n, unsigned long long seed);
int launch_fill_zero(hipStream_t stream, void *x, size_t bytes);
int launch_checksum(hipStream_t stream, const double *x, size_t rows, size_t cols, size_t ld,
                    unsigned long long *out);
struct Server;
int server_create(const cfgmm_server_config *cfg, Server **out);
int server_destroy(Server *sv);
int server_add_molecule(Server *sv, int id, const double *dms, const double *phi, const uint8_t *tab, size_t nao,
                        size_t ngrid, size_t nset, size_t blksize);
int server_warm(Server *sv, const int *ids, size_t n);
int server_run(Server *sv, const cfgmm_request *reqs, size_t n, cfgmm_request_result *res);
int server_stats(const Server *sv, cfgmm_server_stats *out);

// ---- MFMA fragment types -------------------------------------------------------
typedef double f64x4 __attribute__((ext_vector_type(4)));

}  // namespace cfgmm
