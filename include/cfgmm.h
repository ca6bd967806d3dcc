/*
 * cfgmm.h -- C ABI of the MI355X (gfx950) implementation of the two hot paths
 * of ajz34/cfgable-matmul:
 *
 *   1. row-major f64 GEMM   C += op(A) op(B)
 *   2. non0tab-masked AO-DM-AO density contraction
 *        rho[s][g] += sum_{u,v : tab[u][g/B] && tab[v][g/B]} lhs[u][g] D_s[u][v] rhs[v][g]
 *      and the masked GEMM that is its unfused first half.
 *
 * Conventions (identical to the reference):
 *   - row-major; op(A)(i,p) = transa ? a[p*lda+i] : a[i*lda+p],
 *                op(B)(p,j) = transb ? b[j*ldb+p] : b[p*ldb+j]
 *     (src/impl_matmul.rs:198-201, 83-85, 107, 217-220);
 *   - every entry point ACCUMULATES into its output (C += ..., rho += ...);
 *     there is no alpha/beta (src/impl_matmul.rs:33-46, src/impl_aodmao2rho.rs:77-92);
 *   - masks are one byte per entry, 0 = false, anything else = true (Rust
 *     `bool` is one byte); tab[p*ldtab + j/blksize] for the masked GEMM and
 *     tab[u*ldtab + g/blksize] for rho; the mask is authoritative: values under
 *     a false bit are never read;
 *   - only i<m, j<n (resp. g<ngrid) are written; padding columns of C / rho
 *     beyond the logical extent are untouched.
 *
 * The `*_len` arguments are the element counts of the caller's slices (Rust
 * `&[T]::len()`); an extent that the arguments would index past returns
 * CFGMM_EBOUNDS, where the reference would panic with an index-out-of-bounds.
 *
 * Host-pointer entry points are synchronous drop-ins for the reference
 * functions: they copy inputs to HBM, run the HIP kernels on a per-thread
 * stream, and return after the result is back in the caller's buffer.
 * `_dev` entry points take device pointers and a hipStream_t (passed as
 * void*), enqueue the work and return without synchronizing.
 *
 * All functions return a cfgmm status; cfgmm_last_error() gives a message for
 * the last failing call on the calling thread.
 */
#ifndef CFGMM_H
#define CFGMM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum cfgmm_status {
    CFGMM_OK = 0,
    CFGMM_EINVAL = 1,             /* bad argument (blksize, nslices, null pointer) */
    CFGMM_EINVAL_TRANSB_MASK = 2, /* transb with a mask: unimplemented!() at src/impl_matmul_non0tab.rs:181-183,228-230 */
    CFGMM_EINVAL_DIMS = 3,        /* leading dimension / stride smaller than the row length */
    CFGMM_EOOM = 4,               /* device or pinned-host allocation failed */
    CFGMM_EHIP = 5,               /* HIP runtime error */
    CFGMM_ENODEV = 6,             /* no gfx950 device visible */
    CFGMM_EBOUNDS = 7             /* a slice is shorter than the extent the arguments index */
};

/* Message for the last failing call on this thread ("" if none). */
const char *cfgmm_last_error(void);
/* Library version as 10000*major + 100*minor + patch. */
int cfgmm_version(void);

/* ---- precision modes (split-precision variants) -------------------------- */
enum cfgmm_precision {
    CFGMM_PREC_FP64 = 0,        /* plain FP64 MFMA */
    CFGMM_PREC_OZAKI_INT8 = 1   /* Ozaki-scheme splitting of f64 into `nslices` int8 slices (per-row /
                                   per-column power-of-two scaling), slice products on the int8 MFMA with
                                   exact int32 accumulation, FP64 recombination; nslices in [2, 6].  GEMM and
                                   resident rho: balanced radix-254 digits, about 8 bits per slice (5 slices
                                   within 1e-10 on the tested workloads); per-call rho: 7-bit digits (6
                                   slices for 1e-10) */
};

/* ======================================================================== */
/* 1. row-major f64 GEMM                                                    */
/* ======================================================================== */

/* Replaces MatmulLoops::<f64, MC, KC, NC, MR, NR_LANE, LANE>::matmul_loop_parallel_mnk_pack_a
 * (src/impl_matmul.rs:155-231).  The CPU blocking const generics have no GPU
 * meaning and are not arguments. */
int cfgmm_matmul_loop_parallel_mnk_pack_a(double *c, size_t c_len, const double *a, size_t a_len, const double *b,
                                          size_t b_len, size_t m, size_t n, size_t k, size_t lda, size_t ldb,
                                          size_t ldc, int transa, int transb);

/* The README's name for the same function, with an extra MB const generic
 * (README.md:22-31); identical semantics to the above. */
int cfgmm_matmul_loop_macro_mb(double *c, size_t c_len, const double *a, size_t a_len, const double *b, size_t b_len,
                               size_t m, size_t n, size_t k, size_t lda, size_t ldb, size_t ldc, int transa,
                               int transb);

/* Replaces matmul_anyway_full (src/impl_matmul.rs:234-248). */
int cfgmm_matmul_anyway_full(double *c, size_t c_len, const double *a, size_t a_len, const double *b, size_t b_len,
                             size_t m, size_t n, size_t k, size_t lda, size_t ldb, size_t ldc, int transa,
                             int transb);

/* Device-pointer GEMM, enqueued on `stream` (hipStream_t; NULL = null stream). */
int cfgmm_dgemm_dev(void *stream, double *c, const double *a, const double *b, size_t m, size_t n, size_t k,
                    size_t lda, size_t ldb, size_t ldc, int transa, int transb);

/* Split-precision GEMM (no counterpart in the reference): same contract as
 * cfgmm_dgemm_dev, computed with `precision` / `nslices`. */
int cfgmm_dgemm_ex_dev(void *stream, double *c, const double *a, const double *b, size_t m, size_t n, size_t k,
                       size_t lda, size_t ldb, size_t ldc, int transa, int transb, int precision, int nslices);

/* ======================================================================== */
/* 2a. non0tab-masked GEMM                                                  */
/* ======================================================================== */

/* Replaces matmul_loop_parallel_mnk_pack_a_non0tab::<BLKSIZE>
 * (src/impl_matmul_non0tab.rs:210-305):
 *   C(i,j) += sum_{p<k : tab[p*ldtab + j/blksize]} op(A)(i,p) B(p,j).
 * transb != 0 returns CFGMM_EINVAL_TRANSB_MASK; blksize must be a positive
 * multiple of 16 (the reference's BLKSIZE % (NR_LANE*LANE) == 0). */
int cfgmm_matmul_loop_parallel_mnk_pack_a_non0tab(double *c, size_t c_len, const double *a, size_t a_len,
                                                  const double *b, size_t b_len, size_t m, size_t n, size_t k,
                                                  size_t lda, size_t ldb, size_t ldc, int transa, int transb,
                                                  const uint8_t *tab, size_t tab_len, size_t ldtab, size_t blksize);

/* Replaces matmul_anyway_full_non0tab (src/impl_matmul_non0tab.rs:308-326): blksize = 48. */
int cfgmm_matmul_anyway_full_non0tab(double *c, size_t c_len, const double *a, size_t a_len, const double *b,
                                     size_t b_len, size_t m, size_t n, size_t k, size_t lda, size_t ldb, size_t ldc,
                                     int transa, int transb, const uint8_t *tab, size_t tab_len, size_t ldtab);

/* Device-pointer masked GEMM. */
int cfgmm_dgemm_non0tab_dev(void *stream, double *c, const double *a, const double *b, size_t m, size_t n, size_t k,
                            size_t lda, size_t ldb, size_t ldc, int transa, int transb, const uint8_t *tab,
                            size_t ldtab, size_t blksize);

/* Masked GEMM with transb allowed (no counterpart in the reference, which leaves
 * transb unimplemented, src/impl_matmul_non0tab.rs:181-183,228-230):
 *   C(i,j) += sum_{p : tab[p*ldtab + j/blksize]} op(A)(i,p) * op(B)(p,j),
 * op(B)(p,j) = transb ? b[j*ldb + p] : b[p*ldb + j]. */
int cfgmm_dgemm_non0tab_ex_dev(void *stream, double *c, const double *a, const double *b, size_t m, size_t n,
                               size_t k, size_t lda, size_t ldb, size_t ldc, int transa, int transb,
                               const uint8_t *tab, size_t ldtab, size_t blksize);

/* The same masked GEMM with T = f32: the reference's element type is generic
 * (src/impl_matmul_non0tab.rs, `T: TySimd`), but only f64 is implemented there
 * (src/structs.rs:19-21).  Accumulates in f32 (v_mfma_f32_16x16x4f32); transb allowed. */
int cfgmm_sgemm_non0tab_dev(void *stream, float *c, const float *a, const float *b, size_t m, size_t n, size_t k,
                            size_t lda, size_t ldb, size_t ldc, int transa, int transb, const uint8_t *tab,
                            size_t ldtab, size_t blksize);

/* ======================================================================== */
/* 2b. fused masked AO-DM-AO -> rho                                         */
/* ======================================================================== */

/* Replaces aodmao2rho_loop_parallel::<BLKSIZE> (src/impl_aodmao2rho.rs:171-244):
 *   rhos[s*stride_grid + g] += sum_{u,v on} ao_lhs[u*stride_grid+g]
 *                              * dms[s*stride_ao*stride_ao + u*stride_ao + v]
 *                              * ao_rhs[v*stride_grid+g]
 * for s < nset, g < ngrid, with u,v < nao "on" when tab[x*ldtab + g/blksize].
 * `strides = [stride_ao, stride_grid]`, `shape = [nao, ngrid]` in the reference. */
int cfgmm_aodmao2rho_loop_parallel(double *rhos, size_t rhos_len, const double *dms, size_t dms_len,
                                   const double *ao_lhs, size_t ao_lhs_len, const double *ao_rhs, size_t ao_rhs_len,
                                   size_t stride_ao, size_t stride_grid, size_t nao, size_t ngrid, size_t nset,
                                   const uint8_t *tab, size_t tab_len, size_t ldtab, size_t blksize);

/* Replaces aodmao2rho_anyway (src/impl_aodmao2rho.rs:247-261): blksize = 48. */
int cfgmm_aodmao2rho_anyway(double *rhos, size_t rhos_len, const double *dms, size_t dms_len, const double *ao_lhs,
                            size_t ao_lhs_len, const double *ao_rhs, size_t ao_rhs_len, size_t stride_ao,
                            size_t stride_grid, size_t nao, size_t ngrid, size_t nset, const uint8_t *tab,
                            size_t tab_len, size_t ldtab);

/* Device-pointer rho.  When ao_lhs == ao_rhs (one pointer) and every D_s equals its transpose
 * (tested on the device in the call, no host synchronisation; entries equal bitwise or within
 * 2^-50 of the larger magnitude, as a D formed by a general product C C^T is), each block's upper
 * triangle is contracted and the part past the diagonal doubled: the same sum within rounding,
 * half the work.  CFGMM_TUNE_RHO_SYMMETRIC = 1 turns this off. */
int cfgmm_aodmao2rho_dev(void *stream, double *rhos, const double *dms, const double *ao_lhs, const double *ao_rhs,
                         size_t stride_ao, size_t stride_grid, size_t nao, size_t ngrid, size_t nset,
                         const uint8_t *tab, size_t ldtab, size_t blksize);

/* Allocation-free form of cfgmm_aodmao2rho_dev.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and result as cfgmm_aodmao2rho_dev
 * (FP64, the same kernels); it calls no hipMalloc or hipFree and creates no events.  Its
 * scratch -- the compacted AO lists, the task plan and the chunk partials -- is a device array
 * of CFGMM_RHO_NOALLOC_SCRATCH_BYTES in the library (its first 256 bytes a flag word), one per
 * GPU, held until the process ends.
 * A problem needs about 4*nblk*(nao + 3*ceil(nao/64)) + 8*ceil(nao/64)*nset*ngrid bytes
 * (nblk = ceil(ngrid/blksize); 16 MB for nao 2228, ngrid 16384, nset 3); a larger one returns
 * CFGMM_EINVAL with the needed size in cfgmm_last_error().  Every call on a GPU shares that
 * GPU's array: calls that overlap in time on one GPU (other streams or threads) race, and the
 * caller must serialize them (one stream, or events between them). */
#define CFGMM_RHO_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int cfgmm_aodmao2rho_dev_noalloc(void *stream, double *rhos, const double *dms, const double *ao_lhs,
                                 const double *ao_rhs, size_t stride_ao, size_t stride_grid, size_t nao, size_t ngrid,
                                 size_t nset, const uint8_t *tab, size_t ldtab, size_t blksize);

/* Split-precision rho (no counterpart in the reference).  Per call the split form gathers and
 * slices D_s[L_b, L_b] and phi for every grid block before its int8 products, and that costs more
 * than the FP64 evaluation saves: on the tested workloads it runs 4-7x slower than FP64 (config of
 * 2228 AOs x 16384 points: 4.2 ms at 6 slices against 0.6 ms).  It is there for accuracy studies
 * and as the split path of the batched call; for speed, evaluate in FP64 per call, or keep a
 * resident handle (2d) when the same (D, phi, mask) is evaluated repeatedly (about 2x FP64 per
 * evaluation).  6 slices keep 1e-10; 5 do not. */
int cfgmm_aodmao2rho_ex_dev(void *stream, double *rhos, const double *dms, const double *ao_lhs,
                            const double *ao_rhs, size_t stride_ao, size_t stride_grid, size_t nao, size_t ngrid,
                            size_t nset, const uint8_t *tab, size_t ldtab, size_t blksize, int precision,
                            int nslices);

/* Several left-hand factors over the same density matrices, rhs and mask, in one pass
 * (no counterpart in the reference; a GGA density needs phi and its three gradient
 * components as lhs with phi as rhs):
 *   rho_c[s*stride_grid + g] += sum_{u,v in L_b} lhs_c[u][g] D_s[u][v] rhs[v][g],  c < ncomp,
 * rho_c = rhos + c*rho_comp_stride, lhs_c = ao_lhs + c*lhs_comp_stride (doubles).  T = D.rhs
 * is formed once; each component is bitwise the result of cfgmm_aodmao2rho_dev with its lhs.
 * ncomp <= 16; the components' rho ranges must not overlap. */
int cfgmm_aodmao2rho_multi_dev(void *stream, double *rhos, size_t rho_comp_stride, const double *dms,
                               const double *ao_lhs, size_t lhs_comp_stride, size_t ncomp, const double *ao_rhs,
                               size_t stride_ao, size_t stride_grid, size_t nao, size_t ngrid, size_t nset,
                               const uint8_t *tab, size_t ldtab, size_t blksize);

/* The contraction that follows rho in a DFT step (no counterpart in the reference):
 *   V[u*ldv + v] += sum_b sum_{g in block b} [u in L_b][v in L_b] lhs[u][g] * w[g] * rhs[v][g],
 * with the mask, strides and blocks of cfgmm_aodmao2rho_dev (lhs, rhs: nao x stride_grid;
 * w: ngrid weights, or NULL for all ones; V: nao x ldv).  Deterministic (per-block products
 * summed in block order); uses sum_b kpad_b^2 doubles of workspace, kpad_b = |L_b| rounded
 * up to 64.  With ao_lhs == ao_rhs only the upper tiles of each block product are computed
 * (nao^2 more doubles of workspace). */
int cfgmm_aogridao2mat_dev(void *stream, double *V, size_t ldv, const double *ao_lhs, const double *ao_rhs,
                           const double *w, size_t stride_grid, size_t nao, size_t ngrid, const uint8_t *tab,
                           size_t ldtab, size_t blksize);

/* AO tabulation on the grid, the step before rho (no counterpart in the reference, which
 * takes phi and the mask as inputs; the mask plays the role of PySCF's non0tab, README.md:49).
 * Cartesian GTO shells: shell s (host array, nshell entries) centred at R = center, angular
 * momentum l <= 4, contracted over exps/coefs[prim_off .. prim_off + nprim) (host arrays, the
 * coefficients as given: any normalisation is the caller's), owning AOs ao_off .. ao_off +
 * (l+1)(l+2)/2 - 1 in the order lx = l..0, then ly = l-lx..0, lz = l-lx-ly (PySCF's Cartesian
 * order):  phi[(ao_off+c)*stride_grid + g] = dx^lx dy^ly dz^lz sum_k coefs[k] exp(-exps[k] r^2),
 * d = coords[g] - R (coords: device, ngrid x 3).  The mask: tab[(ao_off+c)*ldtab + b] = 1 when
 * max over the points of block b and the components of the shell of |phi| > cutoff, else 0, and
 * then phi is zeroed on that block.  Rows of AOs owned by no shell, and columns g >= ngrid, are
 * left untouched.  Asynchronous on the stream: the host tables (shells, exps, coefs) are copied
 * into pinned staging before the call returns, so the caller may reuse them at once. */
typedef struct cfgmm_gto_shell {
    double center[3];
    int l, nprim, prim_off, ao_off;
} cfgmm_gto_shell;
int cfgmm_eval_gto_dev(void *stream, double *phi, size_t stride_grid, uint8_t *tab, size_t ldtab,
                       const cfgmm_gto_shell *shells, size_t nshell, const double *exps, const double *coefs,
                       size_t nprim_total, const double *coords, size_t ngrid, size_t nao, size_t blksize,
                       double cutoff);

/* ======================================================================== */
/* 2c. batched rho (no counterpart in the reference)                        */
/* ======================================================================== */

/* One problem of a batch: the arguments of cfgmm_aodmao2rho_dev, device pointers. */
typedef struct cfgmm_rho_problem {
    double *rhos;
    const double *dms;
    const double *ao_lhs;
    const double *ao_rhs;
    size_t stride_ao, stride_grid, nao, ngrid, nset;
    const uint8_t *tab;
    size_t ldtab, blksize;
} cfgmm_rho_problem;

/* Evaluates every problem of the batch, each with exactly the contract of
 * cfgmm_aodmao2rho_dev.  FP64: one persistent launch drains a device-built
 * task queue of all (problem, grid block, 64-row chunk, density-matrix group)
 * tasks; split precision: the problems are enqueued one after another.
 * Results are bitwise those of separate cfgmm_aodmao2rho_ex_dev calls. */
int cfgmm_aodmao2rho_batched_dev(void *stream, const cfgmm_rho_problem *problems, size_t nproblems, int precision,
                                 int nslices);

/* ======================================================================== */
/* 2d. resident rho operands (no counterpart in the reference)              */
/* ======================================================================== */

/* Operands of one masked rho problem kept on the device in the form its
 * evaluations read: for every grid block b, the density-matrix sub-blocks
 * D_s[L_b, L_b] and the block's AO values compacted to contiguous tiles and
 * cut into int8 slices (split precision, `nslices` in [2, 6]), with
 * power-of-two scales per block row and grid point.  Creation gathers and
 * slices once (arguments as cfgmm_aodmao2rho_ex_dev, device pointers; the
 * inputs may be freed afterwards); each evaluation then streams the tiles
 * without gathers.  For callers that evaluate the same (D, phi, mask) many
 * times -- a serving loop that keeps a molecule resident.  Memory is
 * sum_b (nset*nslices*kpad_b^2 + blksize*kpad_b*(nslices + 8)) bytes plus small
 * tables, kpad_b = |L_b| rounded up to 64 (cfgmm_rho_resident_bytes).  The slices are
 * balanced digits of about 8 bits: 5 keep rho within 1e-10 on the tested workloads.  With
 * ao_lhs == ao_rhs and every D_s symmetric (tested at creation, as for cfgmm_aodmao2rho_dev)
 * only the upper-triangle tiles are kept (about half the bytes and time), for blocks of up to
 * 8192 active AOs.
 * precision must be CFGMM_PREC_OZAKI_INT8. */
typedef struct cfgmm_rho_resident cfgmm_rho_resident;
int cfgmm_rho_resident_create(void *stream, const double *dms, const double *ao_lhs, const double *ao_rhs,
                              size_t stride_ao, size_t stride_grid, size_t nao, size_t ngrid, size_t nset,
                              const uint8_t *tab, size_t ldtab, size_t blksize, int precision, int nslices,
                              cfgmm_rho_resident **out);
/* rhos[s*stride_rho + g] += rho_s(g) for the leading ngrid_used grid points
 * (ngrid_used == ngrid, or a multiple of blksize below it) and the leading
 * nset_used density matrices.  Evaluations of one handle on different
 * streams may run concurrently (each uses its stream's workspace). */
int cfgmm_rho_resident_eval(void *stream, const cfgmm_rho_resident *h, double *rhos, size_t stride_rho,
                            size_t ngrid_used, size_t nset_used);
/* Several evaluations (of one or many handles) in one launch per (nslices, tile width) group;
 * each entry with the contract of cfgmm_rho_resident_eval and bitwise its result.  Entries
 * must not write overlapping rho ranges.  With up to five slices the launch is persistent: the
 * first call of a batch deals every item to the device's workgroups on the host and uploads that
 * plan (about 128 bytes per item, allocated with hipMalloc); calls with the same handles, ranges
 * and stream workspace reuse it (the plans of the last four batches are kept, process-wide). */
typedef struct cfgmm_rho_resident_job {
    const cfgmm_rho_resident *h;
    double *rhos;
    size_t stride_rho, ngrid_used, nset_used;
} cfgmm_rho_resident_job;
int cfgmm_rho_resident_eval_batched(void *stream, const cfgmm_rho_resident_job *evals, size_t n);
int cfgmm_rho_resident_bytes(const cfgmm_rho_resident *h, size_t *bytes);
/* Frees the device memory; no evaluation of the handle may still be running. */
int cfgmm_rho_resident_destroy(cfgmm_rho_resident *h);

/* ======================================================================== */
/* 2e. serving loop (no counterpart in the reference)                       */
/* ======================================================================== */

/* A request loop on the current device: rho over catalogued molecules and
 * DGEMMs with device-generated operands, admitted as slots (streams) and HBM
 * budget free up, FP64 rho requests queued together fused into one batched
 * launch, molecules kept resident across requests and evicted least recently
 * used when memory is short.  Not thread-safe: one thread drives a server. */
typedef struct cfgmm_server cfgmm_server;

typedef struct cfgmm_server_config {
    size_t hbm_budget_bytes; /* device memory for resident molecules + in-flight requests */
    int slots;               /* concurrent streams, 1..16 */
    int rho_precision;       /* CFGMM_PREC_FP64 (per-call kernels over resident inputs) or
                                CFGMM_PREC_OZAKI_INT8 (resident per-block tiles, see 2d; a molecule whose
                                tiles could take more than split_tile_cap_bytes is served by the FP64
                                kernels instead, from its resident D, phi and mask) */
    int rho_nslices;         /* split precision: slices, 2..6 */
    int max_batch;           /* rho requests dispatched together on one slot, >= 1 */
    size_t split_tile_cap_bytes; /* split precision: largest tile set built for one molecule;
                                    0 = a quarter of hbm_budget_bytes */
} cfgmm_server_config;

enum cfgmm_request_kind { CFGMM_REQ_RHO = 0, CFGMM_REQ_GEMM = 1 };

typedef struct cfgmm_request {
    int kind;
    int molecule;            /* rho: catalogue id */
    size_t ngrid, nset;      /* rho: leading grid points (all, or a multiple of blksize) and density matrices */
    size_t m, n, k;          /* GEMM: C(m x n) = A(m x k) B(k x n), row-major, C starting at zero */
    unsigned long long seed; /* GEMM: A[i] = U(seed' = 2*seed), B[i] = U(2*seed + 1), where
                                U(t)[i] = (splitmix64((t << 32) + i) >> 11) * 2^-52 - 1 */
    double arrival_s;        /* arrival, seconds after the run starts */
    double *out;             /* optional host buffer for the result (rho: nset x ngrid, GEMM: m x n), or NULL */
} cfgmm_request;

typedef struct cfgmm_request_result {
    double admit_s;               /* dispatched to a slot */
    double done_s;                /* completion observed (latency = done_s - arrival_s) */
    unsigned long long checksum;  /* sum over i of splitmix64(bits(x[i]) ^ splitmix64(i)) mod 2^64 */
    int status;                   /* CFGMM_OK */
    int batch;                    /* dispatch the request rode in */
    int batch_size;
} cfgmm_request_result;

typedef struct cfgmm_server_stats {
    unsigned long long rho_requests, gemm_requests, batches, hits, misses, uploads, evictions;
    size_t peak_bytes, held_bytes;
    unsigned long long rho_fp64_requests; /* rho requests run by the FP64 kernels (all of them in FP64) */
} cfgmm_server_stats;

int cfgmm_server_create(const cfgmm_server_config *cfg, cfgmm_server **out);
/* Registers a molecule: host arrays that must stay valid while the server lives
 * (dms nset x nao x nao, phi nao x ngrid, tab nao x ceil(ngrid/blksize), row-major). */
int cfgmm_server_add_molecule(cfgmm_server *s, int id, const double *dms, const double *phi, const uint8_t *tab,
                              size_t nao, size_t ngrid, size_t nset, size_t blksize);
/* Makes the listed molecules resident, in order, while they fit without evicting. */
int cfgmm_server_warm(cfgmm_server *s, const int *ids, size_t n);
/* Replays the requests in real time (open loop) and returns when all are done. */
int cfgmm_server_run(cfgmm_server *s, const cfgmm_request *reqs, size_t n, cfgmm_request_result *results);
int cfgmm_server_get_stats(const cfgmm_server *s, cfgmm_server_stats *out);
int cfgmm_server_destroy(cfgmm_server *s);

/* ---- tunables (set by the autotuner; process-wide) ------------------------ */
enum cfgmm_tunable {
    CFGMM_TUNE_BATCH_WGS_PER_CU = 0, /* batched rho: persistent workgroups per CU, 1..8 (default 2) */
    CFGMM_TUNE_BATCH_MAX_SETS = 1,   /* batched rho: density matrices per task, 1, 2 or 4 (default 1) */
    CFGMM_TUNE_DGEMM_GROUP_M = 2,    /* DGEMM: tile rows per raster group, 1..64 (default 4) */
    CFGMM_TUNE_RHO_MAX_SETS = 3,     /* FP64 rho: density matrices per workgroup, 1..4 (default 1) */
    CFGMM_TUNE_DGEMM_SMALL_TILE = 4, /* DGEMM with fewer 128x128 tiles than slots: tile edge 64 or 128, 0 = heuristic */
    CFGMM_TUNE_DGEMM_SMALL_SPLITS = 5, /* ... and K ranges per tile, 1..32, 0 = heuristic */
    CFGMM_TUNE_RESIDENT_DEPTH = 6,   /* resident rho: k steps of A prefetch, 1 or 2 (default 1) */
    CFGMM_TUNE_RESIDENT_COLS = 7,    /* resident rho: grid points per workgroup, 16/32/48/64, 0 = widest
                                        dividing blksize (read at creation) */
    CFGMM_TUNE_RHO_ORDER = 8,        /* FP64 rho: 0 = chunks ordered by the AO range of their rows
                                        (default), 1 = block-major */
    CFGMM_TUNE_RHO_SYMMETRIC = 9,    /* FP64 rho: 0 = when ao_lhs == ao_rhs, test every D_s for
                                        D_s == D_s^T on the device and, if so, contract only the
                                        upper triangle of each block (default); 1 = never.
                                        Also the Vxc contraction: with ao_lhs == ao_rhs only
                                        the upper tiles of each block product are computed */
    CFGMM_TUNE_DGEMM_KGROUPS = 10,   /* DGEMM too small for 128x128 tiles, A not transposed and B
                                        transposed: 0 = one workgroup per 64-row tile splitting K
                                        among four groups of waves when the tiles then cover 3/4 or
                                        more of the CUs and k <= 1024 (default), 1 = never (K split
                                        over workgroups), 2 = whenever the tiles cover 3/4 of the
                                        CUs, any k */
    CFGMM_TUNE_MASKED_LDS = 11,      /* FP64 masked GEMM with B not transposed and n even: 0 = op(A)
                                        transposed into the workspace, slices copied straight into
                                        LDS, 128-row tiles (default); 2 = the same with 256-row
                                        tiles (blksize not a multiple of 64); 1 = register-staged
                                        slices, as for f32, transposed B and odd n */
    CFGMM_TUNE_RESIDENT_PERSISTENT = 12, /* resident rho, whole evaluations: 0 = two workgroups per CU
                                        each walking a fixed list of items as one pipeline
                                        (default), 1 = one workgroup per item */
    CFGMM_TUNE_COUNT = 13
};
int cfgmm_tune_set(int param, int value);
int cfgmm_tune_get(int param, int *value);

/* ======================================================================== */
/* device plumbing for callers without their own HIP binding                */
/* ======================================================================== */
int cfgmm_device_count(int *count);
int cfgmm_device_malloc(void **ptr, size_t bytes);
int cfgmm_device_free(void *ptr);
int cfgmm_memcpy_htod(void *dst, const void *src, size_t bytes);
int cfgmm_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int cfgmm_memcpy_htod_async(void *dst, const void *src, size_t bytes, void *stream);
int cfgmm_memcpy_dtoh_async(void *dst, const void *src, size_t bytes, void *stream);
int cfgmm_memcpy_dtod_async(void *dst, const void *src, size_t bytes, void *stream);
int cfgmm_memset_dev(void *dst, int value, size_t bytes);
int cfgmm_mem_get_info(size_t *free_bytes, size_t *total_bytes);
int cfgmm_stream_create(void **stream);
int cfgmm_stream_destroy(void *stream);
int cfgmm_stream_synchronize(void *stream);
int cfgmm_device_synchronize(void);
int cfgmm_event_create(void **event);
int cfgmm_event_destroy(void *event);
int cfgmm_event_record(void *event, void *stream);
int cfgmm_event_synchronize(void *event);
int cfgmm_event_query(void *event); /* CFGMM_OK when complete, 1 when pending */
int cfgmm_event_elapsed_ms(float *ms, void *start, void *end);

/* ---- tracing: hipEvent timing of the library's own kernels --------------- */
/* When enabled, every kernel launch of the library is bracketed by two
 * hipEvents recorded on the stream it is launched on.  cfgmm_trace_read waits
 * for the recorded events of kernel `name` (e.g. "dgemm_kernel",
 * "rho_chunk_kernel") and returns their summed duration and launch count
 * since the last reset.  Tracing adds two event records per launch. */
int cfgmm_trace_enable(int on);
int cfgmm_trace_reset(void);
int cfgmm_trace_read(const char *name, double *total_ms, long long *count);

#ifdef __cplusplus
}
#endif

#endif /* CFGMM_H */
