/*
 * cfgmm_oracle.c -- CPU restatement of the reference algorithm.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle and the timed CPU
 * baseline for the two hot paths of ajz34/cfgable-matmul.  Only the tests,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it; the
 * product (libcfgmm.so) never links or calls it.
 *
 * It restates, in plain C with GCC vector extensions, the BLIS-style loop nest
 * of the reference crate with the same blocking constants as its f64 wrappers
 * (MC=252, KC=512, NC=240, MR=14, NR_LANE=2, LANE=8, BLKSIZE=48 by default):
 *
 *   oracle_matmul_loop_parallel_mnk_pack_a     src/impl_matmul.rs:155-231
 *     pack A                                    src/impl_matmul.rs:184-205
 *     task order (m fastest, then n, then k)    src/impl_matmul.rs:207-230
 *     matmul_loop_2nd_nr_pack_b                 src/impl_matmul.rs:120-153
 *     pack_b_no_trans / pack_b_trans            src/impl_matmul.rs:77-117
 *     matmul_loop_1st_mr (locked C tile)        src/impl_matmul.rs:10-74
 *     microkernel (c += a*b, fp-contract)       src/structs.rs:23-46
 *   oracle_matmul_loop_parallel_mnk_pack_a_non0tab
 *                                               src/impl_matmul_non0tab.rs:210-305
 *     get_indices_from_non0tab                  src/impl_matmul_non0tab.rs:119-129
 *     pack_b_no_trans_non0tab                   src/impl_matmul_non0tab.rs:132-161
 *     matmul_loop_2nd_nr_pack_b_non0tab         src/impl_matmul_non0tab.rs:164-208
 *     matmul_loop_1st_mr_mask_k_for_b           src/impl_matmul_non0tab.rs:53-116
 *     microkernel_mask_k_for_b (true FMA)       src/impl_matmul_non0tab.rs:21-41
 *   oracle_aodmao2rho_loop_parallel             src/impl_aodmao2rho.rs:171-244
 *     aodmao2rho_loop_2nd                       src/impl_aodmao2rho.rs:97-169
 *     pack_dm_in_aodmao2rho                     src/impl_aodmao2rho.rs:10-29
 *     aodmao2rho_loop_1st                       src/impl_aodmao2rho.rs:33-95
 *
 * Rayon's pool with an atomic task counter becomes a pthread pool pulling task
 * ids from a C11 atomic counter; the per-tile Mutex<()> becomes a per-tile
 * pthread mutex.  As in the reference, the order in which K blocks (GEMM) or
 * v blocks (rho) add into one output tile follows lock acquisition, so results
 * are reproducible only to rounding.
 *
 * The Rust `bool` mask is passed as one byte per entry (0 = false, else true).
 * The reference's masked drivers need NC % BLKSIZE == 0 (asserted for rho,
 * src/impl_aodmao2rho.rs:188) with NC = 240 and BLKSIZE = 48; for other
 * BLKSIZE values this restatement uses the largest multiple of BLKSIZE not
 * above 240 (or BLKSIZE itself), which changes only the summation blocking,
 * never the terms summed.
 *
 * Parity pins for every function here live in tests/test_oracle.py: numpy
 * float64 matmul/einsum (library routines), exact integer-valued closed forms,
 * brute-force loops on tiny shapes, the masked == zero-filled-dense invariant,
 * mask authority (values under a false bit are ignored), and the reference's
 * own sin/cos/periodic-mask generators.
 */
#include <pthread.h>
#include <stdatomic.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define MC 252
#define KC 512
#define NC 240
#define MR 14
#define NR_LANE 2
#define LANE 8
#define NR (NR_LANE * LANE)

typedef double v8d __attribute__((vector_size(LANE * sizeof(double))));

/* Status codes: same values as include/cfgmm.h. */
#define OR_OK 0
#define OR_EINVAL 1
#define OR_EINVAL_TRANSB_MASK 2
#define OR_EOOM 4

static inline v8d v8_load(const double *p) {
    v8d r;
    memcpy(&r, p, sizeof r);
    return r;
}
static inline void v8_store(double *p, v8d v) { memcpy(p, &v, sizeof v); }
static inline v8d v8_splat(double x) { return (v8d){x, x, x, x, x, x, x, x}; }

static size_t div_ceil(size_t a, size_t b) { return (a + b - 1) / b; }
static size_t min_sz(size_t a, size_t b) { return a < b ? a : b; }

/* ------------------------------------------------------------------------ */
/* thread pool: run fn(ctx, tid) on nthreads threads                         */
/* ------------------------------------------------------------------------ */
typedef void (*par_fn)(void *ctx, int tid);
struct par_arg { par_fn fn; void *ctx; int tid; };

static void *par_trampoline(void *p) {
    struct par_arg *a = (struct par_arg *)p;
    a->fn(a->ctx, a->tid);
    return NULL;
}

int oracle_default_threads(void) {
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    return n > 0 ? (int)n : 1;
}

static int resolve_threads(int nthreads) { return nthreads > 0 ? nthreads : oracle_default_threads(); }

static void par_run(par_fn fn, void *ctx, int nthreads) {
    if (nthreads <= 1) { fn(ctx, 0); return; }
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
    struct par_arg *args = (struct par_arg *)malloc(sizeof(struct par_arg) * (size_t)nthreads);
    for (int t = 1; t < nthreads; ++t) {
        args[t].fn = fn; args[t].ctx = ctx; args[t].tid = t;
        pthread_create(&th[t], NULL, par_trampoline, &args[t]);
    }
    fn(ctx, 0);
    for (int t = 1; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(args);
    free(th);
}

/* ------------------------------------------------------------------------ */
/* microkernels                                                              */
/* ------------------------------------------------------------------------ */

/* src/structs.rs:34-45: for p<kc, i<MR, j_lane: c[i][j_lane] += splat(a[p][i]) * b[p][j_lane].
 * Built with -ffp-contract=fast, as the reference README asks (README.md:36), so
 * the update is one fused multiply-add per element. */
static inline void microkernel(v8d c[MR][NR_LANE], const double (*a)[MR], const v8d (*b)[NR_LANE], size_t kc) {
    v8d r[MR][NR_LANE];
    for (int i = 0; i < MR; ++i)
        for (int j = 0; j < NR_LANE; ++j) r[i][j] = c[i][j];
    for (size_t p = 0; p < kc; ++p) {
        const v8d b0 = b[p][0], b1 = b[p][1];
#pragma GCC unroll 14
        for (int i = 0; i < MR; ++i) {
            const v8d ai = v8_splat(a[p][i]);
            r[i][0] += ai * b0;
            r[i][1] += ai * b1;
        }
    }
    for (int i = 0; i < MR; ++i)
        for (int j = 0; j < NR_LANE; ++j) c[i][j] = r[i][j];
}

/* src/impl_matmul_non0tab.rs:21-41: A indexed by the original p, B by the
 * compressed idx_p; `fma_from` is a true FMA (src/naive_simd.rs:166-172),
 * which -ffp-contract=fast gives here. */
static inline void microkernel_mask_k_for_b(v8d c[MR][NR_LANE], const double (*a)[MR], const v8d (*b)[NR_LANE],
                                            const uint32_t *indices, size_t count) {
    v8d r[MR][NR_LANE];
    for (int i = 0; i < MR; ++i)
        for (int j = 0; j < NR_LANE; ++j) r[i][j] = c[i][j];
    for (size_t ip = 0; ip < count; ++ip) {
        const size_t p = indices[ip];
        const v8d b0 = b[ip][0], b1 = b[ip][1];
#pragma GCC unroll 14
        for (int i = 0; i < MR; ++i) {
            const v8d ai = v8_splat(a[p][i]);
            r[i][0] += ai * b0;
            r[i][1] += ai * b1;
        }
    }
    for (int i = 0; i < MR; ++i)
        for (int j = 0; j < NR_LANE; ++j) c[i][j] = r[i][j];
}

/* ------------------------------------------------------------------------ */
/* C tile load/store around a microkernel call                               */
/* ------------------------------------------------------------------------ */

/* src/impl_matmul.rs:30-71: full tile via vector loads, edge tile element by element. */
static inline void c_tile_load(v8d creg[MR][NR_LANE], const double *c, size_t mr, size_t nr, size_t ldc) {
    if (mr == MR && nr == NR) {
        for (int ii = 0; ii < MR; ++ii)
            for (int jl = 0; jl < NR_LANE; ++jl) creg[ii][jl] = v8_load(c + ii * ldc + jl * LANE);
    } else {
        memset(creg, 0, sizeof(v8d) * MR * NR_LANE);
        for (size_t ii = 0; ii < mr; ++ii)
            for (size_t jj = 0; jj < nr; ++jj) creg[ii][jj / LANE][jj % LANE] = c[ii * ldc + jj];
    }
}

static inline void c_tile_store(double *c, v8d creg[MR][NR_LANE], size_t mr, size_t nr, size_t ldc) {
    if (mr == MR && nr == NR) {
        for (int ii = 0; ii < MR; ++ii)
            for (int jl = 0; jl < NR_LANE; ++jl) v8_store(c + ii * ldc + jl * LANE, creg[ii][jl]);
    } else {
        for (size_t ii = 0; ii < mr; ++ii)
            for (size_t jj = 0; jj < nr; ++jj) c[ii * ldc + jj] = creg[ii][jj / LANE][jj % LANE];
    }
}

/* ------------------------------------------------------------------------ */
/* packing                                                                   */
/* ------------------------------------------------------------------------ */
typedef double apack_t[KC][MR];   /* one packed MR strip of A: [KC][MR] */
typedef v8d bpack_t[NR_LANE];     /* one packed k row of a B strip      */

/* src/impl_matmul.rs:77-98 */
static void pack_b_no_trans(bpack_t *dst, const double *src, size_t kc, size_t nr, size_t ldb) {
    if (nr == NR) {
        for (size_t p = 0; p < kc; ++p)
            for (int jl = 0; jl < NR_LANE; ++jl) dst[p][jl] = v8_load(src + p * ldb + jl * LANE);
    } else {
        for (size_t p = 0; p < kc; ++p)
            for (size_t jj = 0; jj < nr; ++jj) dst[p][jj / LANE][jj % LANE] = src[p * ldb + jj];
    }
}

/* src/impl_matmul.rs:101-117: copy nr rows of B^T to a [NR][KC] buffer, then
 * transpose all KC x NR entries into the packed strip. */
static void pack_b_trans(bpack_t *dst, const double *src, size_t kc, size_t nr, size_t ldb, double *buf) {
    for (size_t j = 0; j < nr; ++j) memcpy(buf + j * KC, src + j * ldb, kc * sizeof(double));
    double *d = (double *)dst;
    for (size_t p = 0; p < KC; ++p)
        for (size_t j = 0; j < NR; ++j) d[p * NR + j] = buf[j * KC + p];
}

/* src/impl_matmul_non0tab.rs:119-129: compact {p < k : tab[p*ldtab]}. */
static size_t get_indices_from_non0tab(uint32_t *indices, const uint8_t *tab, size_t ldtab, size_t k) {
    size_t count = 0;
    for (size_t p = 0; p < k; ++p)
        if (tab[p * ldtab]) indices[count++] = (uint32_t)p;
    return count;
}

/* src/impl_matmul_non0tab.rs:132-161: pack only the listed rows. */
static void pack_b_no_trans_non0tab(bpack_t *dst, const double *src, size_t nr, size_t ldb,
                                    const uint32_t *indices, size_t count) {
    if (nr == NR) {
        for (size_t ip = 0; ip < count; ++ip)
            for (int jl = 0; jl < NR_LANE; ++jl) dst[ip][jl] = v8_load(src + (size_t)indices[ip] * ldb + jl * LANE);
    } else {
        for (size_t ip = 0; ip < count; ++ip)
            for (size_t jj = 0; jj < nr; ++jj) dst[ip][jj / LANE][jj % LANE] = src[(size_t)indices[ip] * ldb + jj];
    }
}

/* ------------------------------------------------------------------------ */
/* GEMM driver (plain and masked share the schedule)                         */
/* ------------------------------------------------------------------------ */
struct gemm_ctx {
    double *c;
    const double *a, *b;
    size_t m, n, k, lda, ldb, ldc;
    int transa, transb;
    const uint8_t *tab; /* NULL for the plain GEMM */
    size_t ldtab, blksize, nc;
    size_t ntask_mc, ntask_nc, ntask_kc, ntask_mr, ntask_nr;
    apack_t *a_pack;
    pthread_mutex_t *barrier;
    atomic_size_t counter;
    size_t ntasks;
};

/* src/impl_matmul.rs:184-205 (and src/impl_matmul_non0tab.rs:245-266). */
static void gemm_pack_a_worker(void *vctx, int tid) {
    (void)tid;
    struct gemm_ctx *x = (struct gemm_ctx *)vctx;
    const size_t total = x->ntask_mc * x->ntask_kc;
    for (;;) {
        const size_t task_id = atomic_fetch_add(&x->counter, 1);
        if (task_id >= total) break;
        const size_t task_mc = task_id / x->ntask_kc, task_kc = task_id % x->ntask_kc;
        const size_t idx_m = task_mc * MC, idx_k = task_kc * KC;
        const size_t mc = min_sz(MC, x->m - idx_m), kc = min_sz(KC, x->k - idx_k);
        for (size_t task_mr = 0; task_mr < div_ceil(mc, MR); ++task_mr) {
            const size_t idx_mr = task_mr * MR;
            apack_t *ap = &x->a_pack[task_mc * (x->ntask_kc * x->ntask_mr) + task_kc * x->ntask_mr + task_mr];
            const size_t mr = min_sz(MR, mc - idx_mr);
            for (size_t p = 0; p < kc; ++p)
                for (size_t i = 0; i < mr; ++i)
                    (*ap)[p][i] = x->transa ? x->a[(idx_k + p) * x->lda + (idx_m + idx_mr + i)]
                                            : x->a[(idx_m + idx_mr + i) * x->lda + (idx_k + p)];
        }
    }
}

/* src/impl_matmul.rs:10-74 with the NR loop of src/impl_matmul.rs:120-153,
 * and the masked variant src/impl_matmul_non0tab.rs:53-116,164-208. */
static void gemm_compute_worker(void *vctx, int tid) {
    (void)tid;
    struct gemm_ctx *x = (struct gemm_ctx *)vctx;
    bpack_t *buf_b = (bpack_t *)aligned_alloc(64, sizeof(bpack_t) * KC);
    double *buf_trans_b = (double *)aligned_alloc(64, sizeof(double) * KC * NR);
    uint32_t *indices = (uint32_t *)malloc(sizeof(uint32_t) * KC);
    memset(buf_b, 0, sizeof(bpack_t) * KC);
    memset(buf_trans_b, 0, sizeof(double) * KC * NR);
    v8d creg[MR][NR_LANE];
    for (;;) {
        const size_t task_id = atomic_fetch_add(&x->counter, 1);
        if (task_id >= x->ntasks) break;
        const size_t task_m = task_id % x->ntask_mc;
        const size_t task_n = (task_id / x->ntask_mc) % x->ntask_nc;
        const size_t task_k = task_id / (x->ntask_mc * x->ntask_nc);
        const size_t mc = min_sz(MC, x->m - task_m * MC);
        const size_t nc = min_sz(x->nc, x->n - task_n * x->nc);
        const size_t kc = min_sz(KC, x->k - task_k * KC);
        const double *b = x->transb ? x->b + task_n * x->nc * x->ldb + task_k * KC
                                    : x->b + task_k * KC * x->ldb + task_n * x->nc;
        apack_t *a_mc_kc = &x->a_pack[task_m * (x->ntask_kc * x->ntask_mr) + task_k * x->ntask_mr];
        pthread_mutex_t *barrier = &x->barrier[(task_m * x->ntask_nc + task_n) * (x->ntask_nr * x->ntask_mr)];
        double *c = x->c + task_m * MC * x->ldc + task_n * x->nc;
        const uint8_t *tab = NULL;
        if (x->tab) tab = x->tab + (task_n * x->nc) / x->blksize + task_k * KC * x->ldtab;
        const size_t blknr = x->tab ? x->blksize / NR : 1;

        for (size_t task_j = 0, j = 0; j < nc; ++task_j, j += NR) {
            const size_t nr = min_sz(NR, nc - j);
            size_t count = 0;
            if (tab) {
                count = get_indices_from_non0tab(indices, tab + task_j / blknr, x->ldtab, kc);
                if (count == 0) continue;
                pack_b_no_trans_non0tab(buf_b, b + j, nr, x->ldb, indices, count);
            } else if (x->transb) {
                pack_b_trans(buf_b, b + j * x->ldb, kc, nr, x->ldb, buf_trans_b);
            } else {
                pack_b_no_trans(buf_b, b + j, kc, nr, x->ldb);
            }
            for (size_t task_i = 0, i = 0; i < mc; ++task_i, i += MR) {
                const size_t mr = min_sz(MR, mc - i);
                pthread_mutex_t *lock = &barrier[task_j * x->ntask_mr + task_i];
                pthread_mutex_lock(lock);
                c_tile_load(creg, c + i * x->ldc + j, mr, nr, x->ldc);
                if (tab)
                    microkernel_mask_k_for_b(creg, (const double(*)[MR])a_mc_kc[task_i], (const bpack_t *)buf_b,
                                             indices, count);
                else
                    microkernel(creg, (const double(*)[MR])a_mc_kc[task_i], (const bpack_t *)buf_b, kc);
                c_tile_store(c + i * x->ldc + j, creg, mr, nr, x->ldc);
                pthread_mutex_unlock(lock);
            }
        }
    }
    free(indices);
    free(buf_trans_b);
    free(buf_b);
}

static int gemm_common(struct gemm_ctx *x, int nthreads) {
    if (x->m == 0 || x->n == 0 || x->k == 0) return OR_OK;
    nthreads = resolve_threads(nthreads);
    x->ntask_mc = div_ceil(x->m, MC);
    x->ntask_nc = div_ceil(x->n, x->nc);
    x->ntask_kc = div_ceil(x->k, KC);
    x->ntask_mr = div_ceil(MC, MR);
    x->ntask_nr = div_ceil(x->nc, NR);
    const size_t nbar = x->ntask_mc * x->ntask_nc * x->ntask_nr * x->ntask_mr;
    const size_t npack = x->ntask_mc * x->ntask_kc * x->ntask_mr;
    x->barrier = (pthread_mutex_t *)malloc(sizeof(pthread_mutex_t) * nbar);
    x->a_pack = (apack_t *)calloc(npack, sizeof(apack_t));
    if (!x->barrier || !x->a_pack) {
        free(x->barrier);
        free(x->a_pack);
        return OR_EOOM;
    }
    for (size_t i = 0; i < nbar; ++i) pthread_mutex_init(&x->barrier[i], NULL);
    atomic_store(&x->counter, 0);
    par_run(gemm_pack_a_worker, x, nthreads);
    atomic_store(&x->counter, 0);
    x->ntasks = x->ntask_mc * x->ntask_nc * x->ntask_kc;
    par_run(gemm_compute_worker, x, nthreads);
    for (size_t i = 0; i < nbar; ++i) pthread_mutex_destroy(&x->barrier[i]);
    free(x->barrier);
    free(x->a_pack);
    return OR_OK;
}

/* C += op(A) op(B), row-major.  src/impl_matmul.rs:155-231 */
int oracle_matmul_loop_parallel_mnk_pack_a(double *c, const double *a, const double *b, size_t m, size_t n, size_t k,
                                           size_t lda, size_t ldb, size_t ldc, int transa, int transb, int nthreads) {
    struct gemm_ctx x;
    memset(&x, 0, sizeof x);
    x.c = c; x.a = a; x.b = b;
    x.m = m; x.n = n; x.k = k; x.lda = lda; x.ldb = ldb; x.ldc = ldc;
    x.transa = transa != 0; x.transb = transb != 0;
    x.nc = NC;
    return gemm_common(&x, nthreads);
}

/* C(i,j) += sum_{p : tab[p*ldtab + j/blksize]} op(A)(i,p) B(p,j).
 * src/impl_matmul_non0tab.rs:210-305; transb is unimplemented there (228-230). */
int oracle_matmul_loop_parallel_mnk_pack_a_non0tab(double *c, const double *a, const double *b, size_t m, size_t n,
                                                   size_t k, size_t lda, size_t ldb, size_t ldc, int transa,
                                                   int transb, const uint8_t *tab, size_t ldtab, size_t blksize,
                                                   int nthreads) {
    if (transb) return OR_EINVAL_TRANSB_MASK;
    /* The reference's NC-block to tab-column map (task_n*NC/BLKSIZE, then
     * task_j/BLKNR) needs BLKSIZE % NR == 0 (debug_assert at
     * src/impl_matmul_non0tab.rs:227) and NC % BLKSIZE == 0; NC is taken as the
     * largest multiple of BLKSIZE not above 240, as for rho. */
    if (blksize == 0 || blksize % NR != 0) return OR_EINVAL;
    struct gemm_ctx x;
    memset(&x, 0, sizeof x);
    x.c = c; x.a = a; x.b = b;
    x.m = m; x.n = n; x.k = k; x.lda = lda; x.ldb = ldb; x.ldc = ldc;
    x.transa = transa != 0; x.transb = 0;
    x.tab = tab; x.ldtab = ldtab; x.blksize = blksize;
    x.nc = NC >= blksize ? (NC / blksize) * blksize : blksize;
    return gemm_common(&x, nthreads);
}

/* ------------------------------------------------------------------------ */
/* fused masked rho                                                          */
/* ------------------------------------------------------------------------ */
struct rho_ctx {
    double *rhos;
    const double *dms, *ao_lhs, *ao_rhs;
    size_t stride_ao, stride_grid, nao, ngrid, nset;
    const uint8_t *tab;
    size_t ldtab, blksize, nc;
    size_t ntask_uc, ntask_vc, ntask_gc, ntask_ur, ntask_gr;
    pthread_mutex_t *barrier;
    atomic_size_t counter;
};

/* src/impl_aodmao2rho.rs:10-29: dm_packed[chunk][idx_p][idx_m] = D[u_m, v_p]. */
static void pack_dm_in_aodmao2rho(apack_t *dm_packed, const double *dm, size_t stride_ao, const uint32_t *idx_u,
                                  size_t count_u, const uint32_t *idx_v, size_t count_v) {
    for (size_t ip = 0; ip < count_v; ++ip)
        for (size_t im = 0; im < count_u; ++im)
            dm_packed[im / MR][ip][im % MR] = dm[(size_t)idx_u[im] * stride_ao + idx_v[ip]];
}

/* src/impl_aodmao2rho.rs:33-95 */
static void aodmao2rho_loop_1st(double *rho, apack_t **dm_packed, const bpack_t *ao_lhs_packed,
                                const bpack_t *ao_rhs_packed, size_t gr, size_t mc, size_t vc, size_t stride_grid,
                                size_t nset, pthread_mutex_t *barrier) {
    if (mc == 0 || vc == 0) return;
    v8d creg[MR][NR_LANE];
    for (size_t task_mr = 0, m_start = 0; m_start < mc; ++task_mr, m_start += MR) {
        const size_t mr = min_sz(MR, mc - m_start);
        for (size_t iset = 0; iset < nset; ++iset) {
            memset(creg, 0, sizeof creg);
            microkernel(creg, (const double(*)[MR])dm_packed[iset][task_mr], ao_rhs_packed, vc);
            v8d rho_reg[NR_LANE];
            memset(rho_reg, 0, sizeof rho_reg);
            for (size_t mm = 0; mm < mr; ++mm)
                for (int jl = 0; jl < NR_LANE; ++jl) rho_reg[jl] += creg[mm][jl] * ao_lhs_packed[m_start + mm][jl];
            pthread_mutex_lock(barrier);
            if (gr == NR) {
                for (int jl = 0; jl < NR_LANE; ++jl) {
                    double *r = rho + iset * stride_grid + jl * LANE;
                    v8_store(r, rho_reg[jl] + v8_load(r));
                }
            } else {
                for (size_t j = 0; j < gr; ++j) rho[iset * stride_grid + j] += rho_reg[j / LANE][j % LANE];
            }
            pthread_mutex_unlock(barrier);
        }
    }
}

/* src/impl_aodmao2rho.rs:97-169 */
static void aodmao2rho_loop_2nd(struct rho_ctx *x, double *rhos, const double *dms, const double *ao_lhs,
                                const double *ao_rhs, size_t uc, size_t vc, size_t gc, pthread_mutex_t *barrier,
                                const uint8_t *tab_lhs, const uint8_t *tab_rhs, apack_t **dm_packed,
                                bpack_t *ao_lhs_packed, bpack_t *ao_rhs_packed, uint32_t *idx_u, uint32_t *idx_v) {
    const size_t blknr = x->blksize / NR;
    for (size_t task_gblk = 0, g = 0; g < gc; ++task_gblk, g += x->blksize) {
        const size_t gblk = min_sz(x->blksize, gc - g);
        const size_t count_v = get_indices_from_non0tab(idx_v, tab_rhs + task_gblk, x->ldtab, vc);
        const size_t count_u = get_indices_from_non0tab(idx_u, tab_lhs + task_gblk, x->ldtab, uc);
        if (count_v == 0 || count_u == 0) continue;
        for (size_t iset = 0; iset < x->nset; ++iset)
            pack_dm_in_aodmao2rho(dm_packed[iset], dms + iset * x->stride_ao * x->stride_ao, x->stride_ao, idx_u,
                                  count_u, idx_v, count_v);
        for (size_t task_g = 0, gg = g; gg < g + gblk; ++task_g, gg += NR) {
            const size_t gr = min_sz(NR, g + gblk - gg);
            pack_b_no_trans_non0tab(ao_rhs_packed, ao_rhs + gg, gr, x->stride_grid, idx_v, count_v);
            pack_b_no_trans_non0tab(ao_lhs_packed, ao_lhs + gg, gr, x->stride_grid, idx_u, count_u);
            aodmao2rho_loop_1st(rhos + gg, dm_packed, ao_lhs_packed, ao_rhs_packed, gr, count_u, count_v,
                                x->stride_grid, x->nset, &barrier[task_gblk * blknr + task_g]);
        }
    }
}

/* src/impl_aodmao2rho.rs:200-243 */
static void rho_worker(void *vctx, int tid) {
    (void)tid;
    struct rho_ctx *x = (struct rho_ctx *)vctx;
    apack_t **dm_packed = (apack_t **)malloc(sizeof(apack_t *) * x->nset);
    for (size_t s = 0; s < x->nset; ++s) dm_packed[s] = (apack_t *)calloc(x->ntask_ur, sizeof(apack_t));
    bpack_t *ao_lhs_packed = (bpack_t *)aligned_alloc(64, sizeof(bpack_t) * MC);
    bpack_t *ao_rhs_packed = (bpack_t *)aligned_alloc(64, sizeof(bpack_t) * KC);
    memset(ao_lhs_packed, 0, sizeof(bpack_t) * MC);
    memset(ao_rhs_packed, 0, sizeof(bpack_t) * KC);
    uint32_t *idx_u = (uint32_t *)malloc(sizeof(uint32_t) * MC);
    uint32_t *idx_v = (uint32_t *)malloc(sizeof(uint32_t) * KC);
    const size_t ntasks = x->ntask_uc * x->ntask_vc * x->ntask_gc;
    for (;;) {
        const size_t task_id = atomic_fetch_add(&x->counter, 1);
        if (task_id >= ntasks) break;
        const size_t task_u = task_id % x->ntask_uc;
        const size_t task_g = (task_id / x->ntask_uc) % x->ntask_gc;
        const size_t task_v = task_id / (x->ntask_uc * x->ntask_gc);
        const size_t task_tab = (task_g * x->nc) / x->blksize;
        const size_t u_start = task_u * MC, v_start = task_v * KC, g_start = task_g * x->nc;
        const size_t u_len = min_sz((task_u + 1) * MC, x->nao) - u_start;
        const size_t v_len = min_sz((task_v + 1) * KC, x->nao) - v_start;
        const size_t g_len = min_sz((task_g + 1) * x->nc, x->ngrid) - g_start;
        aodmao2rho_loop_2nd(x, x->rhos + g_start, x->dms + u_start * x->stride_ao + v_start,
                            x->ao_lhs + u_start * x->stride_grid + g_start,
                            x->ao_rhs + v_start * x->stride_grid + g_start, u_len, v_len, g_len,
                            &x->barrier[task_g * x->ntask_gr], x->tab + u_start * x->ldtab + task_tab,
                            x->tab + v_start * x->ldtab + task_tab, dm_packed, ao_lhs_packed, ao_rhs_packed, idx_u,
                            idx_v);
    }
    free(idx_v);
    free(idx_u);
    free(ao_rhs_packed);
    free(ao_lhs_packed);
    for (size_t s = 0; s < x->nset; ++s) free(dm_packed[s]);
    free(dm_packed);
}

/* rho[s*stride_grid+g] += sum_{u,v on in block g/blksize} lhs[u,g] D_s[u,v] rhs[v,g].
 * src/impl_aodmao2rho.rs:171-244 */
int oracle_aodmao2rho_loop_parallel(double *rhos, const double *dms, const double *ao_lhs, const double *ao_rhs,
                                    size_t stride_ao, size_t stride_grid, size_t nao, size_t ngrid, size_t nset,
                                    const uint8_t *tab, size_t ldtab, size_t blksize, int nthreads) {
    if (blksize == 0 || blksize % NR != 0) return OR_EINVAL;
    if (nao == 0 || ngrid == 0 || nset == 0) return OR_OK;
    struct rho_ctx x;
    memset(&x, 0, sizeof x);
    x.rhos = rhos; x.dms = dms; x.ao_lhs = ao_lhs; x.ao_rhs = ao_rhs;
    x.stride_ao = stride_ao; x.stride_grid = stride_grid; x.nao = nao; x.ngrid = ngrid; x.nset = nset;
    x.tab = tab; x.ldtab = ldtab; x.blksize = blksize;
    x.nc = NC >= blksize ? (NC / blksize) * blksize : blksize;
    x.ntask_uc = div_ceil(nao, MC);
    x.ntask_vc = div_ceil(nao, KC);
    x.ntask_gc = div_ceil(ngrid, x.nc);
    x.ntask_ur = div_ceil(MC, MR);
    x.ntask_gr = div_ceil(x.nc, NR);
    const size_t nbar = x.ntask_gc * x.ntask_gr;
    x.barrier = (pthread_mutex_t *)malloc(sizeof(pthread_mutex_t) * nbar);
    if (!x.barrier) return OR_EOOM;
    for (size_t i = 0; i < nbar; ++i) pthread_mutex_init(&x.barrier[i], NULL);
    atomic_store(&x.counter, 0);
    par_run(rho_worker, &x, resolve_threads(nthreads));
    for (size_t i = 0; i < nbar; ++i) pthread_mutex_destroy(&x.barrier[i]);
    free(x.barrier);
    return OR_OK;
}

/* f64 wrappers with the reference's fixed BLKSIZE = 48
 * (src/impl_matmul.rs:234-248, src/impl_matmul_non0tab.rs:308-326, src/impl_aodmao2rho.rs:247-261). */
int oracle_matmul_anyway_full(double *c, const double *a, const double *b, size_t m, size_t n, size_t k, size_t lda,
                              size_t ldb, size_t ldc, int transa, int transb, int nthreads) {
    return oracle_matmul_loop_parallel_mnk_pack_a(c, a, b, m, n, k, lda, ldb, ldc, transa, transb, nthreads);
}

int oracle_matmul_anyway_full_non0tab(double *c, const double *a, const double *b, size_t m, size_t n, size_t k,
                                      size_t lda, size_t ldb, size_t ldc, int transa, int transb, const uint8_t *tab,
                                      size_t ldtab, int nthreads) {
    return oracle_matmul_loop_parallel_mnk_pack_a_non0tab(c, a, b, m, n, k, lda, ldb, ldc, transa, transb, tab, ldtab,
                                                          48, nthreads);
}

int oracle_aodmao2rho_anyway(double *rhos, const double *dms, const double *ao_lhs, const double *ao_rhs,
                             size_t stride_ao, size_t stride_grid, size_t nao, size_t ngrid, size_t nset,
                             const uint8_t *tab, size_t ldtab, int nthreads) {
    return oracle_aodmao2rho_loop_parallel(rhos, dms, ao_lhs, ao_rhs, stride_ao, stride_grid, nao, ngrid, nset, tab,
                                           ldtab, 48, nthreads);
}
