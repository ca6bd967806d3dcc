// Argument validation of libcfgmm's C ABI under ASan + UBSan, host code only (kernel launchers
// are stubs).  Each case passes arguments the reference would panic on (or that would overflow
// an extent) and expects the documented status, a message, and no launcher reached.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <vector>

#include "cfgmm.h"

extern int g_stub_calls;
static int g_failed = 0, g_cases = 0;

#define EXPECT_STATUS(expr, want)                                                                         \
    do {                                                                                                  \
        ++g_cases;                                                                                        \
        const int calls0 = g_stub_calls;                                                                  \
        const int got_ = (expr);                                                                          \
        const char *msg_ = cfgmm_last_error();                                                            \
        if (got_ != (want) || g_stub_calls != calls0 || !msg_ || !msg_[0]) {                              \
            std::printf("FAIL line %d: %s -> %d (want %d), stub calls %d, message '%s'\n", __LINE__, #expr, \
                        got_, (want), g_stub_calls - calls0, msg_ ? msg_ : "(null)");                     \
            ++g_failed;                                                                                   \
        }                                                                                                 \
    } while (0)

int main() {
    const size_t SZ_MAX = std::numeric_limits<size_t>::max();
    const size_t m = 7, n = 9, k = 11;
    std::vector<double> a(m * k), b(k * n), c(m * n);
    std::vector<uint8_t> tab(k * 1, 1);

    // GEMM host slices: short slices, small leading dimensions, overflowing extents, nulls
    EXPECT_STATUS(cfgmm_matmul_anyway_full(c.data(), c.size(), a.data(), a.size() - 1, b.data(), b.size(), m, n, k,
                                           k, n, n, 0, 0), CFGMM_EBOUNDS);
    EXPECT_STATUS(cfgmm_matmul_anyway_full(c.data(), c.size() - 1, a.data(), a.size(), b.data(), b.size(), m, n, k,
                                           k, n, n, 0, 0), CFGMM_EBOUNDS);
    EXPECT_STATUS(cfgmm_matmul_loop_parallel_mnk_pack_a(c.data(), c.size(), a.data(), a.size(), b.data(), b.size(),
                                                        m, n, k, k - 1, n, n, 0, 0), CFGMM_EINVAL_DIMS);
    EXPECT_STATUS(cfgmm_matmul_loop_macro_mb(c.data(), c.size(), a.data(), a.size(), b.data(), b.size(), m, n, k, k,
                                             n, n - 1, 0, 0), CFGMM_EINVAL_DIMS);
    EXPECT_STATUS(cfgmm_matmul_anyway_full(c.data(), c.size(), a.data(), a.size(), b.data(), b.size(), m, n, k, k,
                                           n - 1, n, 1, 1), CFGMM_EINVAL_DIMS);  // transa: lda >= m
    EXPECT_STATUS(cfgmm_matmul_anyway_full(c.data(), c.size(), a.data(), a.size(), b.data(), b.size(), SZ_MAX / 4, n,
                                           k, SZ_MAX / 2, n, n, 0, 0), CFGMM_EINVAL_DIMS);  // extent overflows
    EXPECT_STATUS(cfgmm_matmul_anyway_full(c.data(), c.size(), nullptr, 0, b.data(), b.size(), m, n, k, k, n, n, 0,
                                           0), CFGMM_EINVAL);
    // masked GEMM: transb, BLKSIZE, short mask
    EXPECT_STATUS(cfgmm_matmul_anyway_full_non0tab(c.data(), c.size(), a.data(), a.size(), b.data(), b.size(), m, n,
                                                   k, k, k, n, 0, 1, tab.data(), tab.size(), 1),
                  CFGMM_EINVAL_TRANSB_MASK);
    EXPECT_STATUS(cfgmm_matmul_loop_parallel_mnk_pack_a_non0tab(c.data(), c.size(), a.data(), a.size(), b.data(),
                                                                b.size(), m, n, k, k, n, n, 0, 0, tab.data(),
                                                                tab.size(), 1, 40), CFGMM_EINVAL);
    EXPECT_STATUS(cfgmm_matmul_loop_parallel_mnk_pack_a_non0tab(c.data(), c.size(), a.data(), a.size(), b.data(),
                                                                b.size(), m, n, k, k, n, n, 0, 0, tab.data(),
                                                                tab.size(), 1, 0), CFGMM_EINVAL);
    EXPECT_STATUS(cfgmm_matmul_anyway_full_non0tab(c.data(), c.size(), a.data(), a.size(), b.data(), b.size(), m, n,
                                                   k, k, n, n, 0, 0, tab.data(), tab.size() - 1, 1), CFGMM_EBOUNDS);

    // rho host slices
    const size_t nao = 5, ngrid = 100, nset = 2, nblk = 3;  // blksize 48
    std::vector<double> rho(nset * ngrid), dms(nset * nao * nao), phi(nao * ngrid);
    std::vector<uint8_t> rtab(nao * nblk, 1);
    EXPECT_STATUS(cfgmm_aodmao2rho_anyway(rho.data(), rho.size(), dms.data(), dms.size() - 1, phi.data(), phi.size(),
                                          phi.data(), phi.size(), nao, ngrid, nao, ngrid, nset, rtab.data(),
                                          rtab.size(), nblk), CFGMM_EBOUNDS);
    EXPECT_STATUS(cfgmm_aodmao2rho_anyway(rho.data(), rho.size() - 1, dms.data(), dms.size(), phi.data(), phi.size(),
                                          phi.data(), phi.size(), nao, ngrid, nao, ngrid, nset, rtab.data(),
                                          rtab.size(), nblk), CFGMM_EBOUNDS);
    EXPECT_STATUS(cfgmm_aodmao2rho_anyway(rho.data(), rho.size(), dms.data(), dms.size(), phi.data(), phi.size(),
                                          phi.data(), phi.size(), nao, ngrid, nao, ngrid, nset, rtab.data(),
                                          rtab.size(), nblk - 1), CFGMM_EINVAL_DIMS);
    EXPECT_STATUS(cfgmm_aodmao2rho_loop_parallel(rho.data(), rho.size(), dms.data(), dms.size(), phi.data(),
                                                 phi.size(), phi.data(), phi.size(), nao, ngrid, nao, ngrid, nset,
                                                 rtab.data(), rtab.size(), nblk, 24), CFGMM_EINVAL);
    EXPECT_STATUS(cfgmm_aodmao2rho_loop_parallel(rho.data(), rho.size(), dms.data(), dms.size(), phi.data(),
                                                 phi.size(), phi.data(), phi.size(), SZ_MAX / 2, ngrid, SZ_MAX / 2,
                                                 ngrid, nset, rtab.data(), rtab.size(), nblk, 48),
                  CFGMM_EINVAL_DIMS);  // stride_ao^2 overflows
    EXPECT_STATUS(cfgmm_aodmao2rho_loop_parallel(rho.data(), rho.size(), dms.data(), dms.size(), phi.data(),
                                                 phi.size(), phi.data(), phi.size(), nao, SZ_MAX / 3, nao,
                                                 ngrid, nset, rtab.data(), rtab.size(), nblk, 48),
                  CFGMM_EINVAL_DIMS);  // rho / phi extents overflow

    // device-pointer entry points
    void *s = nullptr;
    double *dp = reinterpret_cast<double *>(0x1000);
    const uint8_t *tp = reinterpret_cast<const uint8_t *>(0x1000);
    EXPECT_STATUS(cfgmm_dgemm_dev(s, nullptr, dp, dp, m, n, k, k, n, n, 0, 0), CFGMM_EINVAL);
    EXPECT_STATUS(cfgmm_dgemm_dev(s, dp, dp, dp, m, n, k, k - 1, n, n, 0, 0), CFGMM_EINVAL_DIMS);
    EXPECT_STATUS(cfgmm_dgemm_non0tab_dev(s, dp, dp, dp, m, n, k, k, n, n, 0, 1, tp, 1, 48),
                  CFGMM_EINVAL_TRANSB_MASK);
    EXPECT_STATUS(cfgmm_dgemm_non0tab_ex_dev(s, dp, dp, dp, m, n, k, k, n, n, 0, 0, nullptr, 1, 48), CFGMM_EINVAL);
    EXPECT_STATUS(cfgmm_aodmao2rho_dev(s, dp, dp, dp, dp, nao, ngrid, nao, ngrid, nset, nullptr, nblk, 48),
                  CFGMM_EINVAL);
    EXPECT_STATUS(cfgmm_aodmao2rho_dev_noalloc(s, dp, dp, dp, dp, nao, ngrid, nao, ngrid, nset, tp, nblk, 40),
                  CFGMM_EINVAL);
    EXPECT_STATUS(cfgmm_aodmao2rho_ex_dev(s, dp, dp, dp, dp, nao, ngrid, nao, ngrid, nset, tp, nblk, 48, 7, 0),
                  CFGMM_EINVAL);  // unknown precision
    EXPECT_STATUS(cfgmm_aodmao2rho_ex_dev(s, dp, dp, dp, dp, nao, ngrid, nao, ngrid, nset, tp, nblk, 48, 1, 9),
                  CFGMM_EINVAL);  // nslices out of range
    EXPECT_STATUS(cfgmm_dgemm_ex_dev(s, dp, dp, dp, m, n, k, k, n, n, 0, 0, 1, 1), CFGMM_EINVAL);
    EXPECT_STATUS(cfgmm_aodmao2rho_multi_dev(s, dp, nset * ngrid, dp, dp, nao * ngrid, 17, dp, nao, ngrid, nao,
                                             ngrid, nset, tp, nblk, 48), CFGMM_EINVAL);
    EXPECT_STATUS(cfgmm_aodmao2rho_multi_dev(s, dp, ngrid, dp, dp, nao * ngrid, 2, dp, nao, ngrid, nao, ngrid, nset,
                                             tp, nblk, 48), CFGMM_EINVAL_DIMS);  // component rho ranges overlap
    EXPECT_STATUS(cfgmm_aogridao2mat_dev(s, dp, nao - 1, dp, dp, nullptr, ngrid, nao, ngrid, tp, nblk, 48),
                  CFGMM_EINVAL_DIMS);
    cfgmm_rho_problem bad{};
    bad.rhos = dp;
    bad.dms = bad.ao_lhs = bad.ao_rhs = dp;
    bad.tab = tp;
    bad.stride_ao = nao;
    bad.nao = nao;
    bad.stride_grid = ngrid;
    bad.ngrid = ngrid;
    bad.nset = 1;
    bad.ldtab = nblk;
    bad.blksize = 20;
    EXPECT_STATUS(cfgmm_aodmao2rho_batched_dev(s, &bad, 1, 0, 0), CFGMM_EINVAL);
    if (!std::strstr(cfgmm_last_error(), "problem 0")) {
        std::printf("FAIL: batched error does not name the problem: %s\n", cfgmm_last_error());
        ++g_failed;
    }
    EXPECT_STATUS(cfgmm_aodmao2rho_batched_dev(s, nullptr, 3, 0, 0), CFGMM_EINVAL);

    // GTO shells, resident handles, the server, tunables
    cfgmm_gto_shell sh{{0.0, 0.0, 0.0}, 5, 1, 0, 0};
    double e1 = 1.0, c1 = 1.0;
    EXPECT_STATUS(cfgmm_eval_gto_dev(s, dp, ngrid, nullptr, nblk, &sh, 1, &e1, &c1, 1, dp, ngrid, 21, 48, 1e-8),
                  CFGMM_EINVAL);  // l = 5
    sh.l = 1;
    sh.prim_off = 1;
    EXPECT_STATUS(cfgmm_eval_gto_dev(s, dp, ngrid, nullptr, nblk, &sh, 1, &e1, &c1, 1, dp, ngrid, 3, 48, 1e-8),
                  CFGMM_EINVAL);  // primitives outside the table
    sh.prim_off = 0;
    EXPECT_STATUS(cfgmm_eval_gto_dev(s, dp, ngrid, nullptr, nblk, &sh, 1, &e1, &c1, 1, dp, ngrid, 2, 48, 1e-8),
                  CFGMM_EINVAL);  // 3 AOs past nao = 2
    EXPECT_STATUS(cfgmm_eval_gto_dev(s, dp, ngrid, nullptr, nblk, &sh, 1, &e1, &c1, 1, dp, ngrid, 3, 48, -1.0),
                  CFGMM_EINVAL);  // cutoff < 0
    EXPECT_STATUS(cfgmm_rho_resident_eval(s, nullptr, dp, ngrid, ngrid, 1), CFGMM_EINVAL);
    EXPECT_STATUS(cfgmm_rho_resident_create(s, dp, dp, dp, nao, ngrid, nao, ngrid, nset, tp, nblk, 48, 0, 5, nullptr),
                  CFGMM_EINVAL);
    cfgmm_rho_resident *h = nullptr;
    EXPECT_STATUS(cfgmm_rho_resident_create(s, dp, dp, dp, nao, ngrid, nao, ngrid, nset, tp, nblk, 48, 0, 5, &h),
                  CFGMM_EINVAL);  // resident operands exist in split precision only
    EXPECT_STATUS(cfgmm_server_create(nullptr, nullptr), CFGMM_EINVAL);
    EXPECT_STATUS(cfgmm_server_run(nullptr, nullptr, 0, nullptr), CFGMM_EINVAL);
    EXPECT_STATUS(cfgmm_tune_set(CFGMM_TUNE_COUNT, 1), CFGMM_EINVAL);
    EXPECT_STATUS(cfgmm_tune_set(CFGMM_TUNE_BATCH_WGS_PER_CU, 99), CFGMM_EINVAL);
    EXPECT_STATUS(cfgmm_tune_set(CFGMM_TUNE_DGEMM_SMALL_TILE, 96), CFGMM_EINVAL);
    EXPECT_STATUS(cfgmm_tune_set(CFGMM_TUNE_RESIDENT_COLS, 20), CFGMM_EINVAL);

    // a no-op call needs no device and clears the last error
    ++g_cases;
    if (cfgmm_matmul_anyway_full(c.data(), c.size(), a.data(), 0, b.data(), 0, m, n, 0, 1, n, n, 0, 0) != CFGMM_OK ||
        cfgmm_last_error()[0] != '\0') {
        std::printf("FAIL: k = 0 is not a no-op\n");
        ++g_failed;
    }
    // a valid call without a GPU reports the missing device (or, with one, reaches a stub)
    ++g_cases;
    const int st = cfgmm_matmul_anyway_full(c.data(), c.size(), a.data(), a.size(), b.data(), b.size(), m, n, k, k, n,
                                            n, 0, 0);
    if (st == CFGMM_OK || st == CFGMM_EINVAL || st == CFGMM_EBOUNDS || st == CFGMM_EINVAL_DIMS) {
        std::printf("FAIL: a valid call returned %d\n", st);
        ++g_failed;
    }
    std::printf("capi_checks: %d cases, %d failed\n", g_cases, g_failed);
    return g_failed ? 1 : 0;
}
