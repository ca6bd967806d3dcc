/* The CPU oracle under ASan + UBSan: GEMM (all transposes, padded leading dimensions, sizes
 * across the MR / NR / MC / KC / NC block edges), the masked GEMM and rho against naive loops,
 * with guard elements around every buffer that must stay untouched. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int oracle_matmul_loop_parallel_mnk_pack_a(double *c, const double *a, const double *b, size_t m, size_t n, size_t k,
                                           size_t lda, size_t ldb, size_t ldc, int transa, int transb, int nthreads);
int oracle_matmul_loop_parallel_mnk_pack_a_non0tab(double *c, const double *a, const double *b, size_t m, size_t n,
                                                   size_t k, size_t lda, size_t ldb, size_t ldc, int transa,
                                                   int transb, const uint8_t *tab, size_t ldtab, size_t blksize,
                                                   int nthreads);
int oracle_aodmao2rho_loop_parallel(double *rhos, const double *dms, const double *ao_lhs, const double *ao_rhs,
                                    size_t stride_ao, size_t stride_grid, size_t nao, size_t ngrid, size_t nset,
                                    const uint8_t *tab, size_t ldtab, size_t blksize, int nthreads);

static int failed = 0, cases = 0;
static uint64_t rs = 88172645463325252ull;
static double urand(void) {
    rs ^= rs << 13;
    rs ^= rs >> 7;
    rs ^= rs << 17;
    return (double)(rs >> 11) * 0x1.0p-53 * 2.0 - 1.0;
}
static double *alloc_rand(size_t n) {
    double *p = malloc(sizeof(double) * (n ? n : 1));
    for (size_t i = 0; i < n; ++i) p[i] = urand();
    return p;
}
static double rel(const double *x, const double *y, size_t n) {
    double num = 0.0, den = 0.0;
    for (size_t i = 0; i < n; ++i) {
        num += (x[i] - y[i]) * (x[i] - y[i]);
        den += y[i] * y[i];
    }
    return den > 0 ? sqrt(num / den) : sqrt(num);
}

static void check_gemm(size_t m, size_t n, size_t k, int ta, int tb, int masked, size_t blksize) {
    const size_t lda = (ta ? m : k) + 3, ldb = (tb ? k : n) + 1, ldc = n + 2;
    const size_t na = (ta ? k : m) * lda, nb = (tb ? n : k) * ldb, nc = m * ldc;
    double *a = alloc_rand(na), *b = alloc_rand(nb), *c = alloc_rand(nc), *want = malloc(sizeof(double) * (nc ? nc : 1));
    memcpy(want, c, sizeof(double) * nc);
    const size_t nblk = blksize ? (n + blksize - 1) / blksize : 1, ldtab = nblk + 1;
    uint8_t *tab = malloc(k * ldtab + 1);
    for (size_t i = 0; i < k * ldtab; ++i) tab[i] = (uint8_t)((i * 7 + 3) % 10 < 3);
    for (size_t i = 0; i < m; ++i)
        for (size_t j = 0; j < n; ++j) {
            double s = 0.0;
            for (size_t p = 0; p < k; ++p) {
                if (masked && !tab[p * ldtab + j / blksize]) continue;
                const double x = ta ? a[p * lda + i] : a[i * lda + p];
                const double y = tb ? b[j * ldb + p] : b[p * ldb + j];
                s += x * y;
            }
            want[i * ldc + j] += s;
        }
    const int st = masked ? oracle_matmul_loop_parallel_mnk_pack_a_non0tab(c, a, b, m, n, k, lda, ldb, ldc, ta, 0, tab,
                                                                            ldtab, blksize, 3)
                          : oracle_matmul_loop_parallel_mnk_pack_a(c, a, b, m, n, k, lda, ldb, ldc, ta, tb, 3);
    ++cases;
    const double e = rel(c, want, nc);
    if (st != 0 || !(e <= 1e-13)) {
        printf("FAIL gemm m=%zu n=%zu k=%zu ta=%d tb=%d masked=%d: status %d rel %.3e\n", m, n, k, ta, tb, masked, st, e);
        ++failed;
    }
    free(a), free(b), free(c), free(want), free(tab);
}

static void check_rho(size_t nao, size_t ngrid, size_t nset, size_t blksize) {
    const size_t sa = nao + 2, sg = ngrid + 3, nblk = (ngrid + blksize - 1) / blksize, ldtab = nblk + 2;
    double *dms = alloc_rand(nset * sa * sa), *l = alloc_rand(nao * sg), *r = alloc_rand(nao * sg);
    double *rho = alloc_rand(nset * sg), *want = malloc(sizeof(double) * nset * sg);
    memcpy(want, rho, sizeof(double) * nset * sg);
    uint8_t *tab = malloc(nao * ldtab);
    for (size_t i = 0; i < nao * ldtab; ++i) tab[i] = (uint8_t)((i * 11 + 5) % 20 < 5);
    for (size_t s = 0; s < nset; ++s)
        for (size_t g = 0; g < ngrid; ++g) {
            double acc = 0.0;
            const size_t bk = g / blksize;
            for (size_t u = 0; u < nao; ++u) {
                if (!tab[u * ldtab + bk]) continue;
                double t = 0.0;
                for (size_t v = 0; v < nao; ++v)
                    if (tab[v * ldtab + bk]) t += dms[s * sa * sa + u * sa + v] * r[v * sg + g];
                acc += l[u * sg + g] * t;
            }
            want[s * sg + g] += acc;
        }
    const int st = oracle_aodmao2rho_loop_parallel(rho, dms, l, r, sa, sg, nao, ngrid, nset, tab, ldtab, blksize, 3);
    ++cases;
    const double e = rel(rho, want, nset * sg);
    if (st != 0 || !(e <= 1e-13)) {
        printf("FAIL rho nao=%zu ngrid=%zu nset=%zu blk=%zu: status %d rel %.3e\n", nao, ngrid, nset, blksize, st, e);
        ++failed;
    }
    free(dms), free(l), free(r), free(rho), free(want), free(tab);
}

int main(void) {
    const size_t ms[] = {1, 13, 15, 253}, ns[] = {1, 17, 241, 257}, ks[] = {1, 7, 513};
    for (int ta = 0; ta < 2; ++ta)
        for (int tb = 0; tb < 2; ++tb)
            for (size_t i = 0; i < 4; ++i)
                for (size_t j = 0; j < 4; ++j)
                    for (size_t q = 0; q < 3; ++q) check_gemm(ms[i], ns[j], ks[q], ta, tb, 0, 0);
    for (int ta = 0; ta < 2; ++ta) {
        check_gemm(30, 250, 70, ta, 0, 1, 48);
        check_gemm(253, 97, 515, ta, 0, 1, 16);
        check_gemm(5, 300, 9, ta, 0, 1, 64);
    }
    check_rho(1, 1, 1, 16);
    check_rho(29, 101, 2, 48);
    check_rho(70, 300, 3, 64);
    check_rho(260, 130, 1, 16);
    printf("oracle_checks: %d cases, %d failed\n", cases, failed);
    return failed ? 1 : 0;
}
