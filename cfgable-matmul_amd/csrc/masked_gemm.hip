// masked_gemm.hip -- non0tab-masked row-major f64 GEMM on FP64 MFMA (gfx950).
//
//   C(i,j) += sum_{p in K_b} op(A)(i,p) * B(p,j),   b = j / blksize,
//   K_b = { p < k : tab[p*ldtab + b] != 0 }   (ascending p)
//
// Replaces matmul_loop_parallel_mnk_pack_a_non0tab and its helpers
// (src/impl_matmul_non0tab.rs:53-305): the reference rescans the mask per NR
// strip and KC block (get_indices_from_non0tab, 119-129), packs only the
// listed B rows (132-161) and runs the masked microkernel over the compressed
// K list (21-41).  Here the K lists are compacted once on the device per
// column block (launch_build_lists), and each workgroup computes a 128-row x
// (16*NT)-column tile of one column block as a GEMM over the compacted K:
// gathered op(A) columns and gathered B rows are staged through
// double-buffered LDS and accumulated with v_mfma_f64_16x16x4_f64.  A tile
// whose K list is empty is skipped outright (the reference's `count == 0`
// skip, 200-202).  The reference leaves transb unimplemented (228-230), and
// so do the drop-in entry points; the additive cfgmm_dgemm_non0tab_ex_dev takes
// it: B(p, j) = b[j*ldb + p], loaded with p fastest so that a column's listed
// entries are read along its row of b.
//
// The element type is a template parameter (the reference's generic T, of which
// it implements only f64, src/structs.rs:19-21): T = float runs the same tiles
// on v_mfma_f32_16x16x4f32, whose A and B lane layouts are those of the f64
// instruction and whose result rows are 4*(l/16)+r instead of l/16+4r.
#include "internal.hpp"

namespace cfgmm {

// Zero sources for loads outside the problem or the compacted K list.
__device__ const double kZerosM[16] = {};
__device__ const float kZerosMf[16] = {};

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));

template <typename T>
struct MfmaOf;
template <>
struct MfmaOf<double> {
    typedef f64x4 acc_t;
    __device__ static acc_t step(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static int row(int lk, int r) { return lk + 4 * r; }  // D[l/16 + 4r][l%16]
    __device__ static const double *zeros() { return kZerosM; }
};
template <>
struct MfmaOf<float> {
    typedef f32x4 acc_t;
    __device__ static acc_t step(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static int row(int lk, int r) { return 4 * lk + r; }  // D[4(l/16) + r][l%16]
    __device__ static const float *zeros() { return kZerosMf; }
};

constexpr int kThreads = 256;
constexpr int kBM = 128;  // rows per workgroup (4 waves x 32)
constexpr int kBK = 16;   // compacted-K slice depth
constexpr int kSeg = 2048;  // compacted-K list entries staged in LDS at a time

template <typename T, int NT, bool TA, bool TB>
__global__ __launch_bounds__(kThreads, 2) void masked_gemm_kernel(T *__restrict__ c, const T *__restrict__ a,
                                                                  const T *__restrict__ b, int m, int n,
                                                                  long lda, long ldb, long ldc, int kmax,
                                                                  const int *__restrict__ lists,
                                                                  const int *__restrict__ counts, int blksize,
                                                                  int nsub) {
    constexpr int W = NT * 16;
    constexpr int FM = 2;  // 32 rows per wave
    constexpr int LDA_S = kBM + (TA ? 0 : 1);
    constexpr int A_PER_T = kBM * kBK / kThreads;
    constexpr int B_PER_T = kBK * W / kThreads;
    static_assert(B_PER_T * kThreads == kBK * W, "B tile split");

    const int cb = blockIdx.x / nsub, sub = blockIdx.x % nsub;
    const int cnt = counts[cb];
    if (cnt == 0) return;
    const int *__restrict__ list = lists + (long)cb * kmax;
    const int m0 = blockIdx.y * kBM;
    const int n0 = cb * blksize + sub * W;
    if (n0 >= n) return;  // subtile past the end of a partial last column block

    __shared__ T As[2][kBK][LDA_S];  // op(A)[i][list[pp]] stored [pp][i]
    __shared__ T Bs[2][kBK][W];      // B[list[pp]][j]
    __shared__ int vls[kSeg];             // a segment of the compacted K list

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int lrow = lane & 15, lk = lane >> 4;

    // Every load of a slice is issued unconditionally: indices come from LDS,
    // and an entry outside the problem reads kZerosM (the address is
    // selected, not the value, so no load hides behind a branch).

    typedef typename MfmaOf<T>::acc_t acc_t;
    T ra[A_PER_T], rb[B_PER_T];
    acc_t acc[FM][NT];
#pragma unroll
    for (int mi = 0; mi < FM; ++mi)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mi][nt] = acc_t{0, 0, 0, 0};

    for (int seg0 = 0; seg0 < cnt; seg0 += kSeg) {
        const int segn = min(kSeg, cnt - seg0);
        __syncthreads();
        for (int j = tid; j < segn; j += kThreads) vls[j] = list[seg0 + j];
        __syncthreads();

        auto load_slice = [&](int k0) {
#pragma unroll
            for (int e = 0; e < A_PER_T; ++e) {
                const int idx = tid + kThreads * e;
                const int pp = TA ? idx / kBM : idx % kBK;
                const int i = TA ? idx % kBM : idx / kBK;
                const int gi = m0 + i;
                const bool ok = gi < m && k0 + pp < segn;
                const long p = vls[ok ? k0 + pp : 0];
                const T *src = ok ? (TA ? a + p * lda + gi : a + (long)gi * lda + p) : MfmaOf<T>::zeros();
                ra[e] = *src;
            }
#pragma unroll
            for (int e = 0; e < B_PER_T; ++e) {
                const int idx = tid + kThreads * e;
                const int pp = TB ? idx % kBK : idx / W, gj = n0 + (TB ? idx / kBK : idx % W);
                const bool ok = gj < n && k0 + pp < segn;
                const long p = vls[ok ? k0 + pp : 0];
                const T *src = ok ? (TB ? b + (long)gj * ldb + p : b + p * ldb + gj) : MfmaOf<T>::zeros();
                rb[e] = *src;
            }
        };
        auto store_slice = [&](int buf) {
#pragma unroll
            for (int e = 0; e < A_PER_T; ++e) {
                const int idx = tid + kThreads * e;
                const int pp = TA ? idx / kBM : idx % kBK;
                const int i = TA ? idx % kBM : idx / kBK;
                As[buf][pp][i] = ra[e];
            }
#pragma unroll
            for (int e = 0; e < B_PER_T; ++e) {
                const int idx = tid + kThreads * e;
                if (TB)
                    Bs[buf][idx % kBK][idx / kBK] = rb[e];
                else
                    Bs[buf][idx / W][idx % W] = rb[e];
            }
        };

        const int nslices = (segn + kBK - 1) / kBK;
        load_slice(0);
        store_slice(0);
        __syncthreads();
        for (int t = 0; t < nslices; ++t) {
            const int cur = t & 1;
            if (t + 1 < nslices) load_slice((t + 1) * kBK);
#pragma unroll
            for (int kk = 0; kk < kBK / 4; ++kk) {
                T fa[FM], fb[NT];
#pragma unroll
                for (int mi = 0; mi < FM; ++mi) fa[mi] = As[cur][kk * 4 + lk][wave * 32 + mi * 16 + lrow];
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) fb[nt] = Bs[cur][kk * 4 + lk][nt * 16 + lrow];
#pragma unroll
                for (int mi = 0; mi < FM; ++mi)
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt)
                        acc[mi][nt] = MfmaOf<T>::step(fa[mi], fb[nt], acc[mi][nt]);
            }
            if (t + 1 < nslices) store_slice(cur ^ 1);
            __syncthreads();
        }
    }

#pragma unroll
    for (int mi = 0; mi < FM; ++mi)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wave * 32 + mi * 16 + MfmaOf<T>::row(lk, r);
                const int col = n0 + nt * 16 + lrow;
                if (row < m && col < n) c[(long)row * ldc + col] += acc[mi][nt][r];
            }
}

template <typename T, int NT>
int launch_nt(bool ta, bool tb, dim3 grid, hipStream_t stream, T *c, const T *a, const T *b, int m, int n, long lda,
              long ldb, long ldc, int kmax, const int *lists, const int *counts, int blksize, int nsub) {
    hipEvent_t t0 = trace_start(stream);
#define CFGMM_MASKED(TA_, TB_)                                                                                      \
    hipLaunchKernelGGL((masked_gemm_kernel<T, NT, TA_, TB_>), grid, dim3(kThreads), 0, stream, c, a, b, m, n, lda, \
                       ldb, ldc, kmax, lists, counts, blksize, nsub)
    if (ta && tb)
        CFGMM_MASKED(true, true);
    else if (ta)
        CFGMM_MASKED(true, false);
    else if (tb)
        CFGMM_MASKED(false, true);
    else
        CFGMM_MASKED(false, false);
#undef CFGMM_MASKED
    trace_stop(stream, "masked_gemm_kernel", t0);
    return hip_status(hipGetLastError(), "masked_gemm_kernel launch");
}

template <typename T>
int launch_gemm_non0tab(hipStream_t stream, T *c, const T *a, const T *b, size_t m, size_t n, size_t k, size_t lda,
                        size_t ldb, size_t ldc, bool transa, bool transb, const uint8_t *tab, size_t ldtab,
                        size_t blksize) {
    if (m == 0 || n == 0 || k == 0) return CFGMM_OK;
    if (m > INT32_MAX || n > INT32_MAX || k > INT32_MAX) return fail(CFGMM_EINVAL, "masked dgemm: size out of range");
    const size_t nblk = div_up(n, blksize);
    int nt = 1;
    for (int cand = 4; cand >= 1; --cand)
        if ((blksize / 16) % cand == 0) { nt = cand; break; }
    const int nsub = (int)(blksize / (16 * nt));
    const size_t mtiles = div_up(m, kBM);
    if (nblk * nsub > (size_t)INT32_MAX || mtiles > 65535u) return fail(CFGMM_EINVAL, "masked dgemm: grid too large");

    void *ws = nullptr;
    const size_t lists_bytes = sizeof(int) * nblk * k;
    int st = workspace_get(stream, lists_bytes + sizeof(int) * nblk, &ws);
    if (st != CFGMM_OK) return st;
    int *lists = static_cast<int *>(ws);
    int *counts = lists + nblk * k;
    st = launch_build_lists(stream, tab, ldtab, k, nblk, lists, counts);
    if (st != CFGMM_OK) return st;

    dim3 grid((unsigned)(nblk * nsub), (unsigned)mtiles, 1);
    const int im = (int)m, in = (int)n, ik = (int)k, ib = (int)blksize;
    switch (nt) {
        case 1: return launch_nt<T, 1>(transa, transb, grid, stream, c, a, b, im, in, (long)lda, (long)ldb, (long)ldc, ik, lists, counts, ib, nsub);
        case 2: return launch_nt<T, 2>(transa, transb, grid, stream, c, a, b, im, in, (long)lda, (long)ldb, (long)ldc, ik, lists, counts, ib, nsub);
        case 3: return launch_nt<T, 3>(transa, transb, grid, stream, c, a, b, im, in, (long)lda, (long)ldb, (long)ldc, ik, lists, counts, ib, nsub);
        default: return launch_nt<T, 4>(transa, transb, grid, stream, c, a, b, im, in, (long)lda, (long)ldb, (long)ldc, ik, lists, counts, ib, nsub);
    }
}

}  // namespace

int launch_dgemm_non0tab(hipStream_t stream, double *c, const double *a, const double *b, size_t m, size_t n,
                         size_t k, size_t lda, size_t ldb, size_t ldc, bool transa, bool transb, const uint8_t *tab,
                         size_t ldtab, size_t blksize) {
    return launch_gemm_non0tab<double>(stream, c, a, b, m, n, k, lda, ldb, ldc, transa, transb, tab, ldtab, blksize);
}

int launch_sgemm_non0tab(hipStream_t stream, float *c, const float *a, const float *b, size_t m, size_t n, size_t k,
                         size_t lda, size_t ldb, size_t ldc, bool transa, bool transb, const uint8_t *tab,
                         size_t ldtab, size_t blksize) {
    return launch_gemm_non0tab<float>(stream, c, a, b, m, n, k, lda, ldb, ldc, transa, transb, tab, ldtab, blksize);
}

}  // namespace cfgmm
