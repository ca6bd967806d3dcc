// dgemm.hip -- row-major f64 GEMM  C += op(A) op(B)  on FP64 MFMA (gfx950).
//
// Replaces the CPU loop nest of MatmulLoops::matmul_loop_parallel_mnk_pack_a
// (src/impl_matmul.rs:155-231): the A pack (184-205), the per-NR-strip B pack
// (77-117, 147-150) and the MR x NR register microkernel (src/structs.rs:34-45)
// become one kernel in which each workgroup owns a BM x BN tile of C, streams
// BK-deep slices of op(A) and op(B) through double-buffered LDS, and
// accumulates with v_mfma_f64_16x16x4_f64.  The reference's per-tile Mutex
// (src/impl_matmul.rs:26) is unnecessary: each C element is owned by exactly
// one workgroup (or, with split-K, by one deterministic reduction thread).
//
// MFMA operand layout of v_mfma_f64_16x16x4_f64 (wave64):
//   A (16x4):  lane l holds A[l%16][l/16]
//   B (4x16):  lane l holds B[l/16][l%16]
//   D (16x16): lane l, element r holds D[l/16 + 4*r][l%16]
// (measured on gfx950 with one-hot operands: each output register r covers
// rows r*4 .. r*4+3, one row per 16-lane group).
#include "internal.hpp"

#include <algorithm>

namespace cfgmm {
namespace {

constexpr int kThreads = 256;
constexpr int kXcds = 8;  // MI355X: 8 XCDs, workgroups are dealt round-robin across them

// Maps a linear workgroup id to a (tile_m, tile_n) pair so that the
// workgroups resident on one XCD cover a compact group of C tiles (sharing
// A row panels and B column panels in that XCD's L2).
__device__ inline void tile_of(int pid, int total, int tiles_m, int tiles_n, int group_m, int &tm, int &tn) {
    const int full = (total / kXcds) * kXcds;
    if (pid < full) pid = (pid % kXcds) * (total / kXcds) + pid / kXcds;
    const int in_group = group_m * tiles_n;
    const int gid = pid / in_group;
    const int first_m = gid * group_m;
    const int gsize = min(tiles_m - first_m, group_m);
    const int r = pid % in_group;
    tm = first_m + r % gsize;
    tn = r / gsize;
}

template <int BM, int BN, int BK, int WM, int WN, bool TA, bool TB>
__global__ __launch_bounds__(kThreads, 2) void dgemm_kernel(double *__restrict__ c, const double *__restrict__ a,
                                                            const double *__restrict__ b, int m, int n, int k,
                                                            long lda, long ldb, long ldc, int kchunk, int group_m,
                                                            double *__restrict__ partial) {
    constexpr int WAVES_N = BN / WN;
    static_assert((BM / WM) * WAVES_N == kThreads / 64, "four waves per workgroup");
    constexpr int FM = WM / 16, FN = WN / 16;
    // Pad the LDS rows only where the global->LDS copy transposes (store
    // stride = row length), to spread those stores over the banks.
    constexpr int LDA_S = BM + (TA ? 0 : 1);
    constexpr int LDB_S = BN + (TB ? 1 : 0);
    constexpr int A_PER_T = BM * BK / kThreads;
    constexpr int B_PER_T = BN * BK / kThreads;
    static_assert(A_PER_T * kThreads == BM * BK && B_PER_T * kThreads == BN * BK, "tile / thread split");

    __shared__ double As[2][BK][LDA_S];
    __shared__ double Bs[2][BK][LDB_S];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave / WAVES_N, wn = wave % WAVES_N;
    const int tiles_m = (m + BM - 1) / BM, tiles_n = (n + BN - 1) / BN;
    int tm, tn;
    tile_of(blockIdx.x, gridDim.x, tiles_m, tiles_n, group_m, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
    const int kbeg = blockIdx.z * kchunk;
    const int kend = min(k, kbeg + kchunk);
    const int ntiles = kend > kbeg ? (kend - kbeg + BK - 1) / BK : 0;

    double ra[A_PER_T], rb[B_PER_T];

    auto load_tiles = [&](int k0) {
#pragma unroll
        for (int e = 0; e < A_PER_T; ++e) {
            const int idx = tid + kThreads * e;
            const int p = TA ? idx / BM : idx % BK;
            const int i = TA ? idx % BM : idx / BK;
            const int gi = m0 + i, gp = k0 + p;
            double v = 0.0;
            if (gi < m && gp < kend) v = TA ? a[(long)gp * lda + gi] : a[(long)gi * lda + gp];
            ra[e] = v;
        }
#pragma unroll
        for (int e = 0; e < B_PER_T; ++e) {
            const int idx = tid + kThreads * e;
            const int p = TB ? idx % BK : idx / BN;
            const int j = TB ? idx / BK : idx % BN;
            const int gj = n0 + j, gp = k0 + p;
            double v = 0.0;
            if (gj < n && gp < kend) v = TB ? b[(long)gj * ldb + gp] : b[(long)gp * ldb + gj];
            rb[e] = v;
        }
    };
    auto store_tiles = [&](int buf) {
#pragma unroll
        for (int e = 0; e < A_PER_T; ++e) {
            const int idx = tid + kThreads * e;
            const int p = TA ? idx / BM : idx % BK;
            const int i = TA ? idx % BM : idx / BK;
            As[buf][p][i] = ra[e];
        }
#pragma unroll
        for (int e = 0; e < B_PER_T; ++e) {
            const int idx = tid + kThreads * e;
            const int p = TB ? idx % BK : idx / BN;
            const int j = TB ? idx / BK : idx % BN;
            Bs[buf][p][j] = rb[e];
        }
    };

    f64x4 acc[FM][FN];
#pragma unroll
    for (int mi = 0; mi < FM; ++mi)
#pragma unroll
        for (int ni = 0; ni < FN; ++ni) acc[mi][ni] = f64x4{0.0, 0.0, 0.0, 0.0};

    if (ntiles > 0) {
        load_tiles(kbeg);
        store_tiles(0);
    }
    __syncthreads();

    const int lrow = lane & 15, lk = lane >> 4;
    for (int t = 0; t < ntiles; ++t) {
        const int cur = t & 1;
        if (t + 1 < ntiles) load_tiles(kbeg + (t + 1) * BK);
#pragma unroll
        for (int kk = 0; kk < BK / 4; ++kk) {
            double fa[FM], fb[FN];
#pragma unroll
            for (int mi = 0; mi < FM; ++mi) fa[mi] = As[cur][kk * 4 + lk][wm * WM + mi * 16 + lrow];
#pragma unroll
            for (int ni = 0; ni < FN; ++ni) fb[ni] = Bs[cur][kk * 4 + lk][wn * WN + ni * 16 + lrow];
#pragma unroll
            for (int mi = 0; mi < FM; ++mi)
#pragma unroll
                for (int ni = 0; ni < FN; ++ni)
                    acc[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[mi], fb[ni], acc[mi][ni], 0, 0, 0);
        }
        if (t + 1 < ntiles) store_tiles(cur ^ 1);
        __syncthreads();
    }

    // Epilogue: C += acc (one owner per element), or write the split-K partial.
#pragma unroll
    for (int mi = 0; mi < FM; ++mi)
#pragma unroll
        for (int ni = 0; ni < FN; ++ni)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm * WM + mi * 16 + lk + 4 * r;
                const int col = n0 + wn * WN + ni * 16 + lrow;
                if (row < m && col < n) {
                    if (partial)
                        partial[((long)blockIdx.z * m + row) * n + col] = acc[mi][ni][r];
                    else
                        c[(long)row * ldc + col] += acc[mi][ni][r];
                }
            }
}

// C[i][j] += sum_{z < splits} partial[z][i][j], in increasing z (deterministic).
__global__ void splitk_reduce_kernel(double *__restrict__ c, const double *__restrict__ partial, int m, int n,
                                     long ldc, int splits) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)m * n;
    if (idx >= total) return;
    const int row = (int)(idx / n), col = (int)(idx % n);
    double s = 0.0;
    for (int z = 0; z < splits; ++z) s += partial[(long)z * total + idx];
    c[(long)row * ldc + col] += s;
}

struct TileCfg {
    int bm, bn;
};

template <int BM, int BN, int WM, int WN>
int launch_cfg(hipStream_t stream, double *c, const double *a, const double *b, int m, int n, int k, long lda,
               long ldb, long ldc, bool ta, bool tb, int splits, int kchunk, double *partial) {
    constexpr int BK = 16;
    const int tiles = (int)(div_up(m, BM) * div_up(n, BN));
    dim3 grid(tiles, 1, splits), block(kThreads);
    const int group_m = 4;
    hipEvent_t t0 = trace_start(stream);
#define CFGMM_DGEMM_LAUNCH(TA_, TB_)                                                                          \
    hipLaunchKernelGGL((dgemm_kernel<BM, BN, BK, WM, WN, TA_, TB_>), grid, block, 0, stream, c, a, b, m, n, k, \
                       lda, ldb, ldc, kchunk, group_m, partial)
    if (!ta && !tb)
        CFGMM_DGEMM_LAUNCH(false, false);
    else if (!ta && tb)
        CFGMM_DGEMM_LAUNCH(false, true);
    else if (ta && !tb)
        CFGMM_DGEMM_LAUNCH(true, false);
    else
        CFGMM_DGEMM_LAUNCH(true, true);
#undef CFGMM_DGEMM_LAUNCH
    trace_stop(stream, "dgemm_kernel", t0);
    return hip_status(hipGetLastError(), "dgemm_kernel launch");
}

}  // namespace

int launch_dgemm(hipStream_t stream, double *c, const double *a, const double *b, size_t m, size_t n, size_t k,
                 size_t lda, size_t ldb, size_t ldc, bool transa, bool transb) {
    if (m == 0 || n == 0 || k == 0) return CFGMM_OK;
    if (m > INT32_MAX || n > INT32_MAX || k > INT32_MAX) return fail(CFGMM_EINVAL, "dgemm: dimension exceeds 2^31-1");
    const int cus = device_cu_count();
    const int im = (int)m, in = (int)n, ik = (int)k;
    // Tile choice: 128x128 when that already gives about two waves of
    // workgroups per CU; otherwise 64x64, split along K until the grid fills
    // the chip (partials reduced in a fixed order afterwards).
    const long big_tiles = (long)div_up(m, 128) * div_up(n, 128);
    if (big_tiles >= 2L * cus)
        return launch_cfg<128, 128, 64, 64>(stream, c, a, b, im, in, ik, lda, ldb, ldc, transa, transb, 1, ik,
                                            nullptr);
    const long small_tiles = (long)div_up(m, 64) * div_up(n, 64);
    int splits = 1;
    while (small_tiles * splits < 4L * cus && (long)k / (splits * 2) >= 256) splits *= 2;
    if (splits == 1)
        return launch_cfg<64, 64, 32, 32>(stream, c, a, b, im, in, ik, lda, ldb, ldc, transa, transb, 1, ik,
                                          nullptr);
    const int kchunk = (int)(div_up(div_up(k, splits), 16) * 16);
    splits = (int)div_up(k, kchunk);
    void *ws = nullptr;
    int st = workspace_get(stream, sizeof(double) * (size_t)splits * m * n, &ws);
    if (st != CFGMM_OK) return st;
    double *partial = static_cast<double *>(ws);
    st = launch_cfg<64, 64, 32, 32>(stream, c, a, b, im, in, ik, lda, ldb, ldc, transa, transb, splits, kchunk,
                                    partial);
    if (st != CFGMM_OK) return st;
    const long total = (long)m * n;
    hipEvent_t t0 = trace_start(stream);
    hipLaunchKernelGGL(splitk_reduce_kernel, dim3((unsigned)div_up(total, 256)), dim3(256), 0, stream, c, partial,
                       im, in, (long)ldc, splits);
    trace_stop(stream, "splitk_reduce_kernel", t0);
    return hip_status(hipGetLastError(), "splitk_reduce_kernel launch");
}

}  // namespace cfgmm
