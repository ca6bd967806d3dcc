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

// Grouped raster order of C tiles: linear index L -> (tile_m, tile_n), walking
// group_m tile rows at a time so that consecutive tiles share A row panels
// and B column panels.
__device__ inline void raster_of(int L, int tiles_m, int tiles_n, int group_m, int &tm, int &tn) {
    const int in_group = group_m * tiles_n;
    const int gid = L / in_group;
    const int first_m = gid * group_m;
    const int gsize = min(tiles_m - first_m, group_m);
    const int r = L % in_group;
    tm = first_m + r % gsize;
    tn = r / gsize;
}

// Workgroup ids are dealt round-robin across the XCDs; remapping pid so that
// each XCD receives a contiguous range of the raster order keeps the panels
// those workgroups share inside one XCD's L2.
__device__ inline int xcd_remap(int pid, int total) {
    const int full = (total / kXcds) * kXcds;
    return pid < full ? (pid % kXcds) * (total / kXcds) + pid / kXcds : pid;
}

// Work decomposition of one launch: workgroups pid < dp_tiles each own a whole
// C tile (the first dp_tiles tiles of the raster order); the remaining tiles
// are split into `splits` K ranges of `kchunk`, one workgroup per (tile,
// range), whose partial tiles go to `partial` and are summed in a fixed order
// by tail_reduce_kernel.
template <int BM, int BN, int BK, int WM, int WN, bool TA, bool TB>
__global__ __launch_bounds__(kThreads, 2) void dgemm_kernel(double *__restrict__ c, const double *__restrict__ a,
                                                            const double *__restrict__ b, int m, int n, int k,
                                                            long lda, long ldb, long ldc, int dp_tiles, int splits,
                                                            int kchunk, int group_m, double *__restrict__ partial) {
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
    const int pid = blockIdx.x;
    int L, z = 0, kbeg = 0, kend = k;
    if (pid < dp_tiles) {
        L = xcd_remap(pid, dp_tiles);
    } else {
        const int q = pid - dp_tiles;
        L = dp_tiles + q / splits;
        z = q % splits;
        kbeg = z * kchunk;
        kend = min(k, kbeg + kchunk);
    }
    int tm, tn;
    raster_of(L, tiles_m, tiles_n, group_m, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
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

    // Epilogue: C += acc (one owner per element), or store the partial tile.  A fragment row's C
    // values are loaded together from clamped, in-bounds addresses, so that the workgroup waits
    // for one load latency per fragment row instead of one per element.
    if (pid < dp_tiles) {
#pragma unroll
        for (int mi = 0; mi < FM; ++mi) {
            double cv[FN][4];
#pragma unroll
            for (int ni = 0; ni < FN; ++ni)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = min(m0 + wm * WM + mi * 16 + lk + 4 * r, m - 1);
                    const int col = min(n0 + wn * WN + ni * 16 + lrow, n - 1);
                    cv[ni][r] = c[(long)row * ldc + col];
                }
#pragma unroll
            for (int ni = 0; ni < FN; ++ni)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = m0 + wm * WM + mi * 16 + lk + 4 * r;
                    const int col = n0 + wn * WN + ni * 16 + lrow;
                    if (row < m && col < n) c[(long)row * ldc + col] = cv[ni][r] + acc[mi][ni][r];
                }
        }
    } else {
        double *__restrict__ pt = partial + ((long)(pid - dp_tiles)) * (BM * BN);  // [tail tile][split][BM][BN]
#pragma unroll
        for (int mi = 0; mi < FM; ++mi)
#pragma unroll
            for (int ni = 0; ni < FN; ++ni)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    pt[(wm * WM + mi * 16 + lk + 4 * r) * BN + wn * WN + ni * 16 + lrow] = acc[mi][ni][r];
    }
}

// C += sum_{z < splits} partial[tail tile][z] over the tail tiles, in
// increasing z (deterministic, no atomics).
template <int BM, int BN>
__global__ void tail_reduce_kernel(double *__restrict__ c, const double *__restrict__ partial, int m, int n, long ldc,
                                   int dp_tiles, int tail_tiles, int splits, int group_m) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)tail_tiles * BM * BN) return;
    const int t = (int)(idx / (BM * BN)), e = (int)(idx % (BM * BN));
    int tm, tn;
    raster_of(dp_tiles + t, (m + BM - 1) / BM, (n + BN - 1) / BN, group_m, tm, tn);
    const int row = tm * BM + e / BN, col = tn * BN + e % BN;
    if (row >= m || col >= n) return;
    const double *__restrict__ p = partial + (long)t * splits * (BM * BN) + e;
    double s = 0.0;
    for (int z = 0; z < splits; ++z) s += p[(long)z * (BM * BN)];
    c[(long)row * ldc + col] += s;
}

struct DgemmPlan {
    int bm;        // 128 (4 waves of 64x64) or 64 (4 waves of 32x32)
    int dp_tiles;  // whole tiles
    int splits;    // K ranges per remaining tile
    int kchunk;    // K range length (multiple of 16)
    int tiles;     // total C tiles
};

// Resident workgroups per CU of the two tile shapes (LDS / VGPR limits).
constexpr int kSlots128 = 2, kSlots64 = 4;

DgemmPlan plan_dgemm(int m, int n, int k, int cus) {
    DgemmPlan p{};
    const long t128 = (long)div_up(m, 128) * div_up(n, 128);
    const long slots128 = (long)kSlots128 * cus;
    auto split_for = [&](long tail, long slots) {
        // enough K ranges to give the tail about one workgroup per free slot,
        // each range at least 256 deep
        long s = std::max(1L, slots / std::max(1L, tail));
        s = std::min<long>(s, std::max(1L, (long)k / 256));
        return (int)s;
    };
    if (t128 >= slots128) {
        p.bm = 128;
        p.tiles = (int)t128;
        const long rem = t128 % slots128;
        if (rem == 0 || rem > slots128 / 2) {
            p.dp_tiles = p.tiles;
            p.splits = 1;
        } else {
            p.dp_tiles = (int)(t128 - rem);
            p.splits = split_for(rem, slots128);
        }
    } else {
        const long t64 = (long)div_up(m, 64) * div_up(n, 64);
        const long slots64 = (long)kSlots64 * cus;
        // Few tiles: 128x128 tiles split along K fill the chip with fewer,
        // larger pieces when that is possible; otherwise 64x64 tiles.  The
        // autotuner may fix the tile edge and the number of K ranges.
        const int force_tile = tune_get(CFGMM_TUNE_DGEMM_SMALL_TILE);
        const int force_splits = tune_get(CFGMM_TUNE_DGEMM_SMALL_SPLITS);
        const int s128 = split_for(t128, slots128);
        if (force_tile == 128 || (force_tile == 0 && t128 * s128 >= slots128 * 3 / 4)) {
            p.bm = 128;
            p.tiles = (int)t128;
        } else {
            p.bm = 64;
            p.tiles = (int)t64;
        }
        const long slots = p.bm == 128 ? slots128 : slots64;
        if (force_splits > 0) {
            p.dp_tiles = force_splits == 1 ? p.tiles : 0;
            p.splits = std::min(force_splits, std::max(1, k / 16));
        } else if (p.tiles >= slots) {
            p.dp_tiles = p.tiles;
            p.splits = 1;
        } else {
            p.dp_tiles = 0;
            p.splits = split_for(p.tiles, slots);
        }
    }
    if (p.splits <= 1) {
        p.splits = 1;
        p.dp_tiles = p.tiles;
        p.kchunk = k;
    } else {
        p.kchunk = (int)(div_up(div_up(k, p.splits), 16) * 16);
        p.splits = (int)div_up(k, p.kchunk);
        if (p.splits <= 1) {
            p.splits = 1;
            p.dp_tiles = p.tiles;
            p.kchunk = k;
        }
    }
    return p;
}

template <int BM, int BN, int WM, int WN>
int launch_plan(hipStream_t stream, const DgemmPlan &pl, double *c, const double *a, const double *b, int m, int n,
                int k, long lda, long ldb, long ldc, bool ta, bool tb) {
    constexpr int BK = 16;
    const int group_m = tune_get(CFGMM_TUNE_DGEMM_GROUP_M);
    const int tail_tiles = pl.tiles - pl.dp_tiles;
    double *partial = nullptr;
    if (tail_tiles > 0) {
        void *ws = nullptr;
        int st = workspace_get(stream, sizeof(double) * (size_t)tail_tiles * pl.splits * BM * BN, &ws);
        if (st != CFGMM_OK) return st;
        partial = static_cast<double *>(ws);
    }
    dim3 grid((unsigned)(pl.dp_tiles + tail_tiles * pl.splits)), block(kThreads);
    hipEvent_t t0 = trace_start(stream);
#define CFGMM_DGEMM_LAUNCH(TA_, TB_)                                                                          \
    hipLaunchKernelGGL((dgemm_kernel<BM, BN, BK, WM, WN, TA_, TB_>), grid, block, 0, stream, c, a, b, m, n, k, \
                       lda, ldb, ldc, pl.dp_tiles, pl.splits, pl.kchunk, group_m, partial)
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
    int st = hip_status(hipGetLastError(), "dgemm_kernel launch");
    if (st != CFGMM_OK || tail_tiles == 0) return st;
    const long total = (long)tail_tiles * BM * BN;
    t0 = trace_start(stream);
    hipLaunchKernelGGL((tail_reduce_kernel<BM, BN>), dim3((unsigned)div_up(total, 256)), dim3(256), 0, stream, c,
                       partial, m, n, ldc, pl.dp_tiles, tail_tiles, pl.splits, group_m);
    trace_stop(stream, "tail_reduce_kernel", t0);
    return hip_status(hipGetLastError(), "tail_reduce_kernel launch");
}

}  // namespace

int launch_dgemm(hipStream_t stream, double *c, const double *a, const double *b, size_t m, size_t n, size_t k,
                 size_t lda, size_t ldb, size_t ldc, bool transa, bool transb) {
    if (m == 0 || n == 0 || k == 0) return CFGMM_OK;
    if (m > INT32_MAX || n > INT32_MAX || k > INT32_MAX) return fail(CFGMM_EINVAL, "dgemm: dimension exceeds 2^31-1");
    const int im = (int)m, in = (int)n, ik = (int)k;
    const DgemmPlan pl = plan_dgemm(im, in, ik, device_cu_count());
    if (pl.bm == 128)
        return launch_plan<128, 128, 64, 64>(stream, pl, c, a, b, im, in, ik, lda, ldb, ldc, transa, transb);
    return launch_plan<64, 64, 32, 32>(stream, pl, c, a, b, im, in, ik, lda, ldb, ldc, transa, transb);
}

}  // namespace cfgmm
