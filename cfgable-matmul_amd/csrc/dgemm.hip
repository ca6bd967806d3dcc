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

typedef double f64x2 __attribute__((ext_vector_type(2)));

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

// 16 bytes per lane from global memory straight into LDS at dst + 16 * lane (dst uniform across
// the wave; global_load_lds_dwordx4).  Device-only: a kernel template that names the builtin in
// its host pass loses its launch stub.
__device__ __attribute__((always_inline)) inline void copy16_to_lds(const void *src, void *dst) {
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void *)dst, 16, 0, 0);
#endif
}

// LDS layouts of one BK = 16 slice of an operand tile of T = BM (= BN) rows or columns.
//   K-contiguous (A not transposed, B transposed: the 16 k of a row are 128 contiguous bytes):
//     row r is 8 chunks of two k; chunk c sits in slot c ^ ((r >> 1) & 7) of the row's 128
//     bytes, so that the 16 rows one ds_read_b128 phase reads cover all 64 banks once.
//   MN-contiguous (A transposed, B not transposed: a k row of T values is contiguous): copy
//     units of 1 KiB (1024 / (8 T) k rows each) at a pitch of 1024 + PAD bytes, which puts the
//     k rows 4 apart that one ds_read_b64 phase reads on disjoint banks.
// Both are what one wave instruction of global_load_lds_dwordx4 writes: 1 KiB, lane i at 16 i.
template <int T>
struct SliceLayout {
    static constexpr int kc_bytes = T * 128;
    static constexpr int rows_per_unit = 1024 / (T * 8);  // k rows per 1 KiB unit: 1 (T 128), 2 (T 64)
    static constexpr int pad = T == 128 ? 32 : 64;
    static constexpr int mn_bytes = (16 / rows_per_unit) * (1024 + pad);
    __device__ static int kc_offset(int r, int kq) {  // byte of element (row r, k kq)
        return r * 128 + ((((kq >> 1) ^ (r >> 1)) & 7) << 4) + ((kq & 1) << 3);
    }
    __device__ static int mn_offset(int kq, int col) {  // byte of element (k kq, column col)
        return (kq / rows_per_unit) * (1024 + pad) + (kq % rows_per_unit) * (T * 8) + col * 8;
    }
};

// Work decomposition of one launch: workgroups pid < dp_tiles each own a whole
// C tile (the first dp_tiles tiles of the raster order); the remaining tiles
// are split into `splits` K ranges of `kchunk`, one workgroup per (tile,
// range), whose partial tiles go to `partial` and are summed in a fixed order
// by tail_reduce_kernel.
//
// Each k step of 16: the waves copy the next step's A and B slices straight into the idle LDS
// buffers (global_load_lds_dwordx4, 1 KiB per wave instruction), then run the MFMAs of the
// current one; one s_waitcnt vmcnt(0) + barrier per step makes the copies of every wave visible
// and the idle buffers free.  The MFMAs take the k of a step in the order k = 4q + kk (lane group
// q, MFMA kk), the same permutation for A and B, so a lane's four A (or B) values of a
// K-contiguous operand are 32 contiguous bytes: two ds_read_b128.  Row (column) indices past the
// matrix are clamped for K-contiguous operands (their results are never stored); an MN-contiguous
// operand whose tile crosses the matrix edge, and the last k step when k is not a multiple of 16,
// go through registers with zero fill instead.
template <int BM, int WM, bool TA, bool TB>
__global__ __launch_bounds__(kThreads, 2) void dgemm_kernel(double *__restrict__ c, const double *__restrict__ a,
                                                            const double *__restrict__ b, int m, int n, int k,
                                                            long lda, long ldb, long ldc, int dp_tiles, int splits,
                                                            int kchunk, int group_m, double *__restrict__ partial) {
    constexpr int BN = BM, WN = WM, BK = 16;
    constexpr int WAVES_N = BN / WN;
    static_assert((BM / WM) * WAVES_N == kThreads / 64, "four waves per workgroup");
    constexpr int FM = WM / 16, FN = WN / 16;
    using L = SliceLayout<BM>;
    constexpr int A_BYTES = TA ? L::mn_bytes : L::kc_bytes;
    constexpr int B_BYTES = TB ? L::kc_bytes : L::mn_bytes;
    // the two stages are __shared__ objects of their own: the compiler's wait before reading one
    // then does not also wait for the copies into the other
    __shared__ __attribute__((aligned(16))) unsigned char As0[A_BYTES];
    __shared__ __attribute__((aligned(16))) unsigned char As1[A_BYTES];
    __shared__ __attribute__((aligned(16))) unsigned char Bs0[B_BYTES];
    __shared__ __attribute__((aligned(16))) unsigned char Bs1[B_BYTES];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave / WAVES_N, wn = wave % WAVES_N;
    const int lrow = lane & 15, lk = lane >> 4;
    const int tiles_m = (m + BM - 1) / BM, tiles_n = (n + BN - 1) / BN;
    const int pid = blockIdx.x;
    int Lt, z = 0, kbeg = 0, kend = k;
    if (pid < dp_tiles) {
        Lt = xcd_remap(pid, dp_tiles);
    } else {
        const int q = pid - dp_tiles;
        Lt = dp_tiles + q / splits;
        z = q % splits;
        kbeg = z * kchunk;
        kend = min(k, kbeg + kchunk);
    }
    int tm, tn;
    raster_of(Lt, tiles_m, tiles_n, group_m, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
    const int klen = kend > kbeg ? kend - kbeg : 0;
    const int nfull = klen / BK;
    const bool ktail = klen % BK != 0;

    // ---- direct copies of k step t (full steps only) ----
    // K-contiguous operand x (row-major rows of length >= k): BM rows x 8 chunks, BM / 32 wave
    // instructions per wave; instruction u covers rows 8u .. 8u+7
    auto issue_kc = [&](const double *__restrict__ x, long ld, int row0, int nrows, int k0, unsigned char *dst) {
#pragma unroll
        for (int e = 0; e < BM / 32; ++e) {
            const int u = wave * (BM / 32) + e;
            const int P = u * 64 + lane;
            const int r = P >> 3, slot = P & 7;
            const int ch = (slot ^ (r >> 1)) & 7;
            const int gr = min(row0 + r, nrows - 1);
            copy16_to_lds(x + (long)gr * ld + k0 + 2 * ch, dst + u * 1024);
        }
    };
    // MN-contiguous operand x (k rows of stride ld): 16 / rows_per_unit units, a quarter per wave
    auto issue_mn = [&](const double *__restrict__ x, long ld, int col0, int k0, unsigned char *dst) {
        constexpr int UNITS = 16 / L::rows_per_unit;
#pragma unroll
        for (int e = 0; e < UNITS / 4; ++e) {
            const int u = wave * (UNITS / 4) + e;
            const int byte = lane * 16;
            const int kr = u * L::rows_per_unit + byte / (BM * 8);
            const int col = col0 + (byte % (BM * 8)) / 8;
            copy16_to_lds(x + (long)(k0 + kr) * ld + col, dst + u * (1024 + L::pad));
        }
    };
    auto issue = [&](int t, unsigned char *as, unsigned char *bs) {
        const int k0 = kbeg + t * BK;
        if (TA) issue_mn(a, lda, m0, k0, as);
        else issue_kc(a, lda, m0, m, k0, as);
        if (TB) issue_kc(b, ldb, n0, n, k0, bs);
        else issue_mn(b, ldb, n0, k0, bs);
    };

    // ---- register path (zero fill): element e of a thread covers one (row or k row, k or column) ----
    constexpr int PER_T = BM * BK / kThreads;
    double ra[PER_T], rb[PER_T];
    auto load_regs = [&](int k0) {
#pragma unroll
        for (int e = 0; e < PER_T; ++e) {
            const int idx = tid + kThreads * e;
            double v = 0.0;
            if (TA) {  // a[p][i]: k row idx / BM, column idx % BM
                const int p = k0 + idx / BM, i = m0 + idx % BM;
                if (p < kend && i < m) v = a[(long)p * lda + i];
            } else {   // a[i][p]: row idx / 16, k idx % 16
                const int i = m0 + idx / BK, p = k0 + idx % BK;
                if (p < kend && i < m) v = a[(long)i * lda + p];
            }
            ra[e] = v;
            v = 0.0;
            if (TB) {  // b[j][p]
                const int j = n0 + idx / BK, p = k0 + idx % BK;
                if (p < kend && j < n) v = b[(long)j * ldb + p];
            } else {   // b[p][j]
                const int p = k0 + idx / BN, j = n0 + idx % BN;
                if (p < kend && j < n) v = b[(long)p * ldb + j];
            }
            rb[e] = v;
        }
    };
    auto store_regs = [&](unsigned char *as, unsigned char *bs) {
#pragma unroll
        for (int e = 0; e < PER_T; ++e) {
            const int idx = tid + kThreads * e;
            const int ao = TA ? L::mn_offset(idx / BM, idx % BM) : L::kc_offset(idx / BK, idx % BK);
            const int bo = TB ? L::kc_offset(idx / BK, idx % BK) : L::mn_offset(idx / BN, idx % BN);
            *reinterpret_cast<double *>(as + ao) = ra[e];
            *reinterpret_cast<double *>(bs + bo) = rb[e];
        }
    };

    f64x4 acc[FM][FN];
#pragma unroll
    for (int mi = 0; mi < FM; ++mi)
#pragma unroll
        for (int ni = 0; ni < FN; ++ni) acc[mi][ni] = f64x4{0.0, 0.0, 0.0, 0.0};

    // ---- one k step from LDS: lane group q supplies k = 4q + kk to MFMA kk; two halves of two kk
    // each, so that only half the fragments are live at a time ----
    auto compute = [&](const unsigned char *as, const unsigned char *bs) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            f64x2 fa[FM], fb[FN];
#pragma unroll
            for (int mi = 0; mi < FM; ++mi) {
                const int r = wm * WM + mi * 16 + lrow;
                if (TA) {
                    fa[mi][0] = *reinterpret_cast<const double *>(as + L::mn_offset(4 * lk + 2 * h, r));
                    fa[mi][1] = *reinterpret_cast<const double *>(as + L::mn_offset(4 * lk + 2 * h + 1, r));
                } else {
                    fa[mi] = *reinterpret_cast<const f64x2 *>(as + L::kc_offset(r, 4 * lk + 2 * h));
                }
            }
#pragma unroll
            for (int ni = 0; ni < FN; ++ni) {
                const int col = wn * WN + ni * 16 + lrow;
                if (TB) {
                    fb[ni] = *reinterpret_cast<const f64x2 *>(bs + L::kc_offset(col, 4 * lk + 2 * h));
                } else {
                    fb[ni][0] = *reinterpret_cast<const double *>(bs + L::mn_offset(4 * lk + 2 * h, col));
                    fb[ni][1] = *reinterpret_cast<const double *>(bs + L::mn_offset(4 * lk + 2 * h + 1, col));
                }
            }
#pragma unroll
            for (int kk = 0; kk < 2; ++kk)
#pragma unroll
                for (int mi = 0; mi < FM; ++mi)
#pragma unroll
                    for (int ni = 0; ni < FN; ++ni)
                        acc[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[mi][kk], fb[ni][kk], acc[mi][ni], 0, 0, 0);
        }
    };
    // every wave's copies landed (vmcnt(0): expcnt and lgkmcnt left at their maxima) and every
    // wave is done with the buffers read before; an LDS-only barrier, since __syncthreads()
    // would add a fence.  (The scheduler sinks some of the step's MFMAs below the wait; pinning
    // them above it with a scheduling barrier measured no faster here, the second workgroup on
    // the CU covering the copies.)
    auto sync = [&]() {
        __builtin_amdgcn_s_waitcnt((0x7 << 4) | (0xF << 8));
        asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    };

    // an MN-contiguous operand tile that crosses the matrix edge cannot be copied in 16-byte pairs
    const bool direct = (!TA || m0 + BM <= m) && (TB || n0 + BN <= n);
    if (direct) {
        if (nfull > 0) {
            const int last = nfull - 1;
            issue(0, As0, Bs0);
            int t = 0;
            // the prefetch index is clamped rather than guarded (a guard makes the compiler wait
            // for every copy at its join): past the last step it re-copies that step into the idle
            // buffers, whose contents are not used
            for (; t + 2 <= nfull; t += 2) {
                sync();
                issue(min(t + 1, last), As1, Bs1);
                compute(As0, Bs0);
                sync();
                issue(min(t + 2, last), As0, Bs0);
                compute(As1, Bs1);
            }
            if (t < nfull) {
                sync();
                issue(last, As1, Bs1);
                compute(As0, Bs0);
            }
            sync();  // no copy outlives the loop; the buffers are free
        }
        if (ktail) {
            load_regs(kbeg + nfull * BK);
            store_regs(As0, Bs0);
            __syncthreads();
            compute(As0, Bs0);
        }
    } else {
        // registers for every step (edge tiles only: the loads are not overlapped with the MFMAs,
        // which keeps the staging registers out of the MFMA loop); the two buffers alternate, so
        // one barrier per step separates the writes of a buffer from the reads two steps before
        const int nsteps = nfull + (ktail ? 1 : 0);
        for (int t = 0; t < nsteps; ++t) {
            unsigned char *as = (t & 1) ? As1 : As0, *bs = (t & 1) ? Bs1 : Bs0;
            load_regs(kbeg + t * BK);
            store_regs(as, bs);
            __syncthreads();
            compute(as, bs);
        }
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

// Small GEMMs with both operands K-contiguous (A not transposed, B transposed): one workgroup of
// KG groups of four waves per BM x BN tile of C, at most one tile per CU.  Group g takes the k
// steps KG i + g; each group copies its own A and B slices straight into its own pair of LDS
// stages (the layout of the large kernel's K-contiguous operands), and every wave of a group owns
// a 16-row strip of the tile.  The groups' sums are added in group order through LDS at the end:
// no partial tiles in memory and no second launch, which the K split of the large kernel needs
// for shapes with fewer tiles than workgroup slots.  The steps past the last KG-multiple (and the
// k tail) go through registers with zero fill.
__device__ inline int kc_off(int r, int kq) {
    return r * 128 + ((((kq >> 1) ^ (r >> 1)) & 7) << 4) + ((kq & 1) << 3);
}

template <int BN, int KG>
__global__ __launch_bounds__(256 * KG) void dgemm_kgroup_kernel(double *__restrict__ c, const double *__restrict__ a,
                                                               const double *__restrict__ b, int m, int n, int k,
                                                               long lda, long ldb, long ldc, int group_m) {
    constexpr int BM = 64, BK = 16, NW = 4;  // four 16-row strips per group
    constexpr int FN = BN / 16;
    constexpr int A_BYTES = BM * 128, B_BYTES = BN * 128;
    constexpr int UNITS = (BM + BN) / 8;  // 1 KiB copy instructions per slice pair
    constexpr int PER_T = (BM + BN) * BK / 256;
    static_assert(BN % 16 == 0 && (BM + BN) * BK % 256 == 0, "tile shape");
    // the two stages are __shared__ objects of their own (see dgemm_kernel)
    __shared__ __attribute__((aligned(16))) unsigned char S0[KG][A_BYTES + B_BYTES];
    __shared__ __attribute__((aligned(16))) unsigned char S1[KG][A_BYTES + B_BYTES];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = wave / NW, sw = wave % NW, gt = tid % 256;  // group, strip, thread in group
    const int lrow = lane & 15, lk = lane >> 4;
    const int tiles_m = (m + BM - 1) / BM, tiles_n = (n + BN - 1) / BN;
    int tm, tn;  // each XCD takes a contiguous range of the raster order (the panels it shares stay in its L2)
    raster_of(xcd_remap(blockIdx.x, tiles_m * tiles_n), tiles_m, tiles_n, group_m, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
    const int nfull = k / BK;
    const int R = nfull / KG;                    // rounds in which every group has a full step
    const bool tail = R * KG * BK < k;           // one more round through registers

    auto issue = [&](int t, unsigned char *dst) {
        const int k0 = t * BK;
        for (int u = sw; u < UNITS; u += NW) {
            const bool isa = u < BM / 8;
            const int uu = isa ? u : u - BM / 8;
            const int P = uu * 64 + lane;
            const int r = P >> 3, ch = ((P & 7) ^ (r >> 1)) & 7;
            const double *src = isa ? a + (long)min(m0 + r, m - 1) * lda : b + (long)min(n0 + r, n - 1) * ldb;
            copy16_to_lds(src + k0 + 2 * ch, dst + (isa ? 0 : A_BYTES) + uu * 1024);
        }
    };
    f64x4 acc[FN];
#pragma unroll
    for (int ni = 0; ni < FN; ++ni) acc[ni] = f64x4{0.0, 0.0, 0.0, 0.0};
    auto compute = [&](const unsigned char *st) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const f64x2 fa = *reinterpret_cast<const f64x2 *>(st + kc_off(sw * 16 + lrow, 4 * lk + 2 * h));
            f64x2 fb[FN];
#pragma unroll
            for (int ni = 0; ni < FN; ++ni)
                fb[ni] = *reinterpret_cast<const f64x2 *>(st + A_BYTES + kc_off(ni * 16 + lrow, 4 * lk + 2 * h));
#pragma unroll
            for (int kk = 0; kk < 2; ++kk)
#pragma unroll
                for (int ni = 0; ni < FN; ++ni)
                    acc[ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[kk], fb[ni][kk], acc[ni], 0, 0, 0);
        }
    };
    auto sync = [&]() {  // as in dgemm_kernel, the step's MFMAs pinned above the wait (one workgroup per CU)
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_waitcnt((0x7 << 4) | (0xF << 8));
        asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    };

    if (R > 0) {
        const int last = R - 1;
        issue(g, S0[g]);
        int i = 0;
        // prefetch indices clamped rather than guarded (see dgemm_kernel)
        for (; i + 2 <= R; i += 2) {
            sync();
            issue(min(i + 1, last) * KG + g, S1[g]);
            compute(S0[g]);
            sync();
            issue(min(i + 2, last) * KG + g, S0[g]);
            compute(S1[g]);
        }
        if (i < R) {
            sync();
            issue(last * KG + g, S1[g]);
            compute(S0[g]);
        }
        sync();
    }
    if (tail) {
        // step R KG + g of each group through registers, zero past k (and past the matrix)
        const int k0 = (R * KG + g) * BK;
        double v[PER_T];
#pragma unroll
        for (int e = 0; e < PER_T; ++e) {
            const int idx = gt + 256 * e;
            const bool isa = idx < BM * BK;
            const int id2 = isa ? idx : idx - BM * BK;
            const int r = id2 / BK, p = k0 + id2 % BK;
            const int row = (isa ? m0 : n0) + r;
            const bool ok = p < k && row < (isa ? m : n);
            v[e] = ok ? (isa ? a[(long)row * lda + p] : b[(long)row * ldb + p]) : 0.0;
        }
#pragma unroll
        for (int e = 0; e < PER_T; ++e) {
            const int idx = gt + 256 * e;
            const bool isa = idx < BM * BK;
            const int id2 = isa ? idx : idx - BM * BK;
            *reinterpret_cast<double *>(S0[g] + (isa ? 0 : A_BYTES) + kc_off(id2 / BK, id2 % BK)) = v[e];
        }
        __syncthreads();
        compute(S0[g]);
        __syncthreads();
    }

    // groups 1 .. KG-1 hand their sums to group 0 through LDS, one group at a time, in group order
    double *red = reinterpret_cast<double *>(&S1[0][0]);  // NW x FN x 4 x 64 doubles
    static_assert(NW * FN * 4 * 64 * 8 <= KG * (A_BYTES + B_BYTES), "reduction buffer");
    for (int g2 = 1; g2 < KG; ++g2) {
        if (g == g2) {
#pragma unroll
            for (int ni = 0; ni < FN; ++ni)
#pragma unroll
                for (int r = 0; r < 4; ++r) red[((sw * FN + ni) * 4 + r) * 64 + lane] = acc[ni][r];
        }
        __syncthreads();
        if (g == 0) {
#pragma unroll
            for (int ni = 0; ni < FN; ++ni)
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[ni][r] += red[((sw * FN + ni) * 4 + r) * 64 + lane];
        }
        __syncthreads();
    }
    if (g != 0) return;
    // C += acc: a strip row's C values loaded together from clamped addresses, then stored
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int row = m0 + sw * 16 + lk + 4 * r;
        double cv[FN];
#pragma unroll
        for (int ni = 0; ni < FN; ++ni) cv[ni] = c[(long)min(row, m - 1) * ldc + min(n0 + ni * 16 + lrow, n - 1)];
#pragma unroll
        for (int ni = 0; ni < FN; ++ni) {
            const int col = n0 + ni * 16 + lrow;
            if (row < m && col < n) c[(long)row * ldc + col] = cv[ni] + acc[ni][r];
        }
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
    // eight loads in flight at a time (clamped indices, no branch around a load), summed in z order
    double s = 0.0;
    for (int z0 = 0; z0 < splits; z0 += 8) {
        double v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = p[(long)min(z0 + i, splits - 1) * (BM * BN)];
#pragma unroll
        for (int i = 0; i < 8; ++i) s = z0 + i < splits ? s + v[i] : s;
    }
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
        // Few tiles: the autotuner may fix the tile edge and the number of K ranges; otherwise the
        // (edge, ranges) of least modelled time.  Model (fitted to 18 shapes of 256..2000 rows and
        // columns and k 256..6000 on MI355X, within 2 % of the best plan on average): the largest
        // per-CU share of workgroups times one workgroup's multiply-adds at the rate a CU sustains
        // with that many resident (128 x 128: 128 G/s alone, 137 G/s with two; 64 x 64: 91, 123,
        // 131, 128 G/s with one to four), plus, with more than one range, the partial tiles read
        // again by tail_reduce_kernel at 4.5 TB/s and its launch (1.4 us).
        const int force_tile = tune_get(CFGMM_TUNE_DGEMM_SMALL_TILE);
        const int force_splits = tune_get(CFGMM_TUNE_DGEMM_SMALL_SPLITS);
        auto model_us = [&](int bm, int s) {
            const long tiles = (long)div_up(m, bm) * div_up(n, bm);
            const long kc = div_up(div_up(k, s), 16) * 16;
            const long sp = div_up(k, kc);
            const long per_cu = div_up(tiles * sp, (long)cus);
            static const double us_per_kmac128[2] = {0.0078, 0.0073};
            static const double us_per_kmac64[4] = {0.0110, 0.0081, 0.0077, 0.0078};
            const double rate = bm == 128 ? us_per_kmac128[std::min(per_cu, 2L) - 1]
                                          : us_per_kmac64[std::min(per_cu, 4L) - 1];
            double t = rate * (double)per_cu * bm * bm * kc / 1e3;
            if (sp > 1) t += 0.2224 * (double)tiles * sp * bm * bm * 8 / 1e6 + 1.4;
            return t;
        };
        const int s_cap = std::max(1, k / 16);  // ranges at least 16 deep
        const int s_lo = force_splits > 0 ? std::min(force_splits, s_cap) : 1;
        const int s_hi = force_splits > 0 ? s_lo : std::min(16, s_cap);
        int best_bm = 0, best_s = 1;
        double best_t = 0.0;
        for (int bm : {128, 64}) {
            if (force_tile != 0 && bm != force_tile) continue;
            for (int s = s_lo; s <= s_hi; ++s) {
                const double t = model_us(bm, s);
                if (best_bm == 0 || t < best_t) {
                    best_t = t;
                    best_bm = bm;
                    best_s = s;
                }
            }
        }
        p.bm = best_bm;
        p.tiles = (int)((long)div_up(m, best_bm) * div_up(n, best_bm));
        p.splits = best_s;
        p.dp_tiles = best_s == 1 ? p.tiles : 0;
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

template <int BM, int WM>
int launch_plan(hipStream_t stream, const DgemmPlan &pl, double *c, const double *a, const double *b, int m, int n,
                int k, long lda, long ldb, long ldc, bool ta, bool tb) {
    constexpr int BN = BM;
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
    hipLaunchKernelGGL((dgemm_kernel<BM, WM, TA_, TB_>), grid, block, 0, stream, c, a, b, m, n, k,             \
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
    const int cus = device_cu_count();
    // too few 128x128 tiles to fill the chip, both operands K-contiguous: the in-workgroup K split
    // with the narrowest tile width that gives every tile a CU of its own, if the tiles then cover
    // at least 3/4 of the CUs.  By default only up to k = 1024: deeper, the K split over
    // workgroups is 8 - 25 % faster on such shapes (it keeps more waves per CU busy), shallower
    // its partial tiles and second launch cost more than the lower rate here (measured on MI355X).
    const int kgroups = tune_get(CFGMM_TUNE_DGEMM_KGROUPS);
    if (!transa && transb && (kgroups == 2 || (kgroups == 0 && k <= 1024)) &&
        (long)div_up(m, 128) * div_up(n, 128) < (long)kSlots128 * cus) {
        const long tm = (long)div_up(m, 64);
        for (int bn : {64, 80, 96}) {
            const long tiles = tm * (long)div_up(n, (size_t)bn);
            if (tiles > cus) continue;
            if (4 * tiles < 3L * cus) break;
            const int group_m = tune_get(CFGMM_TUNE_DGEMM_GROUP_M);
            hipEvent_t t0 = trace_start(stream);
            if (bn == 64)
                hipLaunchKernelGGL((dgemm_kgroup_kernel<64, 4>), dim3((unsigned)tiles), dim3(1024), 0, stream, c, a, b,
                                   im, in, ik, (long)lda, (long)ldb, (long)ldc, group_m);
            else if (bn == 80)
                hipLaunchKernelGGL((dgemm_kgroup_kernel<80, 4>), dim3((unsigned)tiles), dim3(1024), 0, stream, c, a, b,
                                   im, in, ik, (long)lda, (long)ldb, (long)ldc, group_m);
            else
                hipLaunchKernelGGL((dgemm_kgroup_kernel<96, 4>), dim3((unsigned)tiles), dim3(1024), 0, stream, c, a, b,
                                   im, in, ik, (long)lda, (long)ldb, (long)ldc, group_m);
            trace_stop(stream, "dgemm_kgroup_kernel", t0);
            return hip_status(hipGetLastError(), "dgemm_kgroup_kernel launch");
        }
    }
    const DgemmPlan pl = plan_dgemm(im, in, ik, cus);
    if (pl.bm == 128)
        return launch_plan<128, 64>(stream, pl, c, a, b, im, in, ik, lda, ldb, ldc, transa, transb);
    return launch_plan<64, 32>(stream, pl, c, a, b, im, in, ik, lda, ldb, ldc, transa, transb);
}

}  // namespace cfgmm
