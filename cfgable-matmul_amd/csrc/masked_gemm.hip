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

#include <type_traits>

namespace cfgmm {

// Zero sources for loads outside the problem or the compacted K list.
__device__ const double kZerosM[16] = {};
__device__ const float kZerosMf[16] = {};

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

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
constexpr int kXcds = 8;  // MI355X: 8 XCDs, each with its own L2 and 4 shader engines
constexpr int kSes = 4;

// Task (column subtile, row tile) of this workgroup on a 1-D grid of ncolt x (row tiles) tasks,
// column subtile fastest.  The dispatcher deals workgroup i to XCD i % 8 and, within it, to shader
// engine (i / 8) % 4, in order: when that engine has no free slot, the XCD's later workgroups wait
// too.  So each (XCD, engine) pair takes a contiguous run of the task order, the four runs of an
// XCD adjacent: every engine then gets the same mix of long and short lists whatever their
// pattern over the column blocks (dealing consecutive tasks round-robin would put every 32nd
// column block on one engine), and the workgroups of an XCD that run together share the op(A)
// rows of one or two row tiles in its L2.
__device__ inline void masked_task(int ncolt, int &colt, int &mt) {
    const int pid = blockIdx.x, total = gridDim.x;
    constexpr int kClasses = kXcds * kSes;
    const int per = total / kClasses;
    const int v = pid < per * kClasses ? ((pid % kXcds) * kSes + (pid / kXcds) % kSes) * per + pid / kClasses : pid;
    colt = v % ncolt;
    mt = v / ncolt;
}
constexpr int kBM = 128;  // rows per workgroup (4 waves x 32)
constexpr int kBK = 16;   // compacted-K slice depth
constexpr int kSeg = 2048;  // compacted-K list entries staged in LDS at a time

template <typename T, int NT, bool TA, bool TB>
__global__ __launch_bounds__(kThreads, 2) void masked_gemm_kernel(T *__restrict__ c, const T *__restrict__ a,
                                                                  const T *__restrict__ b, int m, int n,
                                                                  long lda, long ldb, long ldc, int kmax,
                                                                  const int *__restrict__ lists,
                                                                  const int *__restrict__ counts, int blksize,
                                                                  int nsub, int ncolt) {
    constexpr int W = NT * 16;
    constexpr int FM = 2;  // 32 rows per wave
    constexpr int LDA_S = kBM + (TA ? 0 : 1);
    constexpr int A_PER_T = kBM * kBK / kThreads;
    constexpr int B_PER_T = kBK * W / kThreads;
    static_assert(B_PER_T * kThreads == kBK * W, "B tile split");

    int colt, mt;
    masked_task(ncolt, colt, mt);
    const int cb = colt / nsub, sub = colt % nsub;
    const int cnt = counts[cb];
    if (cnt == 0) return;
    const int *__restrict__ list = lists + (long)cb * kmax;
    const int m0 = mt * kBM;
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

    // C += acc: a fragment row's C values are loaded together from clamped, in-bounds addresses and
    // then stored where in range, so that the workgroup waits for one load latency per fragment
    // row instead of one per element
#pragma unroll
    for (int mi = 0; mi < FM; ++mi) {
        T cv[NT][4];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = min(m0 + wave * 32 + mi * 16 + MfmaOf<T>::row(lk, r), m - 1);
                const int col = min(n0 + nt * 16 + lrow, n - 1);
                cv[nt][r] = c[(long)row * ldc + col];
            }
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wave * 32 + mi * 16 + MfmaOf<T>::row(lk, r);
                const int col = n0 + nt * 16 + lrow;
                if (row < m && col < n) c[(long)row * ldc + col] = cv[nt][r] + acc[mi][nt][r];
            }
    }
}

// ---- FP64 masked GEMM with B not transposed and n even: direct copies into LDS ----
//
// op(A) is first written transposed into a k x ldt workspace (ldt = m rounded up to the tile
// rows, the padding zero), so that the listed k of a slice are whole rows of both operands and
// both go from global memory straight into LDS (global_load_lds_dwordx4, 1 KiB per wave
// instruction).  The column block at the end of B is read as clamped 16-byte chunks, exact for
// even n (a chunk never straddles the last column); odd n keeps masked_gemm_kernel.

__device__ __attribute__((always_inline)) inline void copy16_to_lds(const void *src, void *dst) {
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void *)dst, 16, 0, 0);
#endif
}

// at[p][i] = op(A)[i][p] for i < m (0 for m <= i < ldt), p < k; 64 x 64 tiles
template <bool TA>
__global__ __launch_bounds__(256) void masked_prep_a(double *__restrict__ at, long ldt, const double *__restrict__ a,
                                                     long lda, int m, int k) {
    __shared__ double t[64][65];
    const int p0 = blockIdx.y * 64, i0 = blockIdx.x * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    double v[16];
    if (TA) {  // a[p][i]: already k-major, copied with the padding
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int p = min(p0 + ty + 4 * q, k - 1), i = min(i0 + tx, m - 1);
            v[q] = a[(long)p * lda + i];
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int p = p0 + ty + 4 * q, i = i0 + tx;
            if (p < k) at[(long)p * ldt + i] = i < m ? v[q] : 0.0;
        }
    } else {  // a[i][p]: read along p, write along i through the LDS tile
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int i = min(i0 + ty + 4 * q, m - 1), p = min(p0 + tx, k - 1);
            v[q] = a[(long)i * lda + p];
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) t[ty + 4 * q][tx] = i0 + ty + 4 * q < m ? v[q] : 0.0;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int p = p0 + ty + 4 * q, i = i0 + tx;
            if (p < k) at[(long)p * ldt + i] = t[tx][ty + 4 * q];
        }
    }
}

// ---- FP64 masked GEMM, 64 x FM rows per workgroup, two LDS stages filled by direct copies ----
//
// Wave w owns rows 16 FM w .. 16 FM w + 16 FM - 1 of a 64 FM-row tile and all W columns: FM x NT
// MFMAs per k step of 4, 4 FM NT per slice of 16 list entries between two barriers.  op(A) is read
// from the transposed workspace (ldt a multiple of 64 FM), one 1 KiB copy per 128 values of a
// listed k row; B's 16 listed rows are packed (W doubles per row, lane i of copy u writing chunk
// 64 u + i).  Odd k rows swap the 128-byte halves of each 256 bytes (lane i loads chunk i ^ 8) for
// A, and for B when a row is a multiple of 256 bytes, so that the two k rows of one ds_read_b64
// phase fall on disjoint banks.  The list entries a wave copies next are loaded one slice ahead
// straight from the compacted list (no LDS copy of the list); one s_waitcnt vmcnt(0) + barrier per
// slice makes the copies of every wave visible and frees the other stage.  The last, partial
// slice is copied with its entries clamped to the list's last one, and its rows past the list
// are zeroed in LDS before its MFMAs.
template <int NT, int FM>
__global__ __launch_bounds__(kThreads, 2) void masked_pipe_kernel(double *__restrict__ c,
                                                                  const double *__restrict__ at, long ldt,
                                                                  const double *__restrict__ b, int m, int n,
                                                                  long ldb, long ldc, int kmax,
                                                                  const int *__restrict__ lists,
                                                                  const int *__restrict__ counts, int blksize,
                                                                  int nsub, int ncolt) {
    constexpr int W = NT * 16, BMR = 64 * FM;
    constexpr int A_ROW = BMR * 8, A_BYTES = kBK * A_ROW;  // one k row: FM / 2 copies
    constexpr int ACP = A_ROW / 1024;
    constexpr int B_ROW = W * 8, B_BYTES = kBK * B_ROW;    // packed: W / 8 copies
    constexpr int BC = B_BYTES / 1024, B_PER_WAVE = (BC + 3) / 4;
    constexpr int CPR = W / 2;                              // 16-byte chunks per B row
    constexpr int BSW = B_ROW % 256 == 0 ? 128 : 0;
    static_assert(ACP * 1024 == A_ROW && BC * 1024 == B_BYTES, "copy shape");
    __shared__ __attribute__((aligned(16))) unsigned char A0[A_BYTES];
    __shared__ __attribute__((aligned(16))) unsigned char A1[A_BYTES];
    __shared__ __attribute__((aligned(16))) unsigned char B0[B_BYTES];
    __shared__ __attribute__((aligned(16))) unsigned char B1[B_BYTES];

    int colt, mt;
    masked_task(ncolt, colt, mt);
    const int cb = colt / nsub, sub = colt % nsub;
    const int cnt = counts[cb];
    if (cnt == 0) return;
    const int *__restrict__ list = lists + (long)cb * kmax;
    const int m0 = mt * BMR;
    const int n0 = cb * blksize + sub * W;
    if (n0 >= n) return;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lrow = lane & 15, lk = lane >> 4;
    const int nslices = (cnt + kBK - 1) / kBK, nrem = cnt % kBK;

    f64x4 acc[FM][NT];
#pragma unroll
    for (int mi = 0; mi < FM; ++mi)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mi][nt] = f64x4{0.0, 0.0, 0.0, 0.0};

    // this lane's B chunk of each of the wave's copies: k row and (clamped, even n) column
    int brow[B_PER_WAVE], bcol[B_PER_WAVE];
#pragma unroll
    for (int e = 0; e < B_PER_WAVE; ++e) {
        const int u = min(wave + 4 * e, BC - 1), P = u * 64 + lane;
        brow[e] = P / CPR;
        bcol[e] = min(n0 + 2 * ((P % CPR) ^ (BSW && brow[e] % 2 ? 8 : 0)), n - 2);
    }
    int ia[4], ib[B_PER_WAVE];  // list entries of the slice copied next
    auto load_idx = [&](int t) {
        const int k0 = t * kBK;
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) ia[jj] = list[min(k0 + 4 * wave + jj, cnt - 1)];
#pragma unroll
        for (int e = 0; e < B_PER_WAVE; ++e) ib[e] = list[min(k0 + brow[e], cnt - 1)];
    };
    auto issue = [&](unsigned char *as, unsigned char *bs) {
#pragma unroll
        for (int jj = 0; jj < 4; ++jj)
#pragma unroll
            for (int h = 0; h < ACP; ++h)
                copy16_to_lds(at + (long)ia[jj] * ldt + m0 + h * 128 + 2 * (lane ^ (jj % 2 ? 8 : 0)),
                              as + (4 * wave + jj) * A_ROW + h * 1024);
#pragma unroll
        for (int e = 0; e < B_PER_WAVE; ++e)
            if (wave + 4 * e < BC) copy16_to_lds(b + (long)ib[e] * ldb + bcol[e], bs + (wave + 4 * e) * 1024);
    };
    auto compute = [&](const unsigned char *as, const unsigned char *bs) {
#pragma unroll
        for (int kk = 0; kk < kBK / 4; ++kk) {
            const int kq = kk * 4 + lk, odd = kq % 2;
            double fa[FM], fb[NT];
#pragma unroll
            for (int mi = 0; mi < FM; ++mi)
                fa[mi] = *reinterpret_cast<const double *>(
                    as + kq * A_ROW + (((wave * 16 * FM + mi * 16 + lrow) * 8) ^ (odd << 7)));
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                fb[nt] = *reinterpret_cast<const double *>(bs + kq * B_ROW +
                                                           (((nt * 16 + lrow) * 8) ^ (odd ? BSW : 0)));
#pragma unroll
            for (int mi = 0; mi < FM; ++mi)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                    acc[mi][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[mi], fb[nt], acc[mi][nt], 0, 0, 0);
        }
    };
    // the copies (and list loads) of this wave landed, and every wave is past the same point
    auto sync = [&]() {
        __builtin_amdgcn_s_waitcnt((0x7 << 4) | (0xF << 8));
        asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    };
    // rows nrem..15 of the partial last slice (copies of the list's last entry) set to zero
    auto zero_tail = [&](unsigned char *as, unsigned char *bs) {
        for (int o = nrem * A_ROW + tid * 16; o < A_BYTES; o += kThreads * 16)
            *reinterpret_cast<f64x2 *>(as + o) = f64x2{0.0, 0.0};
        for (int o = nrem * B_ROW + tid * 16; o < B_BYTES; o += kThreads * 16)
            *reinterpret_cast<f64x2 *>(bs + o) = f64x2{0.0, 0.0};
        __syncthreads();
    };

    load_idx(0);
    issue(A0, B0);
    if (nslices > 1) load_idx(1);
    for (int t = 0; t < nslices; t += 2) {
        sync();
        if (t + 1 < nslices) {
            issue(A1, B1);
            if (t + 2 < nslices) load_idx(t + 2);
        }
        if (t == nslices - 1 && nrem) zero_tail(A0, B0);
        compute(A0, B0);
        if (t + 1 >= nslices) break;
        sync();
        if (t + 2 < nslices) {
            issue(A0, B0);
            if (t + 3 < nslices) load_idx(t + 3);
        }
        if (t + 1 == nslices - 1 && nrem) zero_tail(A1, B1);
        compute(A1, B1);
    }

    // C += acc, a fragment row's C values loaded together (as in masked_gemm_kernel)
#pragma unroll
    for (int mi = 0; mi < FM; ++mi) {
        double cv[NT][4];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = min(m0 + wave * 16 * FM + mi * 16 + lk + 4 * r, m - 1);
                const int col = min(n0 + nt * 16 + lrow, n - 1);
                cv[nt][r] = c[(long)row * ldc + col];
            }
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wave * 16 * FM + mi * 16 + lk + 4 * r;
                const int col = n0 + nt * 16 + lrow;
                if (row < m && col < n) c[(long)row * ldc + col] = cv[nt][r] + acc[mi][nt][r];
            }
    }
}

template <typename T, int NT>
int launch_nt(bool ta, bool tb, dim3 grid, hipStream_t stream, T *c, const T *a, const T *b, int m, int n, long lda,
              long ldb, long ldc, int kmax, const int *lists, const int *counts, int blksize, int nsub, int ncolt) {
    hipEvent_t t0 = trace_start(stream);
#define CFGMM_MASKED(TA_, TB_)                                                                                      \
    hipLaunchKernelGGL((masked_gemm_kernel<T, NT, TA_, TB_>), grid, dim3(kThreads), 0, stream, c, a, b, m, n, lda, \
                       ldb, ldc, kmax, lists, counts, blksize, nsub, ncolt)
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
    if (nblk * nsub * div_up(m, (size_t)64) > (size_t)INT32_MAX)
        return fail(CFGMM_EINVAL, "masked dgemm: grid too large");
    const int ncolt = (int)(nblk * nsub);

    // FP64 with B not transposed and n even: op(A) transposed into the workspace, direct copies
    const int mode = tune_get(CFGMM_TUNE_MASKED_LDS);
    const bool lds = std::is_same<T, double>::value && !transb && n % 2 == 0 && mode != 1;
    const int fm = mode == 2 && nt < 4 ? 4 : 2;  // 64 fm rows per workgroup (fm 4 at NT 4 spills)
    const size_t bmr = 64 * (size_t)fm;
    const size_t ldt = div_up(m, bmr) * bmr;
    void *ws = nullptr;
    const size_t lists_bytes = sizeof(int) * nblk * k;
    const size_t at_off = div_up(lists_bytes + sizeof(int) * nblk, (size_t)256) * 256;
    int st = workspace_get(stream, lds ? at_off + sizeof(double) * k * ldt : lists_bytes + sizeof(int) * nblk, &ws);
    if (st != CFGMM_OK) return st;
    int *lists = static_cast<int *>(ws);
    int *counts = lists + nblk * k;
    st = launch_build_lists(stream, tab, ldtab, k, nblk, lists, counts);
    if (st != CFGMM_OK) return st;

    dim3 grid((unsigned)(nblk * nsub * mtiles), 1, 1);
    const int im = (int)m, in = (int)n, ik = (int)k, ib = (int)blksize;
    if constexpr (std::is_same<T, double>::value) {
        if (lds) {
            if (div_up(k, (size_t)64) > 65535u) return fail(CFGMM_EINVAL, "masked dgemm: k too large");
            double *at = reinterpret_cast<double *>(static_cast<char *>(ws) + at_off);
            hipEvent_t t0 = trace_start(stream);
            const dim3 pg((unsigned)(ldt / 64), (unsigned)div_up(k, (size_t)64));
            if (transa)
                hipLaunchKernelGGL((masked_prep_a<true>), pg, dim3(256), 0, stream, at, (long)ldt, a, (long)lda, im, ik);
            else
                hipLaunchKernelGGL((masked_prep_a<false>), pg, dim3(256), 0, stream, at, (long)ldt, a, (long)lda, im,
                                   ik);
            trace_stop(stream, "masked_prep_a", t0);
            st = hip_status(hipGetLastError(), "masked_prep_a launch");
            if (st != CFGMM_OK) return st;
            t0 = trace_start(stream);
            const dim3 pgrid((unsigned)(nblk * nsub * (ldt / bmr)), 1, 1);
#define CFGMM_MASKED_PIPE(NT_, FM_)                                                                              \
    hipLaunchKernelGGL((masked_pipe_kernel<NT_, FM_>), pgrid, dim3(kThreads), 0, stream, c, at, (long)ldt, b, im, \
                       in, (long)ldb, (long)ldc, ik, lists, counts, ib, nsub, ncolt)
            if (fm == 4) {
                switch (nt) {
                    case 1: CFGMM_MASKED_PIPE(1, 4); break;
                    case 2: CFGMM_MASKED_PIPE(2, 4); break;
                    default: CFGMM_MASKED_PIPE(3, 4); break;
                }
            } else {
                switch (nt) {
                    case 1: CFGMM_MASKED_PIPE(1, 2); break;
                    case 2: CFGMM_MASKED_PIPE(2, 2); break;
                    case 3: CFGMM_MASKED_PIPE(3, 2); break;
                    default: CFGMM_MASKED_PIPE(4, 2); break;
                }
            }
#undef CFGMM_MASKED_PIPE
            trace_stop(stream, "masked_gemm_kernel", t0);
            return hip_status(hipGetLastError(), "masked_gemm_kernel launch");
        }
    }
    switch (nt) {
        case 1: return launch_nt<T, 1>(transa, transb, grid, stream, c, a, b, im, in, (long)lda, (long)ldb, (long)ldc, ik, lists, counts, ib, nsub, ncolt);
        case 2: return launch_nt<T, 2>(transa, transb, grid, stream, c, a, b, im, in, (long)lda, (long)ldb, (long)ldc, ik, lists, counts, ib, nsub, ncolt);
        case 3: return launch_nt<T, 3>(transa, transb, grid, stream, c, a, b, im, in, (long)lda, (long)ldb, (long)ldc, ik, lists, counts, ib, nsub, ncolt);
        default: return launch_nt<T, 4>(transa, transb, grid, stream, c, a, b, im, in, (long)lda, (long)ldb, (long)ldc, ik, lists, counts, ib, nsub, ncolt);
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
