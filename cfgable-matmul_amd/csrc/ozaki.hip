// ozaki.hip -- split-precision ("Ozaki scheme") f64 GEMM on the int8 MFMA (gfx950).
//
// No counterpart in the reference (its GEMM is plain f64, src/impl_matmul.rs:155-231);
// this is the split-precision mode of the same contract C += op(A) op(B).
//
// Method (Ozaki et al.'s error-free splitting with integer slices, as in the
// int8 "Ozaki scheme" GEMM of Ootomo, Ozaki and Yokota):
//   * each row i of op(A) is scaled by 2^-ea[i] and each column j of op(B) by
//     2^-eb[j], with ea/eb the binary exponents of the row / column maxima, so
//     that every scaled entry lies in (-1, 1);
//   * a scaled entry x is cut into S balanced int8 digits of radix 254:
//       v_0 = 127 x,  q_s = rint(v_s),  v_{s+1} = 254 (v_s - q_s),
//       x = sum_{s<S} q_s / (127 * 254^s) + r,   |q_s| <= 127,  |r| <= 254^-S
//     (v_0 lies in (-127, 127) and every later v is 254 times a remainder in
//     [-1/2, 1/2], so each int8 digit carries about 8 bits);
//   * A B ~= diag(2^ea) [ sum_{s+t < S} Q_s P_t / (127^2 254^(s+t)) ] diag(2^eb):
//     every slice product Q_s P_t is computed exactly with
//     v_mfma_i32_16x16x64_i8 (int32 accumulation; products of one level
//     s+t share an accumulator, exact while S * 127^2 * K < 2^31), and the
//     levels are converted to f64, weighted and summed in f64 from the
//     smallest level up.
// Per product term the dropped levels (s+t >= S) and the digit residuals are
// at most (S + 1.01) 254^-S in scaled units; S = 5 gives about 1e-11 relative
// error on well-scaled data, S = 6 about 1e-13.
// Rows or columns containing Inf/NaN produce NaN in the rows/columns of C they touch.
#include "internal.hpp"

#include <algorithm>

namespace cfgmm {
namespace {

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef long long i64x2 __attribute__((ext_vector_type(2)));

constexpr int kThreads = 256;
constexpr int kXcds = 8;
constexpr int kOzBK = 64;      // K per int8 MFMA (16x16x64)
constexpr int kOzKMax = 16384; // K per launch: 6 * 127^2 * 16384 < 2^31 keeps every level exact
constexpr int kTile = 64;      // split-kernel tile (rows x cols)
constexpr int kExpNonFinite = 0x7fffffff;

// ---------------------------------------------------------------------------
// splitting: X is a logical rows x cols matrix, x(r, c) = x[r*rs + c*cs]
// ---------------------------------------------------------------------------
__global__ void zero_u64_kernel(unsigned long long *__restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// Loads the 64x64 tile (r0, c0) of X into LDS, coalesced along whichever of
// the two directions is contiguous in memory; entries outside X are zero.
__device__ inline void load_tile(double (*t)[kTile + 1], const double *__restrict__ x, long rs, long cs, int rows,
                                 int cols, int r0, int c0) {
    const bool col_fast = cs == 1;
    for (int e = threadIdx.x; e < kTile * kTile; e += kThreads) {
        const int a = e % kTile, b = e / kTile;
        const int r = col_fast ? b : a, c = col_fast ? a : b;
        const int gr = r0 + r, gc = c0 + c;
        double v = 0.0;
        if (gr < rows && gc < cols) v = x[(long)gr * rs + (long)gc * cs];
        t[r][c] = v;
    }
}

// rowmax[r] = max(rowmax[r], max_c |x(r, c)|), as the bit pattern of a
// non-negative double (whose integer order equals its numeric order; NaN
// patterns sort above Inf).
__global__ __launch_bounds__(kThreads) void row_absmax_kernel(const double *__restrict__ x, long rs, long cs, int rows,
                                                              int cols, unsigned long long *__restrict__ rowmax) {
    __shared__ double t[kTile][kTile + 1];
    const int r0 = blockIdx.y * kTile, c0 = blockIdx.x * kTile;
    load_tile(t, x, rs, cs, rows, cols, r0, c0);
    __syncthreads();
    if (threadIdx.x < kTile && r0 + (int)threadIdx.x < rows) {
        unsigned long long m = 0ull;
        for (int c = 0; c < kTile; ++c) {
            const unsigned long long bits = (unsigned long long)__double_as_longlong(fabs(t[threadIdx.x][c]));
            m = bits > m ? bits : m;
        }
        atomicMax(rowmax + r0 + threadIdx.x, m);
    }
}

// Binary exponent e with max|row| < 2^e (0 for an all-zero row, a marker for a non-finite row).
__device__ inline int row_exponent(unsigned long long maxbits) {
    const double m = __longlong_as_double((long long)maxbits);
    if (!(m <= 1.79769313486231570815e308)) return kExpNonFinite;  // Inf or NaN
    if (m == 0.0) return 0;
    int e;
    frexp(m, &e);  // m = f * 2^e, f in [0.5, 1)
    return e;
}

__global__ void row_exponent_kernel(const unsigned long long *__restrict__ rowmax, int *__restrict__ ex, int rows) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < rows) ex[i] = row_exponent(rowmax[i]);
}

// Weight of the digit products of level l = s + t: 1 / (127^2 254^l).
__device__ __host__ constexpr double level_weight(int l) {
    double w = 1.0 / (127.0 * 127.0);
    for (int i = 0; i < l; ++i) w /= 254.0;
    return w;
}

// q[s][r][c] (int8, rows_pad x cols_pad per slice, c contiguous) = s-th balanced
// digit of x(r, c) * 2^-ex[r]; zero outside X.
template <int S>
__global__ __launch_bounds__(kThreads) void slice_kernel(const double *__restrict__ x, long rs, long cs, int rows,
                                                         int cols, const int *__restrict__ ex, int8_t *__restrict__ q,
                                                         long rows_pad, long cols_pad) {
    __shared__ double t[kTile][kTile + 1];
    const int r0 = blockIdx.y * kTile, c0 = blockIdx.x * kTile;
    load_tile(t, x, rs, cs, rows, cols, r0, c0);
    __syncthreads();
    const int r = threadIdx.x / 4, chunk = threadIdx.x % 4;  // 16 consecutive columns per thread
    const int gr = r0 + r;
    const int e = gr < rows ? ex[gr] : 0;
    const bool finite = e != kExpNonFinite;
    int packed[S][4];
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
        for (int w = 0; w < 4; ++w) packed[s][w] = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        double v = finite ? ldexp(t[r][chunk * 16 + j], -e) * 127.0 : 0.0;  // |v| < 127
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const double qv = rint(v);  // |qv| <= 127
            v = (v - qv) * 254.0;       // v - qv is exact, |v - qv| <= 1/2
            packed[s][j / 4] |= ((int)qv & 0xff) << (8 * (j % 4));
        }
    }
#pragma unroll
    for (int s = 0; s < S; ++s) {
        i32x4 v = {packed[s][0], packed[s][1], packed[s][2], packed[s][3]};
        *reinterpret_cast<i32x4 *>(q + ((long)s * rows_pad + gr) * cols_pad + c0 + chunk * 16) = v;
    }
}

// ---------------------------------------------------------------------------
// int8 slice GEMM with per-level int32 accumulators and f64 recombination
// ---------------------------------------------------------------------------
constexpr int kOzBM = 128, kOzBN = 64;  // workgroup tile: 2 x 2 waves of 64 x 32
constexpr int kRowBytes = kOzBK;         // LDS row: 64 int8 of K, 16-byte chunks XOR-swizzled

__device__ inline void raster_of(int L, int tiles_m, int tiles_n, int group_m, int &tm, int &tn) {
    const int in_group = group_m * tiles_n;
    const int gid = L / in_group;
    const int first_m = gid * group_m;
    const int gsize = min(tiles_m - first_m, group_m);
    const int r = L % in_group;
    tm = first_m + r % gsize;
    tn = r / gsize;
}

__device__ inline int xcd_remap(int pid, int total) {
    const int full = (total / kXcds) * kXcds;
    return pid < full ? (pid % kXcds) * (total / kXcds) + pid / kXcds : pid;
}

// byte offset of 16-byte chunk `ch` of LDS row `row` (swizzle spreads the 16
// rows a ds_read_b128 phase touches over distinct banks)
__device__ inline int swz(int row, int ch) { return row * kRowBytes + ((ch ^ ((row >> 2) & 3)) << 4); }

template <int S>
__global__ __launch_bounds__(kThreads, 1) void ozaki_gemm_kernel(double *__restrict__ c, long ldc, int m, int n,
                                                                 const int8_t *__restrict__ qa,
                                                                 const int8_t *__restrict__ qb, long m_pad,
                                                                 long n_pad, long k_pad, int k0, int k1,
                                                                 const int *__restrict__ ea,
                                                                 const int *__restrict__ eb, int group_m) {
    constexpr int LEV = S;  // levels s + t = 0 .. S - 1
    __shared__ __attribute__((aligned(16))) int8_t As[2][S][kOzBM * kRowBytes];
    __shared__ __attribute__((aligned(16))) int8_t Bs[2][S][kOzBN * kRowBytes];

    const int tiles_m = (m + kOzBM - 1) / kOzBM, tiles_n = (n + kOzBN - 1) / kOzBN;
    int tm, tn;
    raster_of(xcd_remap(blockIdx.x, gridDim.x), tiles_m, tiles_n, group_m, tm, tn);
    const int m0 = tm * kOzBM, n0 = tn * kOzBN;

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;  // wave tile: rows wm*64.., cols wn*32..
    const int lrow = lane & 15, lk = lane >> 4;

    // global -> register staging: per slice, A tile = 512 chunks (2 per thread), B tile = 256 (1 per thread)
    i32x4 ra[S][2], rb[S];
    auto load_tiles = [&](int kk) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int e = tid + kThreads * h, row = e >> 2, ch = e & 3;
                ra[s][h] = *reinterpret_cast<const i32x4 *>(qa + ((long)s * m_pad + m0 + row) * k_pad + kk + ch * 16);
            }
            const int row = tid >> 2, ch = tid & 3;
            rb[s] = *reinterpret_cast<const i32x4 *>(qb + ((long)s * n_pad + n0 + row) * k_pad + kk + ch * 16);
        }
    };
    auto store_tiles = [&](int buf) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int e = tid + kThreads * h, row = e >> 2, ch = e & 3;
                *reinterpret_cast<i32x4 *>(&As[buf][s][swz(row, ch)]) = ra[s][h];
            }
            *reinterpret_cast<i32x4 *>(&Bs[buf][s][swz(tid >> 2, tid & 3)]) = rb[s];
        }
    };

    i32x4 acc[LEV][4][2];
#pragma unroll
    for (int l = 0; l < LEV; ++l)
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
            for (int ct = 0; ct < 2; ++ct) acc[l][rt][ct] = i32x4{0, 0, 0, 0};

    const int ntiles = (k1 - k0) / kOzBK;
    load_tiles(k0);
    store_tiles(0);
    __syncthreads();
    for (int t = 0; t < ntiles; ++t) {
        const int cur = t & 1;
        if (t + 1 < ntiles) load_tiles(k0 + (t + 1) * kOzBK);
        // B fragments of every slice; A fragments one slice at a time.
        i32x4 fb[S][2];
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int ct = 0; ct < 2; ++ct)
                fb[s][ct] = *reinterpret_cast<const i32x4 *>(&Bs[cur][s][swz(wn * 32 + ct * 16 + lrow, lk)]);
#pragma unroll
        for (int sa = 0; sa < S; ++sa) {
            i32x4 fa[4];
#pragma unroll
            for (int rt = 0; rt < 4; ++rt)
                fa[rt] = *reinterpret_cast<const i32x4 *>(&As[cur][sa][swz(wm * 64 + rt * 16 + lrow, lk)]);
#pragma unroll
            for (int sb = 0; sb + sa < S; ++sb)  // level sa + sb (0-based) <= S - 1
#pragma unroll
                for (int rt = 0; rt < 4; ++rt)
#pragma unroll
                    for (int ct = 0; ct < 2; ++ct)
                        acc[sa + sb][rt][ct] =
                            __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[rt], fb[sb][ct], acc[sa + sb][rt][ct], 0, 0, 0);
        }
        if (t + 1 < ntiles) store_tiles(cur ^ 1);
        __syncthreads();
    }

    // Recombination: C += 2^(ea+eb) * sum_l acc[l] / (127^2 254^l), smallest level first.
    // i32 16x16 output layout: lane l, element r holds D[4*(l/16) + r][l%16].
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm * 64 + rt * 16 + 4 * lk + r;
                const int col = n0 + wn * 32 + ct * 16 + lrow;
                if (row < m && col < n) {
                    double v = 0.0;
#pragma unroll
                    for (int l = LEV - 1; l >= 0; --l) v += (double)acc[l][rt][ct][r] * level_weight(l);
                    const int e1 = ea[row], e2 = eb[col];
                    v = (e1 == kExpNonFinite || e2 == kExpNonFinite) ? __builtin_nan("") : ldexp(v, e1 + e2);
                    c[(long)row * ldc + col] += v;
                }
            }
}

template <int S>
int run_ozaki(hipStream_t stream, double *c, const double *a, const double *b, int m, int n, int k, long lda, long ldb,
              long ldc, bool ta, bool tb) {
    const long m_pad = (long)div_up(m, kOzBM) * kOzBM;
    const long n_pad = (long)div_up(n, kOzBN) * kOzBN;
    const long k_pad = (long)div_up(k, kOzBK) * kOzBK;
    // workspace: qa [S][m_pad][k_pad], qb [S][n_pad][k_pad], rowmax (m+n) u64, exponents (m+n) int
    const size_t qa_bytes = (size_t)S * m_pad * k_pad, qb_bytes = (size_t)S * n_pad * k_pad;
    const size_t off_qb = div_up(qa_bytes, 256) * 256;
    const size_t off_max = off_qb + div_up(qb_bytes, 256) * 256;
    const size_t off_ex = off_max + div_up(sizeof(unsigned long long) * (m + n), 256) * 256;
    const size_t total = off_ex + sizeof(int) * (m + n);
    void *ws = nullptr;
    int st = workspace_get(stream, total, &ws);
    if (st != CFGMM_OK) return st;
    char *base = static_cast<char *>(ws);
    int8_t *qa = reinterpret_cast<int8_t *>(base);
    int8_t *qb = reinterpret_cast<int8_t *>(base + off_qb);
    unsigned long long *rmax = reinterpret_cast<unsigned long long *>(base + off_max);
    int *ex = reinterpret_cast<int *>(base + off_ex);

    // op(A) rows: (i, p) -> a[i*lda + p] or a[p*lda + i]; op(B)^T rows: (j, p) -> b[p*ldb + j] or b[j*ldb + p]
    const long a_rs = ta ? 1 : lda, a_cs = ta ? lda : 1;
    const long b_rs = tb ? ldb : 1, b_cs = tb ? 1 : ldb;
    hipEvent_t t0 = trace_start(stream);
    hipLaunchKernelGGL(zero_u64_kernel, dim3((unsigned)div_up(m + n, 256)), dim3(256), 0, stream, rmax, m + n);
    hipLaunchKernelGGL(row_absmax_kernel, dim3((unsigned)div_up(k, kTile), (unsigned)div_up(m, kTile)),
                       dim3(kThreads), 0, stream, a, a_rs, a_cs, m, k, rmax);
    hipLaunchKernelGGL(row_absmax_kernel, dim3((unsigned)div_up(k, kTile), (unsigned)div_up(n, kTile)),
                       dim3(kThreads), 0, stream, b, b_rs, b_cs, n, k, rmax + m);
    hipLaunchKernelGGL(row_exponent_kernel, dim3((unsigned)div_up(m + n, 256)), dim3(256), 0, stream, rmax, ex,
                       m + n);
    hipLaunchKernelGGL((slice_kernel<S>), dim3((unsigned)(k_pad / kTile), (unsigned)(m_pad / kTile)), dim3(kThreads),
                       0, stream, a, a_rs, a_cs, m, k, ex, qa, m_pad, k_pad);
    hipLaunchKernelGGL((slice_kernel<S>), dim3((unsigned)(k_pad / kTile), (unsigned)(n_pad / kTile)), dim3(kThreads),
                       0, stream, b, b_rs, b_cs, n, k, ex + m, qb, n_pad, k_pad);
    trace_stop(stream, "ozaki_split", t0);
    st = hip_status(hipGetLastError(), "ozaki split launch");
    if (st != CFGMM_OK) return st;

    const int tiles = (int)((m_pad / kOzBM) * (n_pad / kOzBN));
    for (long k0 = 0; k0 < k_pad; k0 += kOzKMax) {
        const long k1 = std::min<long>(k_pad, k0 + kOzKMax);
        t0 = trace_start(stream);
        hipLaunchKernelGGL((ozaki_gemm_kernel<S>), dim3((unsigned)tiles), dim3(kThreads), 0, stream, c, ldc, m, n, qa,
                           qb, m_pad, n_pad, k_pad, (int)k0, (int)k1, ex, ex + m, 4);
        trace_stop(stream, "ozaki_gemm_kernel", t0);
        st = hip_status(hipGetLastError(), "ozaki_gemm_kernel launch");
        if (st != CFGMM_OK) return st;
    }
    return CFGMM_OK;
}

}  // namespace

int launch_dgemm_ozaki(hipStream_t stream, double *c, const double *a, const double *b, size_t m, size_t n, size_t k,
                       size_t lda, size_t ldb, size_t ldc, bool transa, bool transb, int nslices) {
    if (m == 0 || n == 0 || k == 0) return CFGMM_OK;
    if (m > INT32_MAX / 2 || n > INT32_MAX / 2 || k > INT32_MAX / 2)
        return fail(CFGMM_EINVAL, "split-precision dgemm: dimension out of range");
    const int im = (int)m, in = (int)n, ik = (int)k;
    const long la = (long)lda, lb = (long)ldb, lc = (long)ldc;
    switch (nslices) {
        case 2: return run_ozaki<2>(stream, c, a, b, im, in, ik, la, lb, lc, transa, transb);
        case 3: return run_ozaki<3>(stream, c, a, b, im, in, ik, la, lb, lc, transa, transb);
        case 4: return run_ozaki<4>(stream, c, a, b, im, in, ik, la, lb, lc, transa, transb);
        case 5: return run_ozaki<5>(stream, c, a, b, im, in, ik, la, lb, lc, transa, transb);
        case 6: return run_ozaki<6>(stream, c, a, b, im, in, ik, la, lb, lc, transa, transb);
        default: return fail(CFGMM_EINVAL, "nslices must be in [2, 6]");
    }
}

}  // namespace cfgmm
