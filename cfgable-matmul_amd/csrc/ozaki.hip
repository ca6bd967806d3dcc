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
constexpr int kOzKMax = 16384; // K per launch: 6 * 127^2 * 16384 < 2^31 keeps every level exact
constexpr int kTile = 64;      // split-kernel tile (rows x cols)
constexpr int kExpNonFinite = 0x7fffffff;

// ---------------------------------------------------------------------------
// splitting: X is a logical rows x cols matrix, x(r, c) = x[r*rs + c*cs]; the two operands of
// a call (the rows of op(A) and the columns of op(B)) are sliced by the same launches,
// blockIdx.z selecting the operand.
// ---------------------------------------------------------------------------
struct SliceOperand {
    const double *x;
    long rs, cs;
    int rows;
    int ntiles;                 // 128-row tiles of the digit planes
    unsigned long long *pmax;   // [rows][nkc] partial row maxima
    int *ex;                    // [rows] row exponents
    int8_t *q;                  // digit planes
};
constexpr int kMaxChunk = 512;  // k columns per partial row maximum

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

__device__ inline unsigned long long absbits(double v) {
    return (unsigned long long)__double_as_longlong(fabs(v));
}

// pmax[r][kc] = max over columns [kc*512, kc*512 + 512) of |x(r, c)|, as the bit pattern of a
// non-negative double (whose integer order is its numeric order; NaN patterns sort above Inf).
// Workgroup (kc, 64-row block, operand); each thread keeps the maxima of the rows it loads and
// the workgroup combines them -- no atomics, no zeroing.
__global__ __launch_bounds__(kThreads) void row_absmax_kernel(SliceOperand opa, SliceOperand opb, int cols, int nkc) {
    const SliceOperand &o = blockIdx.z == 0 ? opa : opb;
    const int r0 = blockIdx.y * kTile, c0 = blockIdx.x * kMaxChunk;
    if (r0 >= o.rows) return;
    __shared__ unsigned long long red[kThreads / 64][kTile];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int c1 = min(cols, c0 + kMaxChunk);
    if (o.cs == 1) {
        // k contiguous: a wave reads 64 consecutive columns of rows wave, wave+4, ...
        unsigned long long m[kTile / 4];
#pragma unroll
        for (int i = 0; i < kTile / 4; ++i) m[i] = 0ull;
        for (int c = c0 + lane; c < c1; c += 64)
#pragma unroll
            for (int i = 0; i < kTile / 4; ++i) {
                const int r = r0 + wave + 4 * i;
                if (r < o.rows) {
                    const unsigned long long b = absbits(o.x[(long)r * o.rs + c]);
                    m[i] = b > m[i] ? b : m[i];
                }
            }
#pragma unroll
        for (int i = 0; i < kTile / 4; ++i) {
            unsigned long long v = m[i];
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) {
                const unsigned long long u = __shfl_xor(v, d);
                v = u > v ? u : v;
            }
            if (lane == 0) red[0][wave + 4 * i] = v;
        }
        __syncthreads();
        if (tid < kTile && r0 + tid < o.rows) o.pmax[(long)(r0 + tid) * nkc + blockIdx.x] = red[0][tid];
    } else {
        // rows contiguous: lane = row, the waves split the columns
        unsigned long long m = 0ull;
        const int r = r0 + lane;
        if (r < o.rows)
            for (int c = c0 + wave; c < c1; c += kThreads / 64) {
                const unsigned long long b = absbits(o.x[(long)r * o.rs + (long)c * o.cs]);
                m = b > m ? b : m;
            }
        red[wave][lane] = m;
        __syncthreads();
        if (tid < kTile && r0 + tid < o.rows) {
            unsigned long long v = red[0][tid];
#pragma unroll
            for (int w = 1; w < kThreads / 64; ++w) v = red[w][tid] > v ? red[w][tid] : v;
            o.pmax[(long)(r0 + tid) * nkc + blockIdx.x] = v;
        }
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

// ex[r] from the partial maxima of row r, both operands (rows of op(A) first).
__global__ void row_exponent_kernel(SliceOperand opa, SliceOperand opb, int nkc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool b = i >= opa.rows;
    const SliceOperand &o = b ? opb : opa;
    const int r = b ? i - opa.rows : i;
    if (r >= o.rows) return;
    unsigned long long v = 0ull;
    for (int kc = 0; kc < nkc; ++kc) {
        const unsigned long long u = o.pmax[(long)r * nkc + kc];
        v = u > v ? u : v;
    }
    o.ex[r] = row_exponent(v);
}

// Weight of the digit products of level l = s + t: 1 / (127^2 254^l).
__device__ __host__ constexpr double level_weight(int l) {
    double w = 1.0 / (127.0 * 127.0);
    for (int i = 0; i < l; ++i) w /= 254.0;
    return w;
}

// Slice planes in the GEMM's LDS order.  An operand X (rows_pad x k_pad: the rows of op(A), or
// the columns of op(B)) is stored as 4 KiB blocks, one per (plane s, 128-row tile t, 32-byte k
// step q), at byte ((s * ntiles + t) * nks + q) * 4096; inside a block the 16-byte chunk holding
// row rr, k bytes [16 h, 16 h + 16) sits at chunk index 2 rr + (h ^ ((rr >> 3) & 1)).  A stage of
// the GEMM is then one contiguous copy per block into LDS, and the swizzle spreads the fragment
// reads of 16 consecutive rows over all LDS banks.
constexpr int kGT = 128;               // GEMM tile rows / columns
constexpr int kGK = 32;                // k bytes per step (v_mfma_i32_32x32x32_i8)
constexpr int kBlock = kGT * kGK;      // bytes of one (plane, tile, k step) block
constexpr int kGThreads = 512;         // 8 waves: 2 x 4 wave tiles of 64 x 32

// 16 bytes per lane from global memory straight into LDS at dst + 16 * lane (dst uniform across
// the wave; global_load_lds_dwordx4).  Device-only: a kernel template that names the builtin in
// its host pass loses its launch stub.
__device__ __attribute__((always_inline)) inline void copy16_to_lds(const void *src, void *dst) {
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void *)dst, 16, 0, 0);
#endif
}

__host__ __device__ inline long block_offset(int s, int t, int q, int ntiles, int nks) {
    return (((long)s * ntiles + t) * nks + q) * kBlock;
}
__device__ inline int chunk_in_block(int rr, int h) { return 2 * rr + (h ^ ((rr >> 3) & 1)); }

// Digit planes of X: the s-th balanced digit of x(r, c) * 2^-ex[r] in the block layout above;
// zero outside X.  One workgroup per 64 x 64 tile of X; each thread cuts 16 consecutive k.
template <int S>
__global__ __launch_bounds__(kThreads) void slice_kernel(SliceOperand opa, SliceOperand opb, int cols, int nks) {
    const SliceOperand &o = blockIdx.z == 0 ? opa : opb;
    const int rows = o.rows, ntiles = o.ntiles;
    const int *__restrict__ ex = o.ex;
    int8_t *__restrict__ q = o.q;
    __shared__ double t[kTile][kTile + 1];
    const int r0 = blockIdx.y * kTile, c0 = blockIdx.x * kTile;
    if (r0 >= ntiles * kGT) return;  // past this operand's padded rows (uniform)
    load_tile(t, o.x, o.rs, o.cs, rows, cols, r0, c0);
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
    const int kc = (c0 + chunk * 16) >> 4;  // 16-byte k chunk
    const long off = block_offset(0, gr / kGT, kc >> 1, ntiles, nks) + 16L * chunk_in_block(gr % kGT, kc & 1);
    const long plane = (long)ntiles * nks * kBlock;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        i32x4 v = {packed[s][0], packed[s][1], packed[s][2], packed[s][3]};
        *reinterpret_cast<i32x4 *>(q + s * plane + off) = v;
    }
}

// ---------------------------------------------------------------------------
// int8 slice GEMM with per-level int32 accumulators and f64 recombination
// ---------------------------------------------------------------------------
typedef int i32x16 __attribute__((ext_vector_type(16)));

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

// One 128 x 128 tile of C per workgroup of 8 waves; wave (wm, wn) owns rows wm*64.., columns
// wn*32.. as two 32 x 32 v_mfma_i32_32x32x32_i8 tiles per level.  A stage (32 k bytes of every
// digit plane of the A and B tiles, S * 8 KiB) goes from global memory straight into LDS
// (global_load_lds_dwordx4: no registers held), three stages in flight: stage t+2 is issued
// while stage t is computed.  Wave w copies, per plane, the 1 KiB part (w & 3) of the A (w < 4)
// or B block.  Lane l of an MFMA operand holds row / column l % 32, k bytes 16 (l / 32) ..;
// output element r of lane l is row 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32.
template <int S>
__global__ __launch_bounds__(kGThreads, 1) void ozaki_gemm_kernel(
    double *__restrict__ c, long ldc, int m, int n, const int8_t *__restrict__ qa, const int8_t *__restrict__ qb,
    int ntiles_a, int ntiles_b, int nks, int q0, int q1, const int *__restrict__ ea, const int *__restrict__ eb,
    int group_m) {
    // one __shared__ object per stage, so that the compiler's wait for a stage's copies does not
    // also wait for the copies into the other two
    __shared__ __attribute__((aligned(16))) int8_t St0[2][S][kBlock];
    __shared__ __attribute__((aligned(16))) int8_t St1[2][S][kBlock];
    __shared__ __attribute__((aligned(16))) int8_t St2[2][S][kBlock];

    const int tiles_m = (m + kGT - 1) / kGT, tiles_n = (n + kGT - 1) / kGT;
    int tm, tn;
    raster_of(xcd_remap(blockIdx.x, gridDim.x), tiles_m, tiles_n, group_m, tm, tn);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 2, wn = wave & 3;

    // this wave's copies: operand (wave >> 2), part (wave & 3) of each plane's block
    const int op = wave >> 2, part = wave & 3;
    const int8_t *__restrict__ src =
        (op == 0 ? qa + block_offset(0, tm, 0, ntiles_a, nks) : qb + block_offset(0, tn, 0, ntiles_b, nks)) +
        part * 1024 + lane * 16;
    const long plane = (long)(op == 0 ? ntiles_a : ntiles_b) * nks * kBlock;
    auto issue = [&](int q, int8_t (*st)[S][kBlock]) {
#pragma unroll
        for (int s = 0; s < S; ++s)
            copy16_to_lds(src + s * plane + (long)q * kBlock, &st[op][s][part * 1024]);
    };

    i32x16 acc[S][2];
#pragma unroll
    for (int l = 0; l < S; ++l)
#pragma unroll
        for (int rt = 0; rt < 2; ++rt)
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[l][rt][e] = 0;

    // fragment offsets (bytes) inside a block: rows wm*64 + rt*32 + lane%32 of A, columns
    // wn*32 + lane%32 of B, chunk lane/32
    const int h = lane >> 5;
    int a_off[2];
#pragma unroll
    for (int rt = 0; rt < 2; ++rt) {
        const int rr = wm * 64 + rt * 32 + (lane & 31);
        a_off[rt] = 16 * chunk_in_block(rr, h);
    }
    const int b_off = 16 * chunk_in_block(wn * 32 + (lane & 31), h);
    auto compute = [&](const int8_t (*st)[S][kBlock]) {
        i32x4 fb[S];
#pragma unroll
        for (int sb = 0; sb < S; ++sb) fb[sb] = *reinterpret_cast<const i32x4 *>(&st[1][sb][b_off]);
#pragma unroll
        for (int sa = 0; sa < S; ++sa) {
            i32x4 fa[2];
#pragma unroll
            for (int rt = 0; rt < 2; ++rt) fa[rt] = *reinterpret_cast<const i32x4 *>(&st[0][sa][a_off[rt]]);
#pragma unroll
            for (int sb = 0; sb + sa < S; ++sb)
#pragma unroll
                for (int rt = 0; rt < 2; ++rt)
                    acc[sa + sb][rt] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[rt], fb[sb], acc[sa + sb][rt], 0, 0, 0);
        }
    };
    // Before stage t is read: this wave's copies of stage t have landed (S newer copies, those of
    // stage t+1, may still be in flight), then the barrier makes that true of every wave -- and
    // tells each wave that all have finished reading stage t-1, whose buffer the copies of stage
    // t+2 refill.  s_waitcnt vmcnt(S): vmcnt in bits 3:0, expcnt and lgkmcnt left at their maxima.
    auto sync = [&]() {
        __builtin_amdgcn_s_waitcnt(S | (0x7 << 4) | (0xF << 8));
        asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    };
    // Past the last step the copies re-read the last step into an idle buffer (never computed), so
    // that the number of copies in flight, and the wait above, is the same every step.
    const int nsteps = q1 - q0, last = q1 - 1;
    issue(q0, St0);
    issue(min(q0 + 1, last), St1);
    int t = 0;
    for (; t + 3 <= nsteps; t += 3) {
        sync();
        issue(min(q0 + t + 2, last), St2);
        compute(St0);
        sync();
        issue(min(q0 + t + 3, last), St0);
        compute(St1);
        sync();
        issue(min(q0 + t + 4, last), St1);
        compute(St2);
    }
    if (t < nsteps) {
        sync();
        issue(last, St2);
        compute(St0);
    }
    if (t + 1 < nsteps) {
        sync();
        compute(St1);
    }
    __builtin_amdgcn_s_waitcnt(0);  // no copy into LDS outlives the workgroup

    // Recombination: C += 2^(ea+eb) * sum_l acc[l] / (127^2 254^l), smallest level first.  The
    // loads of a row tile's exponents and C values are issued together (clamped in-bounds
    // addresses), so that the workgroup waits for one load latency, not one per element.
    const int m0 = tm * kGT, n0 = tn * kGT;
    const int col = n0 + wn * 32 + (lane & 31);
    const bool col_ok = col < n;
    const int colc = col_ok ? col : n - 1;
    const int e2 = eb[colc];
#pragma unroll
    for (int rt = 0; rt < 2; ++rt) {
        int e1[16];
        double cv[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = m0 + wm * 64 + rt * 32 + 8 * (r >> 2) + 4 * h + (r & 3);
            const int rowc = row < m ? row : m - 1;
            e1[r] = ea[rowc];
            cv[r] = c[(long)rowc * ldc + colc];
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = m0 + wm * 64 + rt * 32 + 8 * (r >> 2) + 4 * h + (r & 3);
            double v = 0.0;
#pragma unroll
            for (int l = S - 1; l >= 0; --l) v += (double)acc[l][rt][r] * level_weight(l);
            v = (e1[r] == kExpNonFinite || e2 == kExpNonFinite) ? __builtin_nan("") : ldexp(v, e1[r] + e2);
            if (row < m && col_ok) c[(long)row * ldc + col] = cv[r] + v;
        }
    }
}

template <int S>
int run_ozaki(hipStream_t stream, double *c, const double *a, const double *b, int m, int n, int k, long lda, long ldb,
              long ldc, bool ta, bool tb) {
    const long m_pad = (long)div_up(m, kGT) * kGT;
    const long n_pad = (long)div_up(n, kGT) * kGT;
    const long k_pad = (long)div_up(k, kTile) * kTile;  // the slicing grid's k tile; a multiple of kGK
    const int ntiles_a = (int)(m_pad / kGT), ntiles_b = (int)(n_pad / kGT), nks = (int)(k_pad / kGK);
    // workspace: qa [S][m_pad x k_pad], qb [S][n_pad x k_pad] (block layout), partial row maxima
    // (m+n) x nkc u64, exponents (m+n) int
    const int nkc = (int)div_up(k, kMaxChunk);
    const size_t qa_bytes = (size_t)S * m_pad * k_pad, qb_bytes = (size_t)S * n_pad * k_pad;
    const size_t off_qb = div_up(qa_bytes, 256) * 256;
    const size_t off_max = off_qb + div_up(qb_bytes, 256) * 256;
    const size_t off_ex = off_max + div_up(sizeof(unsigned long long) * (m + n) * nkc, 256) * 256;
    const size_t total = off_ex + sizeof(int) * (m + n);
    void *ws = nullptr;
    int st = workspace_get(stream, total, &ws);
    if (st != CFGMM_OK) return st;
    char *base = static_cast<char *>(ws);
    unsigned long long *pmax = reinterpret_cast<unsigned long long *>(base + off_max);
    int *ex = reinterpret_cast<int *>(base + off_ex);
    int8_t *qa = reinterpret_cast<int8_t *>(base);
    int8_t *qb = reinterpret_cast<int8_t *>(base + off_qb);

    // op(A) rows: (i, p) -> a[i*lda + p] or a[p*lda + i]; op(B)^T rows: (j, p) -> b[p*ldb + j] or b[j*ldb + p]
    const SliceOperand oa{a, ta ? 1 : lda, ta ? lda : 1, m, ntiles_a, pmax, ex, qa};
    const SliceOperand ob{b, tb ? ldb : 1, tb ? 1 : ldb, n, ntiles_b, pmax + (size_t)m * nkc, ex + m, qb};
    const unsigned row_blocks = (unsigned)(std::max(m_pad, n_pad) / kTile);
    hipEvent_t t0 = trace_start(stream);
    hipLaunchKernelGGL(row_absmax_kernel, dim3((unsigned)nkc, row_blocks, 2), dim3(kThreads), 0, stream, oa, ob, k,
                       nkc);
    hipLaunchKernelGGL(row_exponent_kernel, dim3((unsigned)div_up(m + n, 256)), dim3(256), 0, stream, oa, ob, nkc);
    hipLaunchKernelGGL((slice_kernel<S>), dim3((unsigned)(k_pad / kTile), row_blocks, 2), dim3(kThreads), 0, stream,
                       oa, ob, k, nks);
    trace_stop(stream, "ozaki_split", t0);
    st = hip_status(hipGetLastError(), "ozaki split launch");
    if (st != CFGMM_OK) return st;

    const int tiles = ntiles_a * ntiles_b;
    const int ks_max = kOzKMax / kGK;  // k steps per launch that keep every level sum exact
    for (int q0 = 0; q0 < nks; q0 += ks_max) {
        const int q1 = std::min(nks, q0 + ks_max);
        t0 = trace_start(stream);
        hipLaunchKernelGGL((ozaki_gemm_kernel<S>), dim3((unsigned)tiles), dim3(kGThreads), 0, stream, c, ldc, m, n, qa,
                           qb, ntiles_a, ntiles_b, nks, q0, q1, ex, ex + m, 4);
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
