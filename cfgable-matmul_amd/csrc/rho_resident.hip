// rho_resident.hip -- masked AO-DM-AO rho over operands held resident on the device
// in per-grid-block compacted, int8-sliced form (split precision, Ozaki scheme).
//
// Same contract as launch_rho (src/impl_aodmao2rho.rs:171-244):
//   rho[s][g] += sum_{u,v in L_b} lhs[u][g] D_s[u][v] rhs[v][g],  b = g / blksize.
//
// The per-call kernels gather D_s[L_b, L_b] from the full matrix for every
// block; those gathers use a third of each cache line and dominate their run
// time.  When the same (D, phi, mask) are evaluated many times -- the serving
// case, where a molecule stays resident across requests -- the gather is done
// once, at creation, into one contiguous tile per block:
//
//   Dq[b][s][p](i, j)  p-th balanced int8 digit of D_s[L_b[i]][L_b[j]] * 2^-ed[b][s][i] (slice4),
//                      i, j < kpad_b = round_up(|L_b|, 64) (zero padded),
//                      ed = exponent of the row maximum over the block's own v;
//                      stored as 1 KiB tiles of 16 rows x 64 k, [i/16][j/64], each
//                      tile ordered [k/16][i%16][k%16] -- one MFMA A fragment per
//                      lane, 64 lanes reading 1 KiB contiguously;
//   Rq[b][p](c, i)     p-th slice of rhs[L_b[i]][b*blksize + c] * 2^-fg[g],
//                      fg = exponent of the column maximum over the block's AOs;
//                      stored [i/64][c][i%64] -- one 64-k step of all columns of
//                      the block is blksize*64 contiguous bytes -- with the four
//                      16-byte chunks of each 64-byte column row permuted (chunk
//                      j at j ^ ((c >> 2) & 3)), the order the evaluation reads
//                      them in from LDS, so that a stage is a plain copy;
//   Lc[b][i][c]        lhs[L_b[i]][b*blksize + c] in f64 (zero padded).
//
// An evaluation then streams Dq once: each workgroup owns 64 compacted rows of
// one block (four waves of 16) and 16..64 of its grid points, loads its A
// fragments straight from global memory (16 contiguous bytes per lane and
// slice), shares the B tile of the block through double-buffered LDS, and
// accumulates every slice product (sa + sb < S) exactly in int32 with
// v_mfma_i32_16x16x64_i8.  The epilogue recombines the levels in f64 with the
// digit weights and the power-of-two scales, multiplies by lhs and reduces the 64 rows; per-tile
// partial rows are summed in a fixed order by launch_rho_reduce.
//
// Symmetric form (every D_s == D_s^T, bitwise or to 2^-50; lhs and rhs one array; kpad <= 8192): only the
// tiles (row tile T of 64 rows, k step ks >= T) are stored, a plane being [T][4 row groups]
// [ks - T] 1 KiB tiles; a workgroup on row tile T streams its steps past the diagonal, doubles
// its int32 sums (exact: 2 * 6 * 8192 * 127^2 < 2^31), then adds the diagonal step -- the pairs
// below the diagonal are the mirror images of those it doubles.  Half the bytes and MFMAs.
//
// A row of D_s or a block column of rhs holding Inf/NaN makes the rho values
// it touches NaN.
#include "internal.hpp"

#include <algorithm>
#include <mutex>
#include <atomic>
#include <type_traits>
#include <vector>

namespace cfgmm {
namespace {

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef unsigned long long u64;

constexpr int kThreads = 256;
constexpr int kRows = 64;   // compacted rows per workgroup (4 waves x 16)
constexpr int kK = 64;      // k per v_mfma_i32_16x16x64_i8
constexpr int kExpNonFinite = 0x7fffffff;
constexpr int kMaxKpad = 16384;  // keeps every int32 level sum below 2^31: 6 * 16384 * 127^2 < 2^31
constexpr int kMaxKpadSym = 8192;  // the same with the off-diagonal sums doubled

struct BlockDesc {
    long d_off;  // bytes: Dq of set 0 ([nset][S] planes of 1 KiB tiles)
    long r_off;  // bytes: Rq ([S][kpad/64][blksize][64] int8)
    long l_off;  // bytes: Lc ([kpad][blksize] double)
    long e_off;  // bytes: ed ([nset][kpad] int)
    int n, kpad;
    int sym;  // upper-triangle tiles only
};

// 1 KiB tiles before row tile T in a plane of the symmetric form (nks k steps): 4 * sum_{t<T} (nks - t)
__host__ __device__ inline long sym_tiles_before(int T, int nks) {
    return 4L * ((long)T * nks - (long)T * (T - 1) / 2);
}
__host__ __device__ inline long dq_plane_bytes(int kpad, int sym) {
    const long nks = kpad / kK;
    return sym ? 1024L * sym_tiles_before((int)nks, (int)nks) : (long)kpad * kpad;
}

__device__ inline int exponent_of(u64 maxbits) {
    const double m = __longlong_as_double((long long)maxbits);
    if (!(m <= 1.79769313486231570815e308)) return kExpNonFinite;
    if (m == 0.0) return 0;
    int e;
    frexp(m, &e);
    return e;
}

// Balanced digits: x * 2^-e = sum_p d_p / (127 * 254^p) + r, d_p = rint of the running value,
// |d_p| <= 127, |r| <= 1 / (2 * 127 * 254^(S-1)).  |x * 2^-e| < 1 makes the first value
// 127 x 2^-e lie in (-127, 127); every later one is 254 times a remainder in [-1/2, 1/2].  About
// 8 bits per int8 slice (7 with truncated digits of 128): five slices hold what six did.  The
// product of digits of levels a and b weighs 1 / (127^2 254^(a+b)): a function of a + b only.
__device__ __host__ constexpr double level_weight(int l) {
    double w = 1.0 / (127.0 * 127.0);
    for (int i = 0; i < l; ++i) w /= 254.0;
    return w;
}

// byte `q` of the result is slice `p` of x * 2^-e, for four values
template <int S>
__device__ inline void slice4(const double x[4], const int e[4], unsigned w[S]) {
#pragma unroll
    for (int p = 0; p < S; ++p) w[p] = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double v = e[q] == kExpNonFinite ? 0.0 : ldexp(x[q], -e[q]) * 127.0;
#pragma unroll
        for (int p = 0; p < S; ++p) {
            const double t = rint(v);
            v = (v - t) * 254.0;
            w[p] |= ((unsigned)(int)t & 0xffu) << (8 * q);
        }
    }
}

// Byte offset of 16-byte chunk `ch` of the 64-byte row `row` (a column of a B tile) in LDS: the
// chunks are XOR-permuted against bank conflicts of the ds_read_b128 fragment reads.
__device__ inline int swz(int row, int ch) { return row * kK + ((ch ^ ((row >> 2) & 3)) << 4); }
// The same permutation inside one row: position of byte k (0..63) of column c's row.
__device__ inline int swz_in_row(int c, int k) { return (((k >> 4) ^ ((c >> 2) & 3)) << 4) | (k & 15); }

// 16 bytes per lane from global memory straight into LDS at dst + 16 * lane (dst uniform across
// the wave; global_load_lds_dwordx4).  Device-only: a kernel template that names the builtin in
// its host pass loses its launch stub.
// a value the same in every lane of the wave, moved to a scalar register (64 bits as two halves)
__device__ __attribute__((always_inline)) inline long uniform64(long x) {
    const unsigned lo = __builtin_amdgcn_readfirstlane((int)(unsigned)x);
    const unsigned hi = __builtin_amdgcn_readfirstlane((int)(x >> 32));
    return (long)(((unsigned long)hi << 32) | lo);
}

__device__ __attribute__((always_inline)) inline void copy16_to_lds(const void *src, void *dst) {
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void *)dst, 16, 0, 0);
#endif
}

// One workgroup per (16-row group g of the padded list, grid block b, density matrix s): the
// exponents ed of its 16 rows (each row's largest magnitude over all its columns), then every
// 1 KiB tile of the group (the k steps from the diagonal's on, in the symmetric form), each
// plane's tile written as 256 consecutive 4-byte words, one per thread (whole cache lines: with
// a row per wave the tiles' 16-byte row pieces were written by 16 different waves).  The loads of
// a batch (list entries, then the D values they index) are issued together; rows and columns past
// |L_b| are zero, their loads clamped.
template <int S>
__global__ __launch_bounds__(kThreads) void prep_dq_kernel(unsigned char *__restrict__ base,
                                                           const BlockDesc *__restrict__ blocks,
                                                           const double *__restrict__ dms, long stride_ao, int nao,
                                                           const int *__restrict__ lists) {
    const int g = blockIdx.x, b = blockIdx.y, s = blockIdx.z;
    const BlockDesc bd = blocks[b];
    const int n = bd.n, kpad = bd.kpad;
    if (g * 16 >= kpad || n == 0) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nks = kpad / kK;
    const long plane = dq_plane_bytes(kpad, bd.sym);
    const int T = g >> 2, ks0 = bd.sym ? T : 0;  // the group's row tile, its first stored k step
    const long row_tile = bd.sym ? sym_tiles_before(T, nks) + (long)(g & 3) * (nks - T) : (long)g * nks;
    unsigned char *__restrict__ out = base + bd.d_off + (long)s * S * plane + row_tile * 1024 + 4 * tid;
    int *__restrict__ ed = reinterpret_cast<int *>(base + bd.e_off) + (long)s * kpad;
    const int *__restrict__ list = lists + (long)b * nao;
    const double *__restrict__ dset = dms + (long)s * stride_ao * stride_ao;
    __shared__ int e_s[16];

    // exponents: wave w takes rows 4 w .. 4 w + 3 of the group, sixteen lanes per row
    {
        constexpr int kU = 8;
        const int r = 4 * wave + (lane >> 4), i = g * 16 + r, l16 = lane & 15;
        const double *__restrict__ drow = dset + (long)list[min(i, n - 1)] * stride_ao;
        u64 m = 0;
        for (int jb = 0; jb < n; jb += 16 * kU) {
            int idx[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) idx[u] = list[min(jb + 16 * u + l16, n - 1)];
            double v[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) v[u] = drow[idx[u]];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const u64 bits = (u64)__double_as_longlong(fabs(v[u]));
                m = jb + 16 * u + l16 < n && bits > m ? bits : m;
            }
        }
#pragma unroll
        for (int off = 8; off > 0; off >>= 1) {
            const u64 o = __shfl_xor(m, off);
            m = o > m ? o : m;
        }
        if (l16 == 0) {
            const int e = i < n ? exponent_of(m) : 0;
            e_s[r] = e;
            ed[i] = e;
        }
    }
    __syncthreads();

    // tiles: thread t holds row r = (t % 64) / 4 of the group and columns 16 (t / 64) + 4 (t % 4)
    // .. + 3 of each k step: bytes 4 t .. 4 t + 3 of the tile
    constexpr int kB = 4;  // tiles per batch
    const int r = lane >> 2, i = g * 16 + r, jc = 16 * wave + 4 * (lane & 3);
    const int e = e_s[r];
    const int ee[4] = {e, e, e, e};
    const double *__restrict__ drow = dset + (long)list[min(i, n - 1)] * stride_ao;
    for (int ks = ks0; ks < nks; ks += kB) {
        int idx[kB][4];
#pragma unroll
        for (int u = 0; u < kB; ++u)
#pragma unroll
            for (int q = 0; q < 4; ++q) idx[u][q] = list[min((ks + u) * kK + jc + q, n - 1)];
        double x[kB][4];
#pragma unroll
        for (int u = 0; u < kB; ++u)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double v = drow[idx[u][q]];
                x[u][q] = i < n && (ks + u) * kK + jc + q < n ? v : 0.0;
            }
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            if (ks + u < nks) {
                unsigned w[S];
                slice4<S>(x[u], ee, w);
#pragma unroll
                for (int p = 0; p < S; ++p)
                    *reinterpret_cast<unsigned *>(out + p * plane + (long)(ks + u - ks0) * 1024) = w[p];
            }
        }
    }
}

// One workgroup per grid block: fg of its points, Rq and Lc.
template <int S>
__global__ __launch_bounds__(kThreads) void prep_phi_kernel(unsigned char *__restrict__ base,
                                                            const BlockDesc *__restrict__ blocks,
                                                            const double *__restrict__ ao_lhs,
                                                            const double *__restrict__ ao_rhs, long stride_grid,
                                                            int ngrid, int nao, int blksize,
                                                            const int *__restrict__ lists, int *__restrict__ fg) {
    extern __shared__ u64 colmax[];  // [blksize]
    const int b = blockIdx.x;
    const BlockDesc bd = blocks[b];
    const int g0 = b * blksize;
    const int ncol = min(blksize, ngrid - g0);
    const int *__restrict__ list = lists + (long)b * nao;
    for (int c = threadIdx.x; c < blksize; c += kThreads) colmax[c] = 0;
    __syncthreads();
    // every pass below loads a batch of kU elements per thread together (list entries, then values;
    // clamped addresses, out-of-range elements masked): two load latencies per batch
    constexpr int kU = 8;
    const int n = bd.n;
    const int nel = n * ncol;
    for (int i0 = threadIdx.x; i0 < nel; i0 += kThreads * kU) {
        int li[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) li[u] = list[min(i0 + u * kThreads, nel - 1) / ncol];
        double v[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u)
            v[u] = ao_rhs[(long)li[u] * stride_grid + g0 + min(i0 + u * kThreads, nel - 1) % ncol];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int idx = i0 + u * kThreads;
            if (idx < nel) atomicMax(&colmax[idx % ncol], (u64)__double_as_longlong(fabs(v[u])));
        }
    }
    __syncthreads();
    for (int c = threadIdx.x; c < ncol; c += kThreads) fg[g0 + c] = exponent_of(colmax[c]);
    // Rq[p][i/64][c][i%64], four consecutive i per thread
    const int quads = bd.kpad / 4;
    const long plane = (long)blksize * bd.kpad;
    unsigned char *__restrict__ rq = base + bd.r_off;
    constexpr int kR = 2;  // iterations of four values each per batch
    const int nrq = blksize * quads;
    for (int i0 = threadIdx.x; i0 < nrq; i0 += kThreads * kR) {
        int li[kR][4];
#pragma unroll
        for (int u = 0; u < kR; ++u) {
            const int idx = min(i0 + u * kThreads, nrq - 1), i = 4 * (idx % quads);
#pragma unroll
            for (int q = 0; q < 4; ++q) li[u][q] = n > 0 ? list[min(i + q, n - 1)] : 0;  // row 0: masked
        }
        double x[kR][4];
#pragma unroll
        for (int u = 0; u < kR; ++u) {
            const int idx = min(i0 + u * kThreads, nrq - 1), c = idx / quads, i = 4 * (idx % quads);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double v = ao_rhs[(long)li[u][q] * stride_grid + g0 + min(c, ncol - 1)];
                x[u][q] = (i + q < n && c < ncol) ? v : 0.0;
            }
        }
#pragma unroll
        for (int u = 0; u < kR; ++u) {
            const int idx = i0 + u * kThreads;
            if (idx < nrq) {
                const int c = idx / quads, i = 4 * (idx % quads);
                const int e = c < ncol ? exponent_of(colmax[c]) : 0;
                const int ee[4] = {e, e, e, e};
                unsigned w[S];
                slice4<S>(x[u], ee, w);
#pragma unroll
                for (int p = 0; p < S; ++p)
                    *reinterpret_cast<unsigned *>(rq + p * plane + ((long)(i >> 6) * blksize + c) * kK +
                                                  swz_in_row(c, i & 63)) = w[p];
            }
        }
    }
    // Lc[i][c]
    double *__restrict__ lc = reinterpret_cast<double *>(base + bd.l_off);
    const int nlc = bd.kpad * blksize;
    for (int i0 = threadIdx.x; i0 < nlc; i0 += kThreads * kU) {
        int li[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u)
            li[u] = n > 0 ? list[min(min(i0 + u * kThreads, nlc - 1) / blksize, n - 1)] : 0;  // row 0: masked
        double v[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u)
            v[u] = ao_lhs[(long)li[u] * stride_grid + g0 + min(min(i0 + u * kThreads, nlc - 1) % blksize, ncol - 1)];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int idx = i0 + u * kThreads;
            if (idx < nlc) {
                const int i = idx / blksize, c = idx % blksize;
                lc[idx] = (i < n && c < ncol) ? v[u] : 0.0;
            }
        }
    }
}



// Task t of an evaluation: tasks[t] = {block, row tile, column subtile}.  Workgroup i takes task
// start_x + i/8 with x = i%8, so consecutive tasks (the tiles of one block, which share its Rq)
// run on one XCD and its L2.
__device__ inline int xcd_task(int i, int ntasks) {
    const int q = ntasks / 8, r = ntasks % 8, x = i % 8;
    return x * q + min(x, r) + i / 8;
}

// One task of density matrix s: block b, column subtile, and 64-row tile task.y -- or, for a
// task of the symmetric form with task.w = 1, the tiles task.y and ntiles-1-task.y one after the
// other (a constant number of k steps per task, one pipeline through both): partial[tile][s][g]
// for the subtile's grid points g < ngrid.
template <int S, int NT, int DEPTH>
__device__ __attribute__((always_inline)) inline void resident_tile(
    double *__restrict__ partial, const unsigned char *__restrict__ base, const BlockDesc *__restrict__ blocks,
    int4 task, int s, const int *__restrict__ fg, int blksize, int ngrid, int nset) {
    constexpr int W = NT * 16;
    // the two B stages are __shared__ objects of their own: the compiler's wait before reading one
    // then does not also wait for the copies into the other
    __shared__ __attribute__((aligned(16))) int8_t Bs0[S][W * kK];
    __shared__ __attribute__((aligned(16))) int8_t Bs1[S][W * kK];
    __shared__ __attribute__((aligned(16))) int8_t Bs2[DEPTH >= 2 ? S : 1][W * kK];
    __shared__ double red[kThreads / 64][W];

    const int b = task.x, sub = task.z;
    const BlockDesc bd = blocks[b];
    const int kpad = bd.kpad;
    const int c0 = sub * W;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lrow = lane & 15, lk = lane >> 4;

    const long dplane = dq_plane_bytes(kpad, bd.sym);
    const int nks = kpad / kK;
    const bool sym = bd.sym != 0;
    // the task's tiles and their k steps: all, or (symmetric form) those past the diagonal, then it
    const int T1 = task.y, T2 = sym && task.w ? kpad / kRows - 1 - task.y : -1;
    const int n1 = sym ? nks - T1 : nks, n2 = T2 > T1 ? nks - T2 : 0;
    const int nsteps = n1 + n2;
    // this wave's 16 rows of tile T: 1 KiB A tiles at (row tile, k step); lane l reads bytes 16l..16l+15
    auto arow_of = [&](int T) {
        const long row_tiles = sym ? sym_tiles_before(T, nks) + (long)wave * (nks - T)
                                   : (long)(T * (kRows / 16) + wave) * nks;
        return reinterpret_cast<const int8_t *>(base + bd.d_off) + (long)s * S * dplane + row_tiles * 1024 + lane * 16;
    };
    const int8_t *__restrict__ arow1 = arow_of(T1);
    const int8_t *__restrict__ arow2 = arow_of(T2 > T1 ? T2 : T1);
    // step t: tile, its local step, k step and A tile within the wave's row of tiles
    auto ks_of = [&](int t) {
        const int T = t < n1 ? T1 : T2, l = t < n1 ? t : t - n1, n = t < n1 ? n1 : n2;
        return sym ? (l < n - 1 ? T + 1 + l : T) : l;
    };
    auto a_of = [&](int t) {
        const int l = t < n1 ? t : t - n1, n = t < n1 ? n1 : n2;
        return sym ? (l < n - 1 ? l + 1 : 0) : l;
    };
    const long rplane = (long)blksize * kpad;
    const int8_t *__restrict__ rq = reinterpret_cast<const int8_t *>(base + bd.r_off) + (long)c0 * kK;

    i32x4 acc[S][NT];
    auto zero_acc = [&]() {
#pragma unroll
        for (int l = 0; l < S; ++l)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) acc[l][nt] = i32x4{0, 0, 0, 0};
    };
    zero_acc();

    i32x4 fa0[S], fa1[S], fa2[S];
    auto load_a = [&](int t, i32x4 *dst) {
        const int8_t *__restrict__ ar = (t < n1 ? arow1 : arow2) + (long)a_of(t) * 1024;
#pragma unroll
        for (int p = 0; p < S; ++p)  // each A tile is read once per evaluation: stream it past the caches
            dst[p] = __builtin_nontemporal_load(reinterpret_cast<const i32x4 *>(ar + p * dplane));
    };
    // B stage of step t: per plane, the W columns x 64 k of its k step are NT KiB contiguous in Rq,
    // already in LDS order; copied straight into LDS, 1 KiB per wave instruction.  Every wave issues
    // the same number of copies (the last part repeated where S*NT is not a multiple of 4), so that
    // the wait below counts the same on every wave.
    constexpr int kBParts = S * NT, kBPerWave = (kBParts + 3) / 4;
    auto issue_b = [&](int t, int8_t (*dst)[W * kK]) {
        const long ks = ks_of(t);
#pragma unroll
        for (int i = 0; i < kBPerWave; ++i) {
            const int j = min(wave + 4 * i, kBParts - 1);
            const int p = j / NT, part = j % NT;
            copy16_to_lds(rq + p * rplane + ks * blksize * kK + part * 1024 + lane * 16, &dst[p][part * 1024]);
        }
    };
    auto compute = [&](const int8_t (*bs)[W * kK], const i32x4 *fa) {
#pragma unroll
        for (int sb = 0; sb < S; ++sb) {
            i32x4 fb[NT];
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                fb[nt] = *reinterpret_cast<const i32x4 *>(&bs[sb][swz(nt * 16 + lrow, lk)]);
#pragma unroll
            for (int sa = 0; sa + sb < S; ++sa)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                    acc[sa + sb][nt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[sa], fb[nt], acc[sa + sb][nt], 0, 0, 0);
        }
    };

    // Epilogue of tile T.  i32 16x16 layout: lane l, element r holds row 4*(l/16)+r, column l%16.
    const int *__restrict__ ed = reinterpret_cast<const int *>(base + bd.e_off) + (long)s * kpad;
    const double *__restrict__ lc = reinterpret_cast<const double *>(base + bd.l_off);
    const int gblk = b * blksize;
    auto epilogue = [&](int T) {
        const double nan = __builtin_nan("");
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int c = c0 + nt * 16 + lrow;
            const int g = gblk + c;
            const bool g_ok = g < ngrid;
            const int eg = g_ok ? fg[g] : 0;
            double part = 0.0;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = T * kRows + wave * 16 + 4 * lk + r;
                if (i < bd.n && g_ok) {
                    const int eu = ed[i];
                    double t = 0.0;
#pragma unroll
                    for (int l = S - 1; l >= 0; --l) t = fma((double)acc[l][nt][r], level_weight(l), t);
                    t = (eu == kExpNonFinite || eg == kExpNonFinite) ? nan : ldexp(t, eu + eg);
                    part = fma(lc[(long)i * blksize + c], t, part);
                }
            }
            part += __shfl_xor(part, 16);
            part += __shfl_xor(part, 32);
            if (lk == 0) red[wave][nt * 16 + lrow] = part;
        }
        // an LDS-only barrier: __syncthreads() would also wait for the prefetches in flight
        asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
        for (int c = tid; c < W; c += kThreads) {
            const int g = gblk + c0 + c;
            if (g < ngrid) {
                double v = red[0][c];
#pragma unroll
                for (int w = 1; w < kThreads / 64; ++w) v += red[w][c];
                partial[((long)T * nset + s) * ngrid + g] = v;
            }
        }
    };

    // Both operands are prefetched DEPTH k steps ahead: the B stage into one of DEPTH+1 LDS buffers,
    // the A fragments into one of DEPTH+1 register sets, their roles static in a loop unrolled by
    // DEPTH+1 (no copy of a register set, which would wait for its prefetch, and each buffer a
    // fixed object, so that the compiler's wait before the LDS reads of one does not also wait
    // for the copies into the others).  A prefetch index is clamped instead of guarded (a guard
    // makes the compiler wait for every load at its join): past the last step it re-reads that
    // step into an idle buffer, and the values are not used.  Each step issues its B copies, then
    // its A loads: at the start of step t the loads of steps t+1 .. t+DEPTH-1 may stay in flight
    // (vmcnt in bits 3:0, expcnt and lgkmcnt left at their maxima), and the barrier makes the
    // arrival of B(t) true of every wave and tells each that all have finished reading the
    // buffer the copies of step t+DEPTH refill.  In the symmetric form the int32 sums of a tile are
    // doubled before its diagonal step.
    const int last = nsteps - 1;
    constexpr int kVmAhead = (DEPTH - 1) * (kBPerWave + S);  // loads of the steps that may stay in flight
    static_assert(kVmAhead < 64, "vmcnt is a 6-bit count");
    auto step = [&](int t, const i32x4 *fa_cur, i32x4 *fa_new, const int8_t (*bcur)[W * kK],
                    int8_t (*bnew)[W * kK]) {
        // the steps past the diagonal count twice: doubled here, not between the loads and the
        // LDS reads, where the join of the branch would make the compiler wait for every load
        if (sym && (t == n1 - 1 || t == last)) {
#pragma unroll
            for (int l = 0; l < S; ++l)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) acc[l][nt] <<= 1;
        }
        __builtin_amdgcn_s_waitcnt((kVmAhead & 15) | ((kVmAhead >> 4) << 14) | (0x7 << 4) | (0xF << 8));
        asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
        issue_b(min(t + DEPTH, last), bnew);
        load_a(min(t + DEPTH, last), fa_new);
        __builtin_amdgcn_sched_barrier(0);  // keep the prefetch ahead of the MFMAs (not sunk to its use)
        compute(bcur, fa_cur);
    };
    auto after = [&](int t) {
        if (t == n1 - 1 && n2 > 0) {  // first tile of a pair done; the pipeline runs on into the second
            epilogue(T1);
            zero_acc();
        }
    };
    issue_b(0, Bs0);
    load_a(0, fa0);
    if constexpr (DEPTH == 1) {
        int t = 0;
        for (; t + 2 <= nsteps; t += 2) {
            step(t, fa0, fa1, Bs0, Bs1);
            after(t);
            step(t + 1, fa1, fa0, Bs1, Bs0);
            after(t + 1);
        }
        if (t < nsteps) {
            step(t, fa0, fa1, Bs0, Bs1);
            after(t);
        }
    } else {
        issue_b(min(1, last), Bs1);
        load_a(min(1, last), fa1);
        int t = 0;
        for (; t + 3 <= nsteps; t += 3) {
            step(t, fa0, fa2, Bs0, Bs2);
            after(t);
            step(t + 1, fa1, fa0, Bs1, Bs0);
            after(t + 1);
            step(t + 2, fa2, fa1, Bs2, Bs1);
            after(t + 2);
        }
        if (t < nsteps) {
            step(t, fa0, fa2, Bs0, Bs2);
            after(t);
        }
        if (t + 1 < nsteps) {
            step(t + 1, fa1, fa0, Bs1, Bs0);
            after(t + 1);
        }
    }
    __builtin_amdgcn_s_waitcnt(0);  // no copy into LDS outlives the task
    epilogue(n2 > 0 ? T2 : T1);
}

// With the handle's order (every grid block evaluated): workgroup i takes item i / 8 of XCD
// i % 8's list of (task, density matrix) items (the XCD it runs on); items of density matrices
// past the evaluated ones and the padding exit at once.  Without it (leading blocks only): the
// first ntasks tasks, task-major, in contiguous equal-count shares per XCD.
template <int S, int NT, int DEPTH>
__global__ __launch_bounds__(kThreads, 2) void rho_resident_kernel(
    double *__restrict__ partial, const unsigned char *__restrict__ base, const BlockDesc *__restrict__ blocks,
    const int4 *__restrict__ tasks, int ntasks, const int2 *__restrict__ order, int order_len,
    const int *__restrict__ fg, int blksize, int ngrid, int nset) {
    int2 item;
    if (order) {
        item = order[(blockIdx.x % 8) * order_len + blockIdx.x / 8];
        if (item.x < 0 || item.y >= nset) return;
    } else {
        const int idx = xcd_task(blockIdx.x, ntasks * nset);
        item = int2{idx / nset, idx % nset};
    }
    resident_tile<S, NT, DEPTH>(partial, base, blocks, tasks[item.x], item.y, fg, blksize, ngrid, nset);
}

// ---------------------------------------------------------------------------------------------
// Persistent evaluation.  With one workgroup per item, every item starts from an empty pipeline
// (its descriptors, then its first A and B loads, each a full memory latency) and ends with its
// epilogue's loads; at about a dozen k steps per item those bubbles cost as much as the
// streaming.  Here the grid is two workgroups per CU and each walks a fixed list of items (dealt
// at creation, longest first, to the least loaded workgroup of the item's XCD) as one pipeline:
// the prefetch of the next item's first step is issued during the last step of the current one.

// One (task, density matrix) item: everything its k steps and epilogues address, in bytes from
// the handle's base.  Tile q of the item is row tile T[q] (q = 1 only for the paired tile of the
// symmetric form; n[1] = 0 otherwise).
struct ResItem {
    long a[2];    // wave 0's row of 1 KiB A tiles of tile T[q], density matrix s, plane 0
    long rq;      // Rq plane 0 at the item's first column
    long ed;      // ed of density matrix s
    long lc;      // Lc of the block
    int aw[2];    // bytes between consecutive waves' rows of A tiles
    int n[2];     // k steps of each tile
    int T[2];     // row tiles
    int dplane;   // bytes between the slice planes of Dq
    int rplane;   // bytes between the slice planes of Rq
    int c0;       // first column of the item in its block
    int b;        // grid block
    int s;        // density matrix
    int nrows;    // active AOs of the block: rows past them are padding
    int sym;      // symmetric form: the k steps past the diagonal first, the diagonal step last
    int pad;
};
static_assert(sizeof(ResItem) == 96, "ResItem is six 16-byte words");

// An item of a batched persistent evaluation (items of several handles and evaluations in one
// launch): self-contained, with absolute addresses, its evaluation's partial rows and sizes.
struct ResItemX {
    long a[2], rq, ed, lc;  // as in ResItem, absolute addresses
    long part;              // absolute: the evaluation's partial + s * ngrid
    long fg;                // absolute: the handle's per-point exponents
    int aw[2], n[2], T[2];
    int dplane, rplane, c0, b, s, nrows, sym;
    int tstride;            // partial elements between row tiles: nset * ngrid of the evaluation
    int blksize, ngrid;     // the handle's block size, the evaluation's grid points
    int pad[2];
};
static_assert(sizeof(ResItemX) == 128, "ResItemX is eight 16-byte words");

constexpr int kItemsLds = 32;  // item records staged in LDS at a time

// DEPTH: k steps prefetched (1 or 2); the B stages are DEPTH + 1 LDS buffers, the A fragments
// DEPTH + 1 register sets, their roles static in a loop unrolled by DEPTH + 1.
// XR = false: the items of one handle (ResItem, offsets from base), evaluated over its leading
// nblk_used blocks and nset density matrices; XR = true: a batch (ResItemX; the other arguments
// unused).
template <int S, int NT, int DEPTH, bool XR>
__global__ __launch_bounds__(kThreads, 2) void rho_resident_persistent_kernel(
    double *__restrict__ partial, const unsigned char *__restrict__ base, const void *__restrict__ items_v,
    const int *__restrict__ slot_begin, const int *__restrict__ fg, int blksize, int ngrid, int nset, int nblk_used) {
    static_assert(DEPTH == 1 || DEPTH == 2, "one or two k steps of prefetch");
    using Rec = typename std::conditional<XR, ResItemX, ResItem>::type;
    constexpr int kWords = (int)sizeof(Rec) / 16;
    const Rec *__restrict__ items = static_cast<const Rec *>(items_v);
    constexpr int W = NT * 16;
    __shared__ __attribute__((aligned(16))) int8_t Bs0[S][W * kK];
    __shared__ __attribute__((aligned(16))) int8_t Bs1[S][W * kK];
    __shared__ __attribute__((aligned(16))) int8_t Bs2[DEPTH >= 2 ? S : 1][W * kK];
    __shared__ double red[kThreads / 64][W];
    __shared__ __attribute__((aligned(16))) Rec its[kItemsLds];
    __shared__ int its_cnt;

    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lrow = lane & 15, lk = lane >> 4;

    i32x4 acc[S][NT];
    auto zero_acc = [&]() {
#pragma unroll
        for (int l = 0; l < S; ++l)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) acc[l][nt] = i32x4{0, 0, 0, 0};
    };

    // k step t of item it: (tile, local step) -> A tile row and k step (the symmetric form visits
    // the steps past the diagonal first, the diagonal last)
    // (the same for every lane: the lane's 16 bytes are added at the load, as a 32-bit offset
    // beside a scalar base)
    struct StepAddr {
        const int8_t *a;  // the wave's A tile, plane 0
        const int8_t *b;  // the B stage, plane 0, part 0
        int dplane, rplane;
    };
    const unsigned lane16 = lane * 16u;
    // The producer keeps its item's fields in registers and reads the item records in LDS only
    // when it moves on to the next item: a read of LDS while copies into LDS are in flight may
    // make the compiler wait for all of them.
    // (a and rq: offsets from base, or with XR absolute addresses)
    struct ProdItem {
        long a[2], rq, bstep;  // bstep (XR): bytes of one k step of all the block's columns in Rq
        int aw[2], n[2], T[2], dplane, rplane, sym;
    };
    const long bstep = (long)blksize * kK;  // the same for a single handle
    // (every field read once per lane from LDS and made scalar: addresses and branches on them
    // are then worked out by the scalar unit)
    auto prod_item = [&](int it) {
        const Rec &R = its[it];
        ProdItem P;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            P.a[q] = uniform64(R.a[q]);
            P.aw[q] = __builtin_amdgcn_readfirstlane(R.aw[q]);
            P.n[q] = __builtin_amdgcn_readfirstlane(R.n[q]);
            P.T[q] = __builtin_amdgcn_readfirstlane(R.T[q]);
        }
        P.rq = uniform64(R.rq);
        if constexpr (XR) P.bstep = (long)__builtin_amdgcn_readfirstlane(R.blksize) * kK;
        P.dplane = __builtin_amdgcn_readfirstlane(R.dplane);
        P.rplane = __builtin_amdgcn_readfirstlane(R.rplane);
        P.sym = __builtin_amdgcn_readfirstlane(R.sym);
        return P;
    };
    // the consumer's item: what the diagonal doubling, the tile ends and the epilogues need (ed
    // and lc as prod_item's a; with XR also the evaluation's partial rows, exponents and sizes)
    struct ConsItem {
        long ed, lc;
        double *part;
        const int *fgp;
        int n[2], T[2], sym, b, c0, s, nrows, tstride, blksize, ngrid;
    };
    auto cons_item = [&](int it) {
        const Rec &R = its[it];
        ConsItem c;
        c.ed = uniform64(R.ed);
        c.lc = uniform64(R.lc);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            c.n[q] = __builtin_amdgcn_readfirstlane(R.n[q]);
            c.T[q] = __builtin_amdgcn_readfirstlane(R.T[q]);
        }
        c.sym = __builtin_amdgcn_readfirstlane(R.sym);
        c.b = __builtin_amdgcn_readfirstlane(R.b);
        c.c0 = __builtin_amdgcn_readfirstlane(R.c0);
        c.s = __builtin_amdgcn_readfirstlane(R.s);
        c.nrows = __builtin_amdgcn_readfirstlane(R.nrows);
        if constexpr (XR) {
            c.part = reinterpret_cast<double *>(uniform64((long)R.part));
            c.fgp = reinterpret_cast<const int *>(uniform64((long)R.fg));
            c.tstride = __builtin_amdgcn_readfirstlane(R.tstride);
            c.blksize = __builtin_amdgcn_readfirstlane(R.blksize);
            c.ngrid = __builtin_amdgcn_readfirstlane(R.ngrid);
        }
        return c;
    };
    auto addr_of = [&](const ProdItem &P, int t) {
        const int q = t < P.n[0] ? 0 : 1;
        const int l = q ? t - P.n[0] : t, n = P.n[q], T = P.T[q];
        const int ks = P.sym ? (l < n - 1 ? T + 1 + l : T) : l;
        const int ai = P.sym ? (l < n - 1 ? l + 1 : 0) : l;
        StepAddr x;
        if constexpr (XR) {
            x.a = reinterpret_cast<const int8_t *>(P.a[q]) + (long)wave * P.aw[q] + (long)ai * 1024;
            x.b = reinterpret_cast<const int8_t *>(P.rq) + ks * P.bstep;
        } else {
            x.a = reinterpret_cast<const int8_t *>(base + P.a[q]) + (long)wave * P.aw[q] + (long)ai * 1024;
            x.b = reinterpret_cast<const int8_t *>(base + P.rq) + ks * bstep;
        }
        x.dplane = P.dplane;
        x.rplane = P.rplane;
        return x;
    };
    auto load_a = [&](const StepAddr &x, i32x4 *dst) {
#pragma unroll
        for (int p = 0; p < S; ++p)  // each A tile is read once per evaluation: stream it past the caches
            dst[p] = __builtin_nontemporal_load(reinterpret_cast<const i32x4 *>(x.a + (long)p * x.dplane + lane16));
    };
    constexpr int kBParts = S * NT, kBPerWave = (kBParts + 3) / 4;
    auto issue_b = [&](const StepAddr &x, int8_t (*dst)[W * kK]) {
#pragma unroll
        for (int i = 0; i < kBPerWave; ++i) {
            const int j = min(wave + 4 * i, kBParts - 1);
            const int p = j / NT, part = j % NT;
            copy16_to_lds(x.b + (long)p * x.rplane + part * 1024 + lane16, &dst[p][part * 1024]);
        }
    };
    auto compute = [&](const int8_t (*bs)[W * kK], const i32x4 *fa) {
#pragma unroll
        for (int sb = 0; sb < S; ++sb) {
            i32x4 fb[NT];
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                fb[nt] = *reinterpret_cast<const i32x4 *>(&bs[sb][swz(nt * 16 + lrow, lk)]);
#pragma unroll
            for (int sa = 0; sa + sb < S; ++sa)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                    acc[sa + sb][nt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[sa], fb[nt], acc[sa + sb][nt], 0, 0, 0);
        }
    };
    // Epilogue of row tile T of item it (the same arithmetic as resident_tile's), in two parts: its
    // loads, issued at the top of the tile's last k step ahead of that step's prefetch (so that they
    // arrive during its MFMAs and the wait for them leaves the prefetch in flight), and the rest.
    // No load sits under a condition (a load inside a divergent branch is waited for there, one
    // after another): rows i < kpad and columns c < blksize are inside ed and Lc (zero padded),
    // the grid point is clamped.
    struct EpiLoads {
        int eu[4], eg[NT];
        double lv[NT][4];
    };
    double pend_v = 0.0;           // a partial value of the last epilogue, not yet stored
    double *pend_dst = nullptr;    // ... and where it goes (threads < W)
    // the item's block size, grid points, exponents and partial row (tile T, grid point 0)
    auto c_blksize = [&](const ConsItem &R) { if constexpr (XR) return R.blksize; else return blksize; };
    auto c_ngrid = [&](const ConsItem &R) { if constexpr (XR) return R.ngrid; else return ngrid; };
    auto c_fg = [&](const ConsItem &R) { if constexpr (XR) return R.fgp; else return fg; };
    auto c_part = [&](const ConsItem &R, int T) {
        if constexpr (XR) return R.part + (long)T * R.tstride;
        else return partial + ((long)T * nset + R.s) * ngrid;
    };
    auto epi_load = [&](const ConsItem &R, int T, EpiLoads &e) {
        const unsigned char *b0 = XR ? nullptr : base;
        const int *__restrict__ ed = reinterpret_cast<const int *>(b0 + R.ed);
        const double *__restrict__ lc = reinterpret_cast<const double *>(b0 + R.lc);
        const int bs = c_blksize(R), ng = c_ngrid(R);
        const int *__restrict__ fgp = c_fg(R);
        const int gblk = R.b * bs, i0 = T * kRows + wave * 16 + 4 * lk;
#pragma unroll
        for (int r = 0; r < 4; ++r) e.eu[r] = ed[i0 + r];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int c = R.c0 + nt * 16 + lrow;
            e.eg[nt] = fgp[min(gblk + c, ng - 1)];
#pragma unroll
            for (int r = 0; r < 4; ++r) e.lv[nt][r] = lc[(long)(i0 + r) * bs + c];
        }
    };
    auto epilogue = [&](const ConsItem &R, int T, const EpiLoads &e) {
        const int ng = c_ngrid(R);
        const int gblk = R.b * c_blksize(R), c0 = R.c0, nrows = R.nrows;
        const int i0 = T * kRows + wave * 16 + 4 * lk;
        const double nan = __builtin_nan("");
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int c = c0 + nt * 16 + lrow;
            const bool g_ok = gblk + c < ng;
            double part = 0.0;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                double t = 0.0;
#pragma unroll
                for (int l = S - 1; l >= 0; --l) t = fma((double)acc[l][nt][r], level_weight(l), t);
                t = (e.eu[r] == kExpNonFinite || e.eg[nt] == kExpNonFinite) ? nan : ldexp(t, e.eu[r] + e.eg[nt]);
                part = i0 + r < nrows && g_ok ? fma(e.lv[nt][r], t, part) : part;
            }
            part += __shfl_xor(part, 16);
            part += __shfl_xor(part, 32);
            if (lk == 0) red[wave][nt * 16 + lrow] = part;
        }
        asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");  // LDS only: prefetches stay in flight
        // thread c < W sums column c; its store is issued by the next step after that step's prefetch
        // (stores count in vmcnt too: stored here, the next step's wait for its operands would also
        // wait for the store)
        if (tid < W) {
            const int g = gblk + c0 + tid;
            double v = red[0][tid];
#pragma unroll
            for (int w = 1; w < kThreads / 64; ++w) v += red[w][tid];
            pend_v = v;
            pend_dst = g < ng ? c_part(R, T) + g : nullptr;
        }
    };

    const int i_begin = slot_begin[blockIdx.x], i_end = slot_begin[blockIdx.x + 1];
    for (int c_begin = i_begin; c_begin < i_end; c_begin += kItemsLds) {
        const int raw = min(kItemsLds, i_end - c_begin);
        __syncthreads();  // the previous chunk's readers of its, Bs and red are done, its loads drained
        // the chunk's items inside the evaluated range (grid blocks < nblk_used, density matrices
        // < nset), compacted in order by wave 0: lane l keeps item l
        if (wave == 0) {
            const Rec *__restrict__ src = items + c_begin + min(lane, raw - 1);
            const bool keep = lane < raw && (XR || (src->b < nblk_used && src->s < nset));
            const unsigned long long m = __ballot(keep);
            if (keep) {
                const int pos = __popcll(m & ((1ull << lane) - 1ull));
#pragma unroll
                for (int u = 0; u < kWords; ++u)
                    reinterpret_cast<int4 *>(its + pos)[u] = reinterpret_cast<const int4 *>(src)[u];
            }
            if (lane == 0) its_cnt = __popcll(m);
        }
        __syncthreads();
        const int cnt = __builtin_amdgcn_readfirstlane(its_cnt);
        if (cnt == 0) continue;
        int total = 0;
        for (int it = 0; it < cnt; ++it) total += its[it].n[0] + its[it].n[1];
        total = __builtin_amdgcn_readfirstlane(total);

        // producer (the step whose loads are issued next) and consumer (the step computed next)
        int pi = 0, pt = 0, ci = 0, ct = 0;
        ProdItem P = prod_item(0);
        auto advance_p = [&]() {
            if (++pt == P.n[0] + P.n[1] && pi + 1 < cnt) {
                ++pi;
                pt = 0;
                P = prod_item(pi);
            }
        };
        ConsItem C = cons_item(0);
        auto advance_c = [&]() {
            if (++ct == C.n[0] + C.n[1] && ci + 1 < cnt) {
                ++ci;
                ct = 0;
                C = cons_item(ci);
            }
        };
        i32x4 fa0[S], fa1[S], fa2[DEPTH >= 2 ? S : 1];
        zero_acc();
        StepAddr nx = addr_of(P, 0);
        issue_b(nx, Bs0);
        load_a(nx, fa0);
        if (total > 1) advance_p();
        nx = addr_of(P, pt);
        if constexpr (DEPTH >= 2) {
            issue_b(nx, Bs1);
            load_a(nx, fa1);
            if (total > 2) advance_p();
            nx = addr_of(P, pt);
        }
        // nx: what step 0 prefetches, step DEPTH (or the last step again)
        // Step j: wait for its operands, prefetch step j+DEPTH (past the last step the last again,
        // into an idle buffer, unused; its addresses were worked out at the end of step j-1, off
        // the path from the barrier to the loads), compute, then the epilogue of a tile that ends
        // here.  The loads of step j+1 may stay in flight at the wait: kBPerWave + S operations
        // per step, and the epilogue loads and the deferred store of a step are issued before
        // and after its prefetch, so that waiting for all but the newest kBPerWave + S leaves at
        // most one operation more than needed.
        constexpr int kVmAhead = (DEPTH - 1) * (kBPerWave + S);
        static_assert(kVmAhead < 64, "vmcnt is a 6-bit count");
        auto step = [&](int j, const i32x4 *fa_cur, i32x4 *fa_new, const int8_t (*bcur)[W * kK],
                        int8_t (*bnew)[W * kK]) {
            const int n0 = C.n[0], n1 = C.n[1];
            if (C.sym && (ct == n0 - 1 || (n1 > 0 && ct == n0 + n1 - 1))) {  // the diagonal step: past it x2
#pragma unroll
                for (int l = 0; l < S; ++l)
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt) acc[l][nt] <<= 1;
            }
            __builtin_amdgcn_s_waitcnt((kVmAhead & 15) | ((kVmAhead >> 4) << 14) | (0x7 << 4) | (0xF << 8));
            asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
            // a tile ends with this step: its epilogue's loads first
            const bool epi = ct == n0 - 1 || (n1 > 0 && ct == n0 + n1 - 1);
            const int Te = ct == n0 - 1 ? C.T[0] : C.T[1];
            EpiLoads el;
            if (epi) epi_load(C, Te, el);
            issue_b(nx, bnew);
            load_a(nx, fa_new);
            if (pend_dst) {  // the previous epilogue's value, behind this step's prefetch
                *pend_dst = pend_v;
                pend_dst = nullptr;
            }
            __builtin_amdgcn_sched_barrier(0);  // keep the prefetch ahead of the MFMAs
            compute(bcur, fa_cur);
            if (epi) {
                // the epilogue's loads were issued before this step's prefetch (kBPerWave + S
                // operations, and a pending store): wait for them only, the prefetch stays in flight
                __builtin_amdgcn_s_waitcnt((kBPerWave + S) | (0x7 << 4) | (0xF << 8));
                epilogue(C, Te, el);
                zero_acc();
            }
            advance_c();
            if (j + DEPTH + 1 < total) advance_p();
            nx = addr_of(P, pt);
        };
        // a counted loop over groups of DEPTH + 1 steps: with `j += 2` the compiler peeled a step
        // and left one per iteration, whose register set then had to be copied (waiting for its
        // prefetch)
        if constexpr (DEPTH == 1) {
            const int npairs = total >> 1;
            for (int q = 0; q < npairs; ++q) {
                step(2 * q, fa0, fa1, Bs0, Bs1);
                step(2 * q + 1, fa1, fa0, Bs1, Bs0);
            }
            if (total & 1) step(total - 1, fa0, fa1, Bs0, Bs1);
        } else {
            const int ntri = total / 3;
            for (int q = 0; q < ntri; ++q) {
                step(3 * q, fa0, fa2, Bs0, Bs2);
                step(3 * q + 1, fa1, fa0, Bs1, Bs0);
                step(3 * q + 2, fa2, fa1, Bs2, Bs1);
            }
            const int r = total - 3 * ntri;
            if (r >= 1) step(3 * ntri, fa0, fa2, Bs0, Bs2);
            if (r >= 2) step(3 * ntri + 1, fa1, fa0, Bs1, Bs0);
        }
        if (pend_dst) {  // the last epilogue's value
            *pend_dst = pend_v;
            pend_dst = nullptr;
        }
        __builtin_amdgcn_s_waitcnt(0);  // no copy into LDS outlives the chunk
    }
}

// One evaluation of a batch (device table).
struct EvalDesc {
    const unsigned char *base;
    const BlockDesc *blocks;
    const int4 *tasks;
    const int *fg, *counts;
    double *partial, *rhos;
    long stride_rho;
    int blksize, ngrid, nset, ntasks;
    int wg_begin;  // first workgroup of this evaluation in the batch grid
};

// Every (task, density matrix) of every evaluation in one grid.  The XCD-contiguous order runs
// over the whole batch, evaluations laid out one after another, so the tiles of a block still
// meet on one XCD.
template <int S, int NT, int DEPTH>
__global__ __launch_bounds__(kThreads, 2) void rho_resident_batched_kernel(const EvalDesc *__restrict__ evals,
                                                                           int nevals) {
    const int L = xcd_task(blockIdx.x, gridDim.x);
    int lo = 0, hi = nevals - 1;  // last evaluation with wg_begin <= L
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (evals[mid].wg_begin <= L) lo = mid;
        else hi = mid - 1;
    }
    const EvalDesc e = evals[lo];
    const int t = L - e.wg_begin;  // task-major (task, density matrix), as in the single evaluation
    resident_tile<S, NT, DEPTH>(e.partial, e.base, e.blocks, e.tasks[t / e.nset], t % e.nset, e.fg, e.blksize,
                                e.ngrid, e.nset);
}

// rho[s*stride + g] += sum over the 64-row tiles of g's block of partial[tile][s][g], per evaluation.
__global__ void resident_batched_reduce(const EvalDesc *__restrict__ evals) {
    const EvalDesc e = evals[blockIdx.y];
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)e.nset * e.ngrid) return;
    const int s = (int)(idx / e.ngrid), g = (int)(idx % e.ngrid);
    const int nch = (e.counts[g / e.blksize] + kRows - 1) / kRows;
    const double *__restrict__ part = e.partial + (long)s * e.ngrid + g;
    const long step = (long)e.nset * e.ngrid;
    // eight loads in flight at a time (clamped indices, no branch around a load), summed in tile order
    double v = 0.0;
    for (int c0 = 0; c0 < nch; c0 += 8) {
        double x[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) x[i] = part[min(c0 + i, nch - 1) * step];
#pragma unroll
        for (int i = 0; i < 8; ++i) v = c0 + i < nch ? v + x[i] : v;
    }
    e.rhos[(long)s * e.stride_rho + g] += v;
}

int tile_nt(size_t blksize) {
    const size_t q = blksize / 16;
    for (int nt = 4; nt >= 1; --nt)
        if (q % nt == 0) return nt;
    return 1;
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct ResidentRho {
    int device = 0;
    int nslices = 0;
    size_t nao = 0, ngrid = 0, nset = 0, blksize = 0, nblk = 0;
    int nt = 4, nsub = 1;
    int max_tiles = 0;
    bool symmetric = false;         // the upper-triangle form
    unsigned char *base = nullptr;  // one allocation: tiles, tables, exponents
    size_t bytes = 0;
    BlockDesc *d_blocks = nullptr;
    int *d_counts = nullptr;
    int *d_fg = nullptr;
    int4 *d_tasks = nullptr;
    std::vector<int> task_prefix;  // tasks of blocks < b, for b = 0..nblk
    int2 *d_order = nullptr;        // [8][order_len] (task, density matrix) items per XCD, (-1, 0) padded
    int order_len = 0;
    // persistent evaluation: the items of workgroup i are d_items[d_slot_begin[i] .. d_slot_begin[i+1])
    int nslots = 0;
    ResItem *d_items = nullptr;
    std::vector<ResItem> h_items;  // every item, task-major (block order), for batched evaluations
    unsigned long long uid = 0;    // unique per handle created in this process (batch plan cache keys)
    int *d_slot_begin = nullptr;
};

template <int S>
static int prep_launch(hipStream_t stream, ResidentRho *h, const double *dms, const double *ao_lhs,
                       const double *ao_rhs, size_t stride_ao, size_t stride_grid, const int *lists, int max_kpad) {
    hipEvent_t t0 = trace_start(stream);
    hipLaunchKernelGGL((prep_dq_kernel<S>), dim3((unsigned)div_up(max_kpad, 16), (unsigned)h->nblk,
                                                 (unsigned)h->nset),
                       dim3(kThreads), 0, stream, h->base, h->d_blocks, dms, (long)stride_ao, (int)h->nao, lists);
    hipLaunchKernelGGL((prep_phi_kernel<S>), dim3((unsigned)h->nblk), dim3(kThreads), sizeof(u64) * h->blksize, stream,
                       h->base, h->d_blocks, ao_lhs, ao_rhs, (long)stride_grid, (int)h->ngrid, (int)h->nao,
                       (int)h->blksize, lists, h->d_fg);
    trace_stop(stream, "rho_resident_prep", t0);
    return hip_status(hipGetLastError(), "rho resident prep launch");
}

int resident_create(hipStream_t stream, const double *dms, const double *ao_lhs, const double *ao_rhs,
                    size_t stride_ao, size_t stride_grid, size_t nao, size_t ngrid, size_t nset, const uint8_t *tab,
                    size_t ldtab, size_t blksize, int nslices, ResidentRho **out) {
    *out = nullptr;
    if (nslices < 2 || nslices > 6) return fail(CFGMM_EINVAL, "nslices must be in [2, 6]");
    if (nao > INT32_MAX || ngrid > INT32_MAX || nset > 65535 || blksize > 4096)
        return fail(CFGMM_EINVAL, "resident rho: size out of range");
    auto h = new ResidentRho();
    static std::atomic<unsigned long long> next_uid{1};
    h->uid = next_uid++;
    CFGMM_HIP_TRY(hipGetDevice(&h->device));
    h->nslices = nslices;
    h->nao = nao;
    h->ngrid = ngrid;
    h->nset = nset;
    h->blksize = blksize;
    h->nblk = ngrid ? div_up(ngrid, blksize) : 0;
    h->nt = tile_nt(blksize);
    const int want_nt = tune_get(CFGMM_TUNE_RESIDENT_COLS) / 16;  // 0: the widest tile
    if (want_nt > 0 && want_nt < h->nt && (blksize / 16) % want_nt == 0) h->nt = want_nt;
    h->nsub = (int)(blksize / (16 * h->nt));
    const size_t nblk = h->nblk;
    if (nao == 0 || ngrid == 0 || nset == 0) {
        h->task_prefix.assign(nblk + 1, 0);
        *out = h;
        return CFGMM_OK;
    }

    // 1. compacted AO lists of every block (device), counts to the host; with one array on both
    //    sides, whether every D_s is symmetric
    void *ws = nullptr;
    const size_t lists_bytes = sizeof(int) * nblk * nao;
    int st = workspace_get(stream, lists_bytes + sizeof(int) * (nblk + 1), &ws);
    if (st != CFGMM_OK) {
        delete h;
        return st;
    }
    int *lists = static_cast<int *>(ws);
    int *counts = lists + nblk * nao;
    int *dflag = counts + nblk;
    const bool try_sym = ao_lhs == ao_rhs && tune_get(CFGMM_TUNE_RHO_SYMMETRIC) == 0;
    int hflag = try_sym ? 1 : 0;
    st = launch_build_lists(stream, tab, ldtab, nao, nblk, lists, counts);
    if (st == CFGMM_OK && try_sym) {
        st = hip_status(hipMemcpyAsync(dflag, &hflag, sizeof(int), hipMemcpyHostToDevice, stream), "resident rho: flag");
        if (st == CFGMM_OK) st = launch_dm_symmetry(stream, dms, stride_ao, nao, nset, dflag);
        if (st == CFGMM_OK)
            st = hip_status(hipMemcpyAsync(&hflag, dflag, sizeof(int), hipMemcpyDeviceToHost, stream),
                            "resident rho: flag D2H");
    }
    std::vector<int> hcounts(nblk);
    if (st == CFGMM_OK)
        st = hip_status(hipMemcpyAsync(hcounts.data(), counts, sizeof(int) * nblk, hipMemcpyDeviceToHost, stream),
                        "resident rho: counts D2H");
    if (st == CFGMM_OK) st = hip_status(hipStreamSynchronize(stream), "resident rho: sync");
    if (st != CFGMM_OK) {
        delete h;
        return st;
    }
    const int max_n = nblk ? *std::max_element(hcounts.begin(), hcounts.end()) : 0;
    h->symmetric = hflag != 0 && (int)div_up((size_t)max_n, kK) * kK <= kMaxKpadSym;

    // 2. layout
    auto al = [](size_t x) { return div_up(x, 256) * 256; };
    const int S = nslices;
    std::vector<BlockDesc> blocks(nblk);
    std::vector<int4> tasks;
    h->task_prefix.assign(nblk + 1, 0);
    size_t off = 0;
    int max_kpad = 0;
    for (size_t b = 0; b < nblk; ++b) {
        const int n = hcounts[b];
        const int kpad = (int)div_up((size_t)n, kK) * kK;
        if (kpad > kMaxKpad) {
            delete h;
            return fail(CFGMM_EINVAL, "resident rho: a grid block has more than 16384 active AOs");
        }
        max_kpad = std::max(max_kpad, kpad);
        BlockDesc &bd = blocks[b];
        bd.n = n;
        bd.kpad = kpad;
        bd.sym = h->symmetric ? 1 : 0;
        bd.d_off = (long)off;
        off += al(nset * S * (size_t)dq_plane_bytes(kpad, bd.sym));
        bd.r_off = (long)off;
        off += al((size_t)S * blksize * kpad);
        bd.l_off = (long)off;
        off += al(sizeof(double) * kpad * blksize);
        bd.e_off = (long)off;
        off += al(sizeof(int) * nset * kpad);
        const int ntiles = kpad / kRows;
        h->max_tiles = std::max(h->max_tiles, ntiles);
        // symmetric form: tile t paired with ntiles-1-t (nks+1 k steps per task)
        const int ntask_tiles = h->symmetric ? (ntiles + 1) / 2 : ntiles;
        for (int t = 0; t < ntask_tiles; ++t)
            for (int sub = 0; sub < h->nsub; ++sub) tasks.push_back(int4{(int)b, t, sub, h->symmetric ? 1 : 0});
        h->task_prefix[b + 1] = (int)tasks.size();
    }
    const size_t o_blocks = off;
    off += al(sizeof(BlockDesc) * nblk);
    const size_t o_counts = off;
    off += al(sizeof(int) * nblk);
    const size_t o_fg = off;
    off += al(sizeof(int) * nblk * blksize);
    const size_t o_tasks = off;
    off += al(sizeof(int4) * std::max<size_t>(tasks.size(), 1));
    // Evaluation order of the (task, density matrix) items.  Workgroup i runs on XCD i % 8 and
    // an XCD starts its workgroups in index order as slots free up; a contiguous equal-count share
    // per XCD leaves the XCDs with unequal work (a block's tasks take nks + 1 k steps each) and
    // ends on long tasks.  Whole blocks (their tasks share Rq and Lc in the XCD's L2) go to the
    // XCD with the least work so far, largest first, and each XCD's items run longest first.
    std::vector<int2> order;  // [8][order_len], copied to the device below
    std::vector<ResItem> ritems;  // persistent evaluation: items of slot 0, then of slot 1, ...
    std::vector<int> slot_begin;
    {
        std::vector<std::vector<int2>> xcd_items(8);
        auto steps_of = [&](const int4 &t) {  // the k steps resident_tile runs for task t
            const int kp = blocks[t.x].kpad, nks = kp / kK, ntl = kp / kRows;
            if (!t.w) return nks;
            const int T2 = ntl - 1 - t.y;
            return (nks - t.y) + (T2 > t.y ? nks - T2 : 0);
        };
        std::vector<int> steps(tasks.size());
        for (size_t i = 0; i < tasks.size(); ++i) steps[i] = steps_of(tasks[i]);
        std::vector<int> order_b;
        std::vector<long> work_b(nblk, 0);
        for (size_t b = 0; b < nblk; ++b) {
            for (int i = h->task_prefix[b]; i < h->task_prefix[b + 1]; ++i) work_b[b] += (long)steps[i] * nset;
            if (h->task_prefix[b + 1] > h->task_prefix[b]) order_b.push_back((int)b);
        }
        std::stable_sort(order_b.begin(), order_b.end(), [&](int x, int y) { return work_b[x] > work_b[y]; });
        long load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int b : order_b) {
            const int x = (int)(std::min_element(load, load + 8) - load);
            load[x] += work_b[b];
            for (int i = h->task_prefix[b]; i < h->task_prefix[b + 1]; ++i)
                for (size_t sset = 0; sset < nset; ++sset) xcd_items[x].push_back(int2{i, (int)sset});
        }
        for (auto &l : xcd_items) {
            std::stable_sort(l.begin(), l.end(), [&](const int2 &u, const int2 &v) { return steps[u.x] > steps[v.x]; });
            h->order_len = std::max(h->order_len, (int)l.size());
        }
        order.assign((size_t)8 * h->order_len, int2{-1, 0});
        for (int x = 0; x < 8; ++x) std::copy(xcd_items[x].begin(), xcd_items[x].end(), order.begin() + (size_t)x * h->order_len);

        // Persistent evaluation: two workgroups per CU, workgroup i on XCD i % 8; each XCD's items,
        // longest first, go to its least loaded workgroup.
        const int per_xcd = std::max(1, 2 * device_cu_count() / 8);
        h->nslots = 8 * per_xcd;
        std::vector<std::vector<int2>> slot_items((size_t)h->nslots);
        for (int x = 0; x < 8; ++x) {
            std::vector<long> wl((size_t)per_xcd, 0);
            for (const int2 &it : xcd_items[x]) {
                const int k = (int)(std::min_element(wl.begin(), wl.end()) - wl.begin());
                wl[(size_t)k] += steps[(size_t)it.x];
                slot_items[(size_t)x + 8 * (size_t)k].push_back(it);
            }
        }
        const int W = 16 * h->nt;
        // the record of (task i, density matrix s): offsets from the handle's base
        auto make_item = [&](int i, int s) {
            const int4 &t = tasks[(size_t)i];
            const BlockDesc &bd = blocks[(size_t)t.x];
            const int kp = bd.kpad, nks = kp / kK, ntl = kp / kRows;
            const long dplane = dq_plane_bytes(kp, bd.sym);
            // first 1 KiB tile of wave 0's row of tile T, and the bytes between the waves' rows
            auto arow = [&](int T) { return bd.sym ? sym_tiles_before(T, nks) : (long)T * (kRows / 16) * nks; };
            auto awave = [&](int T) { return (bd.sym ? nks - T : nks) * 1024; };
            const int T1 = t.y, T2 = bd.sym && t.w ? ntl - 1 - t.y : -1;
            const int n1 = bd.sym ? nks - T1 : nks, n2 = T2 > T1 ? nks - T2 : 0;
            const int Tq[2] = {T1, n2 > 0 ? T2 : T1};
            ResItem r{};
            for (int q = 0; q < 2; ++q) {
                r.a[q] = bd.d_off + (long)s * S * dplane + arow(Tq[q]) * 1024;
                r.aw[q] = awave(Tq[q]);
                r.T[q] = Tq[q];
            }
            r.n[0] = n1;
            r.n[1] = n2;
            r.rq = bd.r_off + (long)t.z * W * kK;
            r.ed = bd.e_off + (long)s * kp * (long)sizeof(int);
            r.lc = bd.l_off;
            r.dplane = (int)dplane;
            r.rplane = (int)(blksize * (size_t)kp);
            r.c0 = t.z * W;
            r.b = t.x;
            r.s = s;
            r.nrows = bd.n;
            r.sym = bd.sym;
            return r;
        };
        ritems.reserve(tasks.size() * nset);
        slot_begin.assign((size_t)h->nslots + 1, 0);
        for (int i = 0; i < h->nslots; ++i) {
            for (const int2 &it : slot_items[(size_t)i]) ritems.push_back(make_item(it.x, it.y));
            slot_begin[(size_t)i + 1] = (int)ritems.size();
        }
        // batched evaluations deal the items of several handles per call: task-major, blocks in order
        h->h_items.reserve(tasks.size() * nset);
        for (size_t i = 0; i < tasks.size(); ++i)
            for (size_t sset = 0; sset < nset; ++sset) h->h_items.push_back(make_item((int)i, (int)sset));
    }
    if ((size_t)8 * h->order_len > INT32_MAX) {
        delete h;
        return fail(CFGMM_EINVAL, "resident rho: too many (task, density matrix) items");
    }
    const size_t o_order = off;
    off += al(sizeof(int2) * 8 * std::max(h->order_len, 1));
    const size_t o_items = off;
    off += al(sizeof(ResItem) * std::max<size_t>(ritems.size(), 1));
    const size_t o_slots = off;
    off += al(sizeof(int) * slot_begin.size());
    h->bytes = off;

    // 3. allocate, upload the tables, slice
    st = hip_status(hipMalloc(&h->base, h->bytes), "resident rho: hipMalloc");
    if (st != CFGMM_OK) {
        h->base = nullptr;
        delete h;
        return st;
    }
    h->d_blocks = reinterpret_cast<BlockDesc *>(h->base + o_blocks);
    h->d_counts = reinterpret_cast<int *>(h->base + o_counts);
    h->d_fg = reinterpret_cast<int *>(h->base + o_fg);
    h->d_tasks = reinterpret_cast<int4 *>(h->base + o_tasks);
    h->d_order = reinterpret_cast<int2 *>(h->base + o_order);
    h->d_items = reinterpret_cast<ResItem *>(h->base + o_items);
    h->d_slot_begin = reinterpret_cast<int *>(h->base + o_slots);
    st = hip_status(hipMemcpyAsync(h->d_blocks, blocks.data(), sizeof(BlockDesc) * nblk, hipMemcpyHostToDevice, stream),
                    "resident rho: table H2D");
    if (st == CFGMM_OK)
        st = hip_status(hipMemcpyAsync(h->d_counts, hcounts.data(), sizeof(int) * nblk, hipMemcpyHostToDevice, stream),
                        "resident rho: counts H2D");
    if (st == CFGMM_OK && !tasks.empty())
        st = hip_status(hipMemcpyAsync(h->d_tasks, tasks.data(), sizeof(int4) * tasks.size(), hipMemcpyHostToDevice,
                                       stream),
                        "resident rho: tasks H2D");
    if (st == CFGMM_OK && !order.empty())
        st = hip_status(hipMemcpyAsync(h->d_order, order.data(), sizeof(int2) * order.size(), hipMemcpyHostToDevice,
                                       stream),
                        "resident rho: order H2D");
    if (st == CFGMM_OK && !ritems.empty())
        st = hip_status(hipMemcpyAsync(h->d_items, ritems.data(), sizeof(ResItem) * ritems.size(),
                                       hipMemcpyHostToDevice, stream),
                        "resident rho: items H2D");
    if (st == CFGMM_OK)
        st = hip_status(hipMemcpyAsync(h->d_slot_begin, slot_begin.data(), sizeof(int) * slot_begin.size(),
                                       hipMemcpyHostToDevice, stream),
                        "resident rho: slots H2D");
    if (st == CFGMM_OK && max_kpad > 0) {  // an all-empty mask has no tiles to fill
        switch (S) {
            case 2: st = prep_launch<2>(stream, h, dms, ao_lhs, ao_rhs, stride_ao, stride_grid, lists, max_kpad); break;
            case 3: st = prep_launch<3>(stream, h, dms, ao_lhs, ao_rhs, stride_ao, stride_grid, lists, max_kpad); break;
            case 4: st = prep_launch<4>(stream, h, dms, ao_lhs, ao_rhs, stride_ao, stride_grid, lists, max_kpad); break;
            case 5: st = prep_launch<5>(stream, h, dms, ao_lhs, ao_rhs, stride_ao, stride_grid, lists, max_kpad); break;
            default: st = prep_launch<6>(stream, h, dms, ao_lhs, ao_rhs, stride_ao, stride_grid, lists, max_kpad); break;
        }
    }
    // the host tables above must outlive their copies
    if (st == CFGMM_OK) st = hip_status(hipStreamSynchronize(stream), "resident rho: sync");
    if (st != CFGMM_OK) {
        (void)hipStreamSynchronize(stream);
        (void)hipFree(h->base);
        delete h;
        return st;
    }
    *out = h;
    return CFGMM_OK;
}

int resident_destroy(ResidentRho *h) {
    if (!h) return CFGMM_OK;
    int st = CFGMM_OK;
    if (h->base) st = hip_status(hipFree(h->base), "resident rho: hipFree");
    delete h;
    return st;
}

size_t resident_bytes(const ResidentRho *h) { return h ? h->bytes : 0; }

size_t resident_bytes_bound(const uint8_t *tab, size_t ldtab, size_t nao, size_t nblk, size_t nset, size_t blksize,
                            int nslices, bool symmetric) {
    // the layout of resident_create, from the mask on the host: the upper-triangle D planes when
    // the caller knows every D_s to be symmetric (and lhs is rhs) and no block is too long for
    // that form, else full planes
    auto al = [](size_t x) { return div_up(x, 256) * 256; };
    std::vector<size_t> counts(nblk, 0);
    for (size_t a = 0; a < nao; ++a)
        for (size_t b = 0; b < nblk; ++b) counts[b] += tab[a * ldtab + b] != 0;
    size_t max_kpad = 0;
    for (size_t b = 0; b < nblk; ++b) max_kpad = std::max(max_kpad, div_up(counts[b], (size_t)kK) * kK);
    const int sym = symmetric && max_kpad <= (size_t)kMaxKpadSym ? 1 : 0;
    size_t bytes = al(sizeof(BlockDesc) * nblk) + al(sizeof(int) * nblk) + al(sizeof(int) * nblk * blksize) +
                   al(sizeof(int2) * 8 * (1 + nblk * div_up(max_kpad, (size_t)kRows) * div_up(blksize, (size_t)16) * nset));
    const size_t S = (size_t)nslices;
    for (size_t b = 0; b < nblk; ++b) {
        const size_t kpad = div_up(counts[b], (size_t)kK) * kK;
        bytes += al(nset * S * (size_t)dq_plane_bytes((int)kpad, sym)) + al(S * blksize * kpad) +
                 al(sizeof(double) * kpad * blksize) + al(sizeof(int) * nset * kpad) +
                 al(sizeof(int4) * (kpad / kRows) * div_up(blksize, (size_t)16));
    }
    return bytes;
}
size_t resident_ngrid(const ResidentRho *h) { return h->ngrid; }
size_t resident_nset(const ResidentRho *h) { return h->nset; }
size_t resident_blksize(const ResidentRho *h) { return h->blksize; }
int resident_device(const ResidentRho *h) { return h->device; }

template <int S>
static int eval_launch(hipStream_t stream, const ResidentRho *h, double *partial, int ntasks, int ngrid, int nset) {
    hipEvent_t t0 = trace_start(stream);
    const int depth = tune_get(CFGMM_TUNE_RESIDENT_DEPTH);
    const bool all = ntasks == h->task_prefix[h->nblk];
    // The persistent kernel walks the items dealt at creation, skipping those outside a leading
    // range: taken when at least half of them are evaluated (each workgroup's list then keeps
    // roughly its share; a short leading range is better spread by one workgroup per item).
    const long nitems_all = (long)h->task_prefix[h->nblk] * (long)h->nset;
    const bool persistent = tune_get(CFGMM_TUNE_RESIDENT_PERSISTENT) == 0 && 2L * ntasks * nset >= nitems_all;
    if (persistent) {
        const int nblk_used = (int)div_up((size_t)ngrid, h->blksize);
#define CFGMM_RESP_LAUNCH(NT_, D_)                                                                                 \
    hipLaunchKernelGGL((rho_resident_persistent_kernel<S, NT_, D_, false>), dim3((unsigned)h->nslots), dim3(kThreads), 0, \
                       stream, partial, h->base, h->d_items, h->d_slot_begin, h->d_fg, (int)h->blksize, ngrid, nset, \
                       nblk_used)
        if (depth >= 2 && S <= 5) {  // with six slices the third register set spills
            switch (h->nt) {
                case 1: CFGMM_RESP_LAUNCH(1, 2); break;
                case 2: CFGMM_RESP_LAUNCH(2, 2); break;
                case 3: CFGMM_RESP_LAUNCH(3, 2); break;
                default: CFGMM_RESP_LAUNCH(4, 2); break;
            }
        } else {
            switch (h->nt) {
                case 1: CFGMM_RESP_LAUNCH(1, 1); break;
                case 2: CFGMM_RESP_LAUNCH(2, 1); break;
                case 3: CFGMM_RESP_LAUNCH(3, 1); break;
                default: CFGMM_RESP_LAUNCH(4, 1); break;
            }
        }
#undef CFGMM_RESP_LAUNCH
        trace_stop(stream, "rho_resident_kernel", t0);
        return hip_status(hipGetLastError(), "rho_resident_persistent_kernel launch");
    }
    const dim3 grid((unsigned)(all ? 8 * h->order_len : ntasks * nset));
    const int2 *order = all ? h->d_order : nullptr;
#define CFGMM_RES_LAUNCH(NT_, D_)                                                                                    \
    hipLaunchKernelGGL((rho_resident_kernel<S, NT_, D_>), grid, dim3(kThreads), 0, stream, partial, h->base,          \
                       h->d_blocks, h->d_tasks, ntasks, order, h->order_len, h->d_fg, (int)h->blksize, ngrid, nset)
    if (depth >= 2) {
        switch (h->nt) {
            case 1: CFGMM_RES_LAUNCH(1, 2); break;
            case 2: CFGMM_RES_LAUNCH(2, 2); break;
            case 3: CFGMM_RES_LAUNCH(3, 2); break;
            default: CFGMM_RES_LAUNCH(4, 2); break;
        }
    } else {
        switch (h->nt) {
            case 1: CFGMM_RES_LAUNCH(1, 1); break;
            case 2: CFGMM_RES_LAUNCH(2, 1); break;
            case 3: CFGMM_RES_LAUNCH(3, 1); break;
            default: CFGMM_RES_LAUNCH(4, 1); break;
        }
    }
#undef CFGMM_RES_LAUNCH
    trace_stop(stream, "rho_resident_kernel", t0);
    return hip_status(hipGetLastError(), "rho_resident_kernel launch");
}

// rho[s*stride_rho + g] += ... for g < ngrid_used, s < nset_used (leading grid points and density matrices).
int resident_eval(hipStream_t stream, const ResidentRho *h, double *rhos, size_t stride_rho, size_t ngrid_used,
                  size_t nset_used) {
    if (ngrid_used == 0 || nset_used == 0 || h->nao == 0) return CFGMM_OK;
    const size_t nblk_used = div_up(ngrid_used, h->blksize);
    const int ntasks = h->task_prefix[nblk_used];
    if (ntasks == 0) return CFGMM_OK;
    if (nset_used > 65535) return fail(CFGMM_EINVAL, "resident rho: too many density matrices");
    void *ws = nullptr;
    int st = workspace_get(stream, sizeof(double) * h->max_tiles * nset_used * ngrid_used, &ws);
    if (st != CFGMM_OK) return st;
    double *partial = static_cast<double *>(ws);
    if ((size_t)ntasks * nset_used > INT32_MAX) return fail(CFGMM_EINVAL, "resident rho: too many workgroups");
    const int ng = (int)ngrid_used, ns = (int)nset_used;
    switch (h->nslices) {
        case 2: st = eval_launch<2>(stream, h, partial, ntasks, ng, ns); break;
        case 3: st = eval_launch<3>(stream, h, partial, ntasks, ng, ns); break;
        case 4: st = eval_launch<4>(stream, h, partial, ntasks, ng, ns); break;
        case 5: st = eval_launch<5>(stream, h, partial, ntasks, ng, ns); break;
        default: st = eval_launch<6>(stream, h, partial, ntasks, ng, ns); break;
    }
    if (st != CFGMM_OK) return st;
    return launch_rho_reduce(stream, rhos, partial, h->d_counts, stride_rho, ngrid_used, nset_used, h->blksize);
}

// Batched persistent evaluations: the items of every evaluation of a group, as self-contained
// records, dealt to two workgroups per CU (workgroup i on XCD i % 8): each (evaluation, grid block)
// -- its items share Rq and Lc -- to the XCD with the least work so far, its items round-robin over
// that XCD's workgroups.  The records depend on the handles, their leading ranges and the partial
// buffers (in the stream's workspace), not on rho: a plan is built on the host and uploaded once
// per distinct batch, and kept for the next calls with the same batch (the last few batches).
namespace {
struct BatchPlan {
    int device = -1;
    std::vector<unsigned long long> key;
    unsigned char *dev = nullptr;  // [slot_begin: nslots + 1 ints, padded to 256 bytes][ResItemX records]
    int nslots = 0;
};
std::mutex g_plan_mutex;
std::vector<BatchPlan> g_plans;  // least recently used first
constexpr size_t kPlanCache = 4;
}  // namespace

struct BatchEvalRef {
    const ResidentRho *h;
    double *partial;
    int ngrid, nset;
};

// The device plan of `evals` (records and slot ranges), from the cache or built and uploaded now.
// Returns with `lock` holding the cache: the caller enqueues the launch that reads the plan before
// releasing it, so that an eviction (which synchronizes the plan's device first) cannot free a plan
// whose launch is not yet enqueued.
static int batch_plan(hipStream_t stream, const std::vector<BatchEvalRef> &evals, const unsigned char **dev,
                      int *nslots, std::unique_lock<std::mutex> &lock) {
    int device = 0;
    CFGMM_HIP_TRY(hipGetDevice(&device));
    std::vector<unsigned long long> key;
    for (const BatchEvalRef &e : evals) {
        key.push_back(e.h->uid);
        key.push_back(reinterpret_cast<unsigned long long>(e.partial));
        key.push_back(((unsigned long long)(unsigned)e.ngrid << 32) | (unsigned)e.nset);
    }
    lock = std::unique_lock<std::mutex>(g_plan_mutex);
    for (size_t i = 0; i < g_plans.size(); ++i) {
        if (g_plans[i].device == device && g_plans[i].key == key) {
            std::rotate(g_plans.begin() + (long)i, g_plans.begin() + (long)i + 1, g_plans.end());
            *dev = g_plans.back().dev;
            *nslots = g_plans.back().nslots;
            return CFGMM_OK;
        }
    }
    const int per_xcd = std::max(1, 2 * device_cu_count() / 8), ns_all = 8 * per_xcd;
    std::vector<std::vector<ResItemX>> slots((size_t)ns_all);
    long xload[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int rr[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    size_t nitems = 0;
    for (const BatchEvalRef &e : evals) {
        const ResidentRho *h = e.h;
        const int nblk_used = (int)div_up((size_t)e.ngrid, h->blksize);
        const long hb = reinterpret_cast<long>(h->base);
        size_t i = 0;
        while (i < h->h_items.size()) {
            const int b = h->h_items[i].b;
            size_t j = i;
            while (j < h->h_items.size() && h->h_items[j].b == b) ++j;  // the items of block b
            if (b < nblk_used) {
                const int x = (int)(std::min_element(xload, xload + 8) - xload);
                for (size_t k = i; k < j; ++k) {
                    const ResItem &r = h->h_items[k];
                    if (r.s >= e.nset) continue;
                    ResItemX q{};
                    for (int t = 0; t < 2; ++t) {
                        q.a[t] = hb + r.a[t];
                        q.aw[t] = r.aw[t];
                        q.n[t] = r.n[t];
                        q.T[t] = r.T[t];
                    }
                    q.rq = hb + r.rq;
                    q.ed = hb + r.ed;
                    q.lc = hb + r.lc;
                    q.part = reinterpret_cast<long>(e.partial + (long)r.s * e.ngrid);
                    q.fg = reinterpret_cast<long>(h->d_fg);
                    q.dplane = r.dplane;
                    q.rplane = r.rplane;
                    q.c0 = r.c0;
                    q.b = r.b;
                    q.s = r.s;
                    q.nrows = r.nrows;
                    q.sym = r.sym;
                    q.tstride = e.nset * e.ngrid;
                    q.blksize = (int)h->blksize;
                    q.ngrid = e.ngrid;
                    slots[(size_t)x + 8 * (size_t)(rr[x]++ % per_xcd)].push_back(q);
                    xload[x] += r.n[0] + r.n[1];
                    ++nitems;
                }
            }
            i = j;
        }
    }
    if (nitems > (size_t)INT32_MAX) return fail(CFGMM_EINVAL, "resident rho: batch too large");
    const size_t head = div_up(sizeof(int) * (size_t)(ns_all + 1), 256) * 256;
    std::vector<unsigned char> host(head + sizeof(ResItemX) * std::max<size_t>(nitems, 1));
    int *sb = reinterpret_cast<int *>(host.data());
    ResItemX *rec = reinterpret_cast<ResItemX *>(host.data() + head);
    size_t pos = 0;
    sb[0] = 0;
    for (int i = 0; i < ns_all; ++i) {
        for (const ResItemX &q : slots[(size_t)i]) rec[pos++] = q;
        sb[i + 1] = (int)pos;
    }
    if (g_plans.size() == kPlanCache) {  // the least recently used plan goes
        // no launch on its device still reads it (every launch that read it was enqueued under the
        // cache lock, before this point)
        const int pdev = g_plans.front().device;
        if (pdev != device) (void)hipSetDevice(pdev);
        (void)hipDeviceSynchronize();
        (void)hipFree(g_plans.front().dev);
        if (pdev != device) (void)hipSetDevice(device);
        g_plans.erase(g_plans.begin());
    }
    BatchPlan pl;
    pl.device = device;
    pl.key = key;
    pl.nslots = ns_all;
    CFGMM_HIP_TRY(hipMalloc(&pl.dev, host.size()));
    // pageable source: the copy is complete when hipMemcpyAsync returns
    const int st = hip_status(hipMemcpyAsync(pl.dev, host.data(), host.size(), hipMemcpyHostToDevice, stream),
                              "resident rho: batch plan H2D");
    if (st != CFGMM_OK) {
        (void)hipFree(pl.dev);
        return st;
    }
    g_plans.push_back(std::move(pl));
    *dev = g_plans.back().dev;
    *nslots = ns_all;
    return CFGMM_OK;
}

// Evaluations of several handles in one launch per (slices, tile width) group; each exactly as
// resident_eval would compute it (the same tile code and the same fixed reduction order).
int resident_eval_batched(hipStream_t stream, const ResidentEvalReq *reqs, size_t n) {
    std::vector<size_t> todo;
    for (size_t i = 0; i < n; ++i) {
        const ResidentRho *h = reqs[i].h;
        if (reqs[i].ngrid == 0 || reqs[i].nset == 0 || h->nao == 0) continue;
        if (h->task_prefix[div_up(reqs[i].ngrid, h->blksize)] == 0) continue;
        todo.push_back(i);
    }
    while (!todo.empty()) {
        const int S = reqs[todo[0]].h->nslices, nt = reqs[todo[0]].h->nt;
        std::vector<size_t> group, rest;
        for (size_t i : todo) (reqs[i].h->nslices == S && reqs[i].h->nt == nt ? group : rest).push_back(i);
        todo.swap(rest);
        // workspace: [descriptors][partials of each evaluation]
        auto al = [](size_t x) { return div_up(x, 256) * 256; };
        std::vector<EvalDesc> d(group.size());
        size_t off = al(sizeof(EvalDesc) * group.size());
        long wg = 0;
        std::vector<size_t> part_off(group.size());
        for (size_t j = 0; j < group.size(); ++j) {
            const ResidentEvalReq &q = reqs[group[j]];
            part_off[j] = off;
            off += al(sizeof(double) * q.h->max_tiles * q.nset * q.ngrid);
        }
        void *ws = nullptr;
        int st = workspace_get(stream, off, &ws);
        if (st != CFGMM_OK) return st;
        char *base = static_cast<char *>(ws);
        size_t max_elems = 0;
        for (size_t j = 0; j < group.size(); ++j) {
            const ResidentEvalReq &q = reqs[group[j]];
            const ResidentRho *h = q.h;
            EvalDesc &e = d[j];
            e.base = h->base;
            e.blocks = h->d_blocks;
            e.tasks = h->d_tasks;
            e.fg = h->d_fg;
            e.counts = h->d_counts;
            e.partial = reinterpret_cast<double *>(base + part_off[j]);
            e.rhos = q.rhos;
            e.stride_rho = (long)q.stride_rho;
            e.blksize = (int)h->blksize;
            e.ngrid = (int)q.ngrid;
            e.nset = (int)q.nset;
            e.ntasks = h->task_prefix[div_up(q.ngrid, h->blksize)];
            e.wg_begin = (int)wg;
            wg += (long)e.ntasks * e.nset;
            max_elems = std::max(max_elems, q.nset * q.ngrid);
        }
        if (wg > INT32_MAX || group.size() > 65535) return fail(CFGMM_EINVAL, "resident rho: batch too large");
        // pageable source: the copy is complete when hipMemcpyAsync returns
        CFGMM_HIP_TRY(hipMemcpyAsync(base, d.data(), sizeof(EvalDesc) * d.size(), hipMemcpyHostToDevice, stream));
        const EvalDesc *dd = reinterpret_cast<const EvalDesc *>(base);
        const bool deep = tune_get(CFGMM_TUNE_RESIDENT_DEPTH) >= 2;
        // the persistent batch for up to five slices (with six its register sets spill)
        if (tune_get(CFGMM_TUNE_RESIDENT_PERSISTENT) == 0 && S <= 5) {
            std::vector<BatchEvalRef> refs(group.size());
            for (size_t j = 0; j < group.size(); ++j)
                refs[j] = BatchEvalRef{reqs[group[j]].h, d[j].partial, d[j].ngrid, d[j].nset};
            const unsigned char *plan = nullptr;
            int nslots = 0;
            std::unique_lock<std::mutex> plan_lock;  // held until the launch that reads the plan is enqueued
            st = batch_plan(stream, refs, &plan, &nslots, plan_lock);
            if (st != CFGMM_OK) return st;
            const size_t head = div_up(sizeof(int) * (size_t)(nslots + 1), 256) * 256;
            const void *recs = plan + head;
            const int *sb = reinterpret_cast<const int *>(plan);
            hipEvent_t t0 = trace_start(stream);
#define CFGMM_RESBP(S_, NT_, D_)                                                                              \
    hipLaunchKernelGGL((rho_resident_persistent_kernel<S_, NT_, D_, true>), dim3((unsigned)nslots),            \
                       dim3(kThreads), 0, stream, nullptr, nullptr, recs, sb, nullptr, 0, 0, 0, 0)
#define CFGMM_RESBP_NT(S_, D_)               \
    switch (nt) {                            \
        case 1: CFGMM_RESBP(S_, 1, D_); break; \
        case 2: CFGMM_RESBP(S_, 2, D_); break; \
        case 3: CFGMM_RESBP(S_, 3, D_); break; \
        default: CFGMM_RESBP(S_, 4, D_); break; \
    }
            switch (S) {  // one k step of prefetch: with the batch records a second register set spills
                case 2: CFGMM_RESBP_NT(2, 1) break;
                case 3: CFGMM_RESBP_NT(3, 1) break;
                case 4: CFGMM_RESBP_NT(4, 1) break;
                default: CFGMM_RESBP_NT(5, 1) break;
            }
#undef CFGMM_RESBP_NT
#undef CFGMM_RESBP
            trace_stop(stream, "rho_resident_kernel", t0);
            plan_lock.unlock();
            st = hip_status(hipGetLastError(), "rho_resident_persistent_kernel (batch) launch");
            if (st != CFGMM_OK) return st;
            hipLaunchKernelGGL(resident_batched_reduce, dim3((unsigned)div_up(max_elems, 256), (unsigned)d.size()),
                               dim3(256), 0, stream, dd);
            st = hip_status(hipGetLastError(), "resident_batched_reduce launch");
            if (st != CFGMM_OK) return st;
            continue;
        }
        hipEvent_t t0 = trace_start(stream);
#define CFGMM_RESB(S_, NT_)                                                                                     \
    do {                                                                                                         \
        if (deep)                                                                                                \
            hipLaunchKernelGGL((rho_resident_batched_kernel<S_, NT_, 2>), dim3((unsigned)wg), dim3(kThreads), 0, \
                               stream, dd, (int)d.size());                                                       \
        else                                                                                                     \
            hipLaunchKernelGGL((rho_resident_batched_kernel<S_, NT_, 1>), dim3((unsigned)wg), dim3(kThreads), 0, \
                               stream, dd, (int)d.size());                                                       \
    } while (0)
#define CFGMM_RESB_NT(S_)               \
    switch (nt) {                       \
        case 1: CFGMM_RESB(S_, 1); break; \
        case 2: CFGMM_RESB(S_, 2); break; \
        case 3: CFGMM_RESB(S_, 3); break; \
        default: CFGMM_RESB(S_, 4); break; \
    }
        switch (S) {
            case 2: CFGMM_RESB_NT(2); break;
            case 3: CFGMM_RESB_NT(3); break;
            case 4: CFGMM_RESB_NT(4); break;
            case 5: CFGMM_RESB_NT(5); break;
            default: CFGMM_RESB_NT(6); break;
        }
#undef CFGMM_RESB_NT
#undef CFGMM_RESB
        trace_stop(stream, "rho_resident_kernel", t0);
        st = hip_status(hipGetLastError(), "rho_resident_batched_kernel launch");
        if (st != CFGMM_OK) return st;
        hipLaunchKernelGGL(resident_batched_reduce, dim3((unsigned)div_up(max_elems, 256), (unsigned)d.size()),
                           dim3(256), 0, stream, dd);
        st = hip_status(hipGetLastError(), "resident_batched_reduce launch");
        if (st != CFGMM_OK) return st;
    }
    return CFGMM_OK;
}

}  // namespace cfgmm
