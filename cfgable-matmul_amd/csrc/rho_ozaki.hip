// rho_ozaki.hip -- split-precision (Ozaki-scheme) masked AO-DM-AO rho on the int8 MFMA (gfx950).
//
// Same contract as launch_rho (src/impl_aodmao2rho.rs:171-244):
//   rho[s][g] += sum_{u,v in L_b} lhs[u][g] D_s[u][v] rhs[v][g],  b = g / blksize.
// T_s = D_s[L_b, L_b] rhs[L_b, b] is computed from int8 slices:
//   * every row u of D_s is scaled by 2^-ed[s][u] (row maximum over all v) and
//     cut into S 7-bit slices; the slices of one entry are stored together in
//     one 8-byte word (byte j = slice j) so a gathered entry is one load;
//   * every grid point g is scaled by 2^-fg[g], the maximum of |rhs[v][g]| over
//     the active AOs v of its block only (entries under off mask bits are never
//     read, also not for scaling), and rhs is stored block-compacted:
//     rq[b][i][c] = slices of rhs[L_b[i]][b*blksize + c], with zero rows up to
//     the next multiple of 64 so that the main kernel needs no bounds checks;
//   * slice products are accumulated exactly in int32 per level s+t <= S+1
//     (v_mfma_i32_16x16x64_i8), recombined in f64 with the power-of-two
//     scales, multiplied by lhs[u][g] in f64 and reduced over u in a fixed
//     order (the same chunk partials and reduction as the FP64 kernel).
// A row of D_s or a block column of rhs holding Inf/NaN makes the rho values
// it touches NaN.
#include "internal.hpp"

#include <algorithm>

namespace cfgmm {
namespace {

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef unsigned long long u64;

constexpr int kThreads = 256;
constexpr int kUC = 64;    // u rows per workgroup (4 waves x 16)
constexpr int kKV = 64;    // compacted v per slice (int8 MFMA K)
constexpr int kSeg = 2048; // compacted-list entries staged in LDS at a time
constexpr int kExpNonFinite = 0x7fffffff;

__device__ inline int exponent_of(u64 maxbits) {
    const double m = __longlong_as_double((long long)maxbits);
    if (!(m <= 1.79769313486231570815e308)) return kExpNonFinite;
    if (m == 0.0) return 0;
    int e;
    frexp(m, &e);
    return e;
}

// S 7-bit slices of x * 2^-e (|x| < 2^e) packed into one word, byte j = slice j.
template <int S>
__device__ inline u64 pack_slices(double x, int e) {
    double rest = e == kExpNonFinite ? 0.0 : ldexp(x, -e);
    u64 w = 0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const double v = rest * 128.0;
        const double q = trunc(v);
        rest = v - q;
        w |= (u64)((unsigned)(int)q & 0xffu) << (8 * s);
    }
    return w;
}

__device__ inline u64 block_max_u64(u64 v, u64 *red) {
    for (int off = 32; off > 0; off >>= 1) {
        const u64 o = __shfl_xor(v, off);
        v = o > v ? o : v;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    u64 m = red[0];
    for (int w = 1; w < kThreads / 64; ++w) m = red[w] > m ? red[w] : m;
    __syncthreads();
    return m;
}

// One workgroup per (row u, set s): ed[s][u] and dq[s][u][v] = slices of D_s[u][v].
template <int S>
__global__ __launch_bounds__(kThreads) void dm_slice_kernel(const double *__restrict__ dms, long stride_ao, int nao,
                                                            int *__restrict__ ed, u64 *__restrict__ dq) {
    __shared__ u64 red[kThreads / 64];
    const int u = blockIdx.x, s = blockIdx.y;
    const double *__restrict__ row = dms + (long)s * stride_ao * stride_ao + (long)u * stride_ao;
    u64 m = 0;
    for (int v = threadIdx.x; v < nao; v += kThreads) {
        const u64 bits = (u64)__double_as_longlong(fabs(row[v]));
        m = bits > m ? bits : m;
    }
    const int e = exponent_of(block_max_u64(m, red));
    if (threadIdx.x == 0) ed[(long)s * nao + u] = e;
    u64 *__restrict__ out = dq + ((long)s * nao + u) * nao;
    for (int v = threadIdx.x; v < nao; v += kThreads) out[v] = pack_slices<S>(row[v], e);
}

// One workgroup per grid block b: fg[g] for the block's points and the
// block-compacted slices rq[b][i][c], i < round_up(|L_b|, 64), c < blksize.
template <int S>
__global__ __launch_bounds__(kThreads) void phi_slice_kernel(const double *__restrict__ rhs, long stride_grid,
                                                             int ngrid, int nao, int blksize, long rows_per_blk,
                                                             const int *__restrict__ lists,
                                                             const int *__restrict__ counts, int *__restrict__ fg,
                                                             u64 *__restrict__ rq) {
    extern __shared__ u64 colmax[];  // [blksize]
    const int b = blockIdx.x;
    const int cnt = counts[b];
    const int g0 = b * blksize;
    const int ncol = min(blksize, ngrid - g0);
    const int *__restrict__ list = lists + (long)b * nao;
    for (int c = threadIdx.x; c < blksize; c += kThreads) colmax[c] = 0;
    __syncthreads();
    for (int idx = threadIdx.x; idx < cnt * ncol; idx += kThreads) {
        const int i = idx / ncol, c = idx % ncol;
        const u64 bits = (u64)__double_as_longlong(fabs(rhs[(long)list[i] * stride_grid + g0 + c]));
        atomicMax(&colmax[c], bits);
    }
    __syncthreads();
    for (int c = threadIdx.x; c < ncol; c += kThreads) fg[g0 + c] = exponent_of(colmax[c]);
    const int rows = (cnt + kKV - 1) / kKV * kKV;
    u64 *__restrict__ out = rq + (long)b * rows_per_blk * blksize;
    for (int idx = threadIdx.x; idx < rows * blksize; idx += kThreads) {
        const int i = idx / blksize, c = idx % blksize;
        u64 w = 0;
        if (i < cnt && c < ncol) w = pack_slices<S>(rhs[(long)list[i] * stride_grid + g0 + c], exponent_of(colmax[c]));
        out[idx] = w;
    }
}

// byte offset of 16-byte chunk `ch` of a 64-byte LDS row (XOR swizzle)
__device__ inline int swz(int row, int ch) { return row * kKV + ((ch ^ ((row >> 2) & 3)) << 4); }

// byte j of each of four words, packed into one dword
__device__ inline unsigned gather_bytes(u64 w0, u64 w1, u64 w2, u64 w3, int j) {
    const int sh = 8 * j;
    return (unsigned)((w0 >> sh) & 0xff) | ((unsigned)((w1 >> sh) & 0xff) << 8) |
           ((unsigned)((w2 >> sh) & 0xff) << 16) | ((unsigned)((w3 >> sh) & 0xff) << 24);
}

template <int S, int NT>
__global__ __launch_bounds__(kThreads, 1) void rho_ozaki_kernel(
    double *__restrict__ partial, const u64 *__restrict__ dq, const int *__restrict__ ed,
    const u64 *__restrict__ rq, const int *__restrict__ fg, const double *__restrict__ ao_lhs, long stride_grid,
    int nao, int ngrid, int nset, const int *__restrict__ lists, const int *__restrict__ counts, int blksize, int nsub,
    long rows_per_blk) {
    constexpr int W = NT * 16;
    const int blk = blockIdx.x / nsub, sub = blockIdx.x % nsub;
    const int chunk = blockIdx.y, s = blockIdx.z;
    const int cnt = counts[blk];
    const int u_begin = chunk * kUC;
    if (u_begin >= cnt) return;
    const int c0 = sub * W;
    const int g0 = blk * blksize + c0;
    if (g0 >= ngrid) return;
    const int *__restrict__ list = lists + (long)blk * nao;
    const int nrows = min(kUC, cnt - u_begin);

    __shared__ __attribute__((aligned(16))) int8_t As[S][kUC * kKV];  // D slices [row][k]
    __shared__ __attribute__((aligned(16))) int8_t Bs[S][W * kKV];    // rhs slices [col][k]
    __shared__ int vls[kSeg];
    __shared__ int ulist[kUC];
    __shared__ double red[kThreads / 64][W];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int lrow = lane & 15, lk = lane >> 4;
    if (tid < kUC) ulist[tid] = list[u_begin + (tid < nrows ? tid : 0)];
    __syncthreads();

    // D loads: thread -> row tid/4, 16 consecutive compacted v (chunk tid%4)
    const int d_row = tid >> 2, d_ch = tid & 3;
    const u64 *__restrict__ drow = dq + ((long)s * nao + ulist[d_row]) * nao;
    // rhs loads: thread -> column tid % W, 16 consecutive compacted v (chunk tid / W); threads >= 4W idle
    const bool r_on = tid < 4 * W;
    const int r_col = r_on ? tid % W : 0, r_ch = r_on ? tid / W : 0;
    const u64 *__restrict__ rbase = rq + (long)blk * rows_per_blk * blksize + c0 + r_col;

    i32x4 acc[S][NT];
#pragma unroll
    for (int l = 0; l < S; ++l)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[l][nt] = i32x4{0, 0, 0, 0};

    u64 pd[16], pr[16];
    for (int seg0 = 0; seg0 < cnt; seg0 += kSeg) {
        const int segn = min(kSeg, cnt - seg0);
        __syncthreads();
        for (int j = tid; j < segn; j += kThreads) vls[j] = list[seg0 + j];
        __syncthreads();
        // Entries past the list end read a valid D entry (index clamped) and a
        // zero rhs row, so their products are exactly zero.
        auto load_slice = [&](int k0) {
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int p = min(k0 + d_ch * 16 + j, segn - 1);
                pd[j] = drow[vls[p]];
            }
            if (r_on) {
#pragma unroll
                for (int j = 0; j < 16; ++j) pr[j] = rbase[(long)(seg0 + k0 + r_ch * 16 + j) * blksize];
            }
        };
        auto store_slice = [&]() {
#pragma unroll
            for (int sl = 0; sl < S; ++sl) {
                i32x4 v;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    v[q] = (int)gather_bytes(pd[4 * q], pd[4 * q + 1], pd[4 * q + 2], pd[4 * q + 3], sl);
                *reinterpret_cast<i32x4 *>(&As[sl][swz(d_row, d_ch)]) = v;
            }
            if (r_on) {
#pragma unroll
                for (int sl = 0; sl < S; ++sl) {
                    i32x4 v;
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        v[q] = (int)gather_bytes(pr[4 * q], pr[4 * q + 1], pr[4 * q + 2], pr[4 * q + 3], sl);
                    *reinterpret_cast<i32x4 *>(&Bs[sl][swz(r_col, r_ch)]) = v;
                }
            }
        };

        const int nslices = (segn + kKV - 1) / kKV;
        load_slice(0);
        for (int t = 0; t < nslices; ++t) {
            __syncthreads();  // previous slice's MFMA reads of As/Bs are done
            store_slice();
            __syncthreads();
            if (t + 1 < nslices) load_slice((t + 1) * kKV);
            i32x4 fb[S][NT];
#pragma unroll
            for (int sb = 0; sb < S; ++sb)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                    fb[sb][nt] = *reinterpret_cast<const i32x4 *>(&Bs[sb][swz(nt * 16 + lrow, lk)]);
#pragma unroll
            for (int sa = 0; sa < S; ++sa) {
                const i32x4 fa = *reinterpret_cast<const i32x4 *>(&As[sa][swz(wave * 16 + lrow, lk)]);
#pragma unroll
                for (int sb = 0; sb + sa < S; ++sb)
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt)
                        acc[sa + sb][nt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa, fb[sb][nt], acc[sa + sb][nt], 0, 0, 0);
            }
        }
    }

    // Epilogue. i32 16x16 layout: lane l, element r holds row 4*(l/16)+r, column l%16.
    const double nan = __builtin_nan("");
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        const int g = g0 + nt * 16 + lrow;
        const bool g_ok = g < ngrid;
        const int eg = g_ok ? fg[g] : 0;
        double part = 0.0;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = wave * 16 + 4 * lk + r;
            if (i < nrows && g_ok) {
                const int u = ulist[i];
                const int eu = ed[(long)s * nao + u];
                double t = 0.0;
#pragma unroll
                for (int l = S - 1; l >= 0; --l) t += ldexp((double)acc[l][nt][r], -7 * (l + 2));
                t = (eu == kExpNonFinite || eg == kExpNonFinite) ? nan : ldexp(t, eu + eg);
                part = fma(ao_lhs[(long)u * stride_grid + g], t, part);
            }
        }
        part += __shfl_xor(part, 16);
        part += __shfl_xor(part, 32);
        if (lk == 0) red[wave][nt * 16 + lrow] = part;
    }
    __syncthreads();
    for (int c = tid; c < W; c += kThreads) {
        const int g = g0 + c;
        if (g < ngrid) {
            double v = red[0][c];
#pragma unroll
            for (int w = 1; w < kThreads / 64; ++w) v += red[w][c];
            partial[((long)chunk * nset + s) * ngrid + g] = v;
        }
    }
}

template <int S>
int run_rho_ozaki(hipStream_t stream, double *rhos, const double *dms, const double *ao_lhs, const double *ao_rhs,
                  size_t stride_ao, size_t stride_grid, size_t nao, size_t ngrid, size_t nset, const uint8_t *tab,
                  size_t ldtab, size_t blksize) {
    const size_t nblk = div_up(ngrid, blksize);
    const size_t maxchunks = div_up(nao, kUC);
    const size_t rows_per_blk = div_up(nao, kKV) * kKV;
    int nt = 1;
    for (int cand = 4; cand >= 1; --cand)
        if ((blksize / 16) % cand == 0) { nt = cand; break; }
    const int nsub = (int)(blksize / (16 * nt));
    if (nblk * nsub > (size_t)INT32_MAX || maxchunks > 65535 || nset > 65535 || blksize > 4096)
        return fail(CFGMM_EINVAL, "split-precision rho: problem too large for one launch");

    // workspace: lists, counts | partial | ed | fg | dq | rq
    auto al = [](size_t x) { return div_up(x, 256) * 256; };
    const size_t o_counts = al(sizeof(int) * nblk * nao);
    const size_t o_part = o_counts + al(sizeof(int) * nblk);
    const size_t o_ed = o_part + al(sizeof(double) * maxchunks * nset * ngrid);
    const size_t o_fg = o_ed + al(sizeof(int) * nset * nao);
    const size_t o_dq = o_fg + al(sizeof(int) * nblk * blksize);
    const size_t o_rq = o_dq + al(sizeof(u64) * nset * nao * nao);
    const size_t total = o_rq + sizeof(u64) * nblk * rows_per_blk * blksize;
    void *ws = nullptr;
    int st = workspace_get(stream, total, &ws);
    if (st != CFGMM_OK) return st;
    char *base = static_cast<char *>(ws);
    int *lists = reinterpret_cast<int *>(base);
    int *counts = reinterpret_cast<int *>(base + o_counts);
    double *partial = reinterpret_cast<double *>(base + o_part);
    int *ed = reinterpret_cast<int *>(base + o_ed);
    int *fg = reinterpret_cast<int *>(base + o_fg);
    u64 *dq = reinterpret_cast<u64 *>(base + o_dq);
    u64 *rq = reinterpret_cast<u64 *>(base + o_rq);

    st = launch_build_lists(stream, tab, ldtab, nao, nblk, lists, counts);
    if (st != CFGMM_OK) return st;
    hipEvent_t t0 = trace_start(stream);
    hipLaunchKernelGGL((dm_slice_kernel<S>), dim3((unsigned)nao, (unsigned)nset), dim3(kThreads), 0, stream, dms,
                       (long)stride_ao, (int)nao, ed, dq);
    hipLaunchKernelGGL((phi_slice_kernel<S>), dim3((unsigned)nblk), dim3(kThreads), sizeof(u64) * blksize, stream,
                       ao_rhs, (long)stride_grid, (int)ngrid, (int)nao, (int)blksize, (long)rows_per_blk, lists,
                       counts, fg, rq);
    trace_stop(stream, "rho_ozaki_split", t0);
    st = hip_status(hipGetLastError(), "rho_ozaki split launch");
    if (st != CFGMM_OK) return st;

    dim3 grid((unsigned)(nblk * nsub), (unsigned)maxchunks, (unsigned)nset);
    t0 = trace_start(stream);
#define CFGMM_RHO_OZ(NT_)                                                                                           \
    hipLaunchKernelGGL((rho_ozaki_kernel<S, NT_>), grid, dim3(kThreads), 0, stream, partial, dq, ed, rq, fg, ao_lhs, \
                       (long)stride_grid, (int)nao, (int)ngrid, (int)nset, lists, counts, (int)blksize, nsub,          \
                       (long)rows_per_blk)
    switch (nt) {
        case 1: CFGMM_RHO_OZ(1); break;
        case 2: CFGMM_RHO_OZ(2); break;
        case 3: CFGMM_RHO_OZ(3); break;
        default: CFGMM_RHO_OZ(4); break;
    }
#undef CFGMM_RHO_OZ
    trace_stop(stream, "rho_ozaki_kernel", t0);
    st = hip_status(hipGetLastError(), "rho_ozaki_kernel launch");
    if (st != CFGMM_OK) return st;
    return launch_rho_reduce(stream, rhos, partial, counts, stride_grid, ngrid, nset, blksize);
}

}  // namespace

int launch_rho_ozaki(hipStream_t stream, double *rhos, const double *dms, const double *ao_lhs, const double *ao_rhs,
                     size_t stride_ao, size_t stride_grid, size_t nao, size_t ngrid, size_t nset, const uint8_t *tab,
                     size_t ldtab, size_t blksize, int nslices) {
    if (nao == 0 || ngrid == 0 || nset == 0) return CFGMM_OK;
    if (nao > INT32_MAX || ngrid > INT32_MAX) return fail(CFGMM_EINVAL, "split-precision rho: size out of range");
    switch (nslices) {
        case 2: return run_rho_ozaki<2>(stream, rhos, dms, ao_lhs, ao_rhs, stride_ao, stride_grid, nao, ngrid, nset, tab, ldtab, blksize);
        case 3: return run_rho_ozaki<3>(stream, rhos, dms, ao_lhs, ao_rhs, stride_ao, stride_grid, nao, ngrid, nset, tab, ldtab, blksize);
        case 4: return run_rho_ozaki<4>(stream, rhos, dms, ao_lhs, ao_rhs, stride_ao, stride_grid, nao, ngrid, nset, tab, ldtab, blksize);
        case 5: return run_rho_ozaki<5>(stream, rhos, dms, ao_lhs, ao_rhs, stride_ao, stride_grid, nao, ngrid, nset, tab, ldtab, blksize);
        case 6: return run_rho_ozaki<6>(stream, rhos, dms, ao_lhs, ao_rhs, stride_ao, stride_grid, nao, ngrid, nset, tab, ldtab, blksize);
        default: return fail(CFGMM_EINVAL, "nslices must be in [2, 6]");
    }
}

}  // namespace cfgmm
