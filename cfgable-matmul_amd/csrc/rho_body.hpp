// rho_body.hpp -- the per-chunk body of the fused FP64 masked rho contraction,
// shared by the per-problem kernel (rho.hip) and the persistent batched kernel
// (rho_batched.hip).  See rho.hip for the algorithm.
#pragma once

#include "internal.hpp"

namespace cfgmm {
namespace rho {

constexpr int kThreads = 256;
constexpr int kUC = 64;     // rows of L_b per workgroup (4 waves x 16)
constexpr int kKV = 16;     // v slice depth per LDS stage
constexpr int kSeg = 1536;  // compacted-list entries staged in LDS at a time (keeps <1,4> under 40 KB: 4 WGs per CU)

// One unit of work: rows [chunk*64, chunk*64+64) of the compacted AO list of
// one grid block, grid points [g0, g0 + 16*NT), density matrices [set0, set0+NS).
struct RhoChunk {
    double *partial;  // [chunk][nset][ngrid]
    const double *dms, *ao_lhs, *ao_rhs;
    long stride_ao, stride_grid;
    int ngrid, nset, set0;
    const int *list;  // compacted AO list of the block
    int cnt;          // its length
    int chunk, g0;
    // several left-hand factors sharing D and rhs (e.g. phi and its gradient): component c
    // reads ao_lhs + c*lhs_comp_stride and writes partial + c*part_comp_stride
    int ncomp = 1;
    long lhs_comp_stride = 0, part_comp_stride = 0;
};
// rho_chunk_body<NS, NT, true>: D_s symmetric and lhs == rhs (one component): the chunk contracts
// only list positions from its own first row on, the positions past its 64 rows twice (the pairs
// of those positions with earlier rows are the mirror images of pairs that the earlier chunks count).

template <int NS, int NT>
constexpr size_t chunk_lds_bytes() {
    return sizeof(double) * (2 * NS * kKV * (kUC + 1) + 2 * kKV * NT * 16) + sizeof(int) * (kSeg + kUC);
}

// OFF32: every byte offset into ao_rhs (row x stride_grid + point, times 8) is below 2^32, so the
// slice loads of phi_rhs take 32-bit offsets from its base (fewer VALU instructions in the slice
// loop than 64-bit row addresses); the values loaded are the same.
template <int NS, int NT, bool SYM = false, bool OFF32 = false>
__device__ inline void rho_chunk_body(const RhoChunk &task, unsigned char *__restrict__ smem) {
    constexpr int W = NT * 16;                          // grid points per workgroup
    constexpr int PHI_PER_T = kKV * W / kThreads;       // phi_rhs elements per thread per slice
    constexpr int D_ROWS_PER_T = kUC * kKV / kThreads;  // D rows per thread per slice and set (4)
    static_assert(PHI_PER_T * kThreads == kKV * W, "phi tile split");
    static_assert(D_ROWS_PER_T * kThreads == kUC * kKV, "D tile split");

    double *__restrict__ partial = task.partial;
    const double *__restrict__ dms = task.dms;
    const double *__restrict__ ao_lhs = task.ao_lhs;
    const double *__restrict__ ao_rhs = task.ao_rhs;
    const long stride_ao = task.stride_ao, stride_grid = task.stride_grid;
    const int ngrid = task.ngrid, nset = task.nset, set0 = task.set0;
    const int *__restrict__ list = task.list;
    const int cnt = task.cnt, chunk = task.chunk, g0 = task.g0;
    const int u_begin = chunk * kUC;

    // LDS (carved from smem): double-buffered A (D_s[u][v] stored [v][u]) and B
    // (phi_rhs[v][g]) slices, a segment of the compacted v list, this chunk's
    // u rows.  The cross-wave reduction buffer of the epilogue reuses the A slices.
    constexpr int AS_ELEMS = 2 * NS * kKV * (kUC + 1);
    static_assert(AS_ELEMS >= (kThreads / 64) * NS * W, "epilogue buffer fits in the A slices");
    double *__restrict__ As_raw = reinterpret_cast<double *>(smem);
    double (*Bs)[kKV][W] = reinterpret_cast<double (*)[kKV][W]>(smem + sizeof(double) * AS_ELEMS);
    int *__restrict__ vls = reinterpret_cast<int *>(smem + sizeof(double) * (AS_ELEMS + 2 * kKV * W));
    int *__restrict__ ulist = vls + kSeg;
    auto As = [&](int buf, int s, int p, int i) -> double & {
        return As_raw[((buf * NS + s) * kKV + p) * (kUC + 1) + i];
    };

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int lrow = lane & 15, lk = lane >> 4;
    const int nrows = min(kUC, cnt - u_begin);
    // rows past the end of the list read a valid row (the chunk's first) and are zeroed
    if (tid < kUC) ulist[tid] = list[u_begin + (tid < nrows ? tid : 0)];

    const long dstride = stride_ao * stride_ao;
    // Fixed per-thread D coordinates: column slot p_d, rows tid/16 + 16e.
    const int p_d = tid % kKV;

    // Register sets of slice operands: a slice's loads are issued PF slices before it is
    // stored into LDS.  A loaded global access takes a few microseconds, several slices' worth
    // of MFMAs, but every set costs VGPRs: on the config-2 workload one set at four waves per
    // SIMD (1.05 ms) beats two (1.07 ms) and three (1.11 ms) at three waves per SIMD, the
    // waves of the other workgroups covering the latency.
    constexpr int PF = 1;
    double rphi[PF][PHI_PER_T], rd[PF][NS][D_ROWS_PER_T];
    f64x4 acc[NS][NT];
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[s][nt] = f64x4{0.0, 0.0, 0.0, 0.0};

    // Symmetric form: positions past the chunk's rows first, the accumulators doubled, then the
    // chunk's own 64 positions.  General form: the whole list.
    const int own_end = min(u_begin + kUC, cnt);
    constexpr int nranges = SYM ? 2 : 1;
    for (int rg = 0; rg < nranges; ++rg) {
        const int kb = SYM ? (rg == 0 ? own_end : u_begin) : 0;
        const int ke = SYM && rg == 1 ? own_end : cnt;
        for (int seg0 = kb; seg0 < ke; seg0 += kSeg) {
            const int segn = min(kSeg, ke - seg0);
            __syncthreads();  // previous segment's readers of vls / slices are done
            for (int j = tid; j < segn; j += kThreads) vls[j] = list[seg0 + j];
            __syncthreads();
            const double *__restrict__ drow[NS][D_ROWS_PER_T];
#pragma unroll
            for (int e = 0; e < D_ROWS_PER_T; ++e) {
                const int i = tid / kKV + (kThreads / kKV) * e;
#pragma unroll
                for (int s = 0; s < NS; ++s) drow[s][e] = dms + (long)(set0 + s) * dstride + (long)ulist[i] * stride_ao;
            }

            // Loads of a full 16-deep slice are unconditional: every list entry is
            // valid, rows past the chunk end read a valid fallback row (their
            // products are dropped in the epilogue) and columns past the grid end
            // are clamped (their outputs are never written).  The last, partial
            // slice of a segment is loaded by load_tail, after the main loop, so
            // that no select on loaded values sits in the pipelined path.
            // (set = register set; a compile-time index wherever the loop below is unrolled)
            auto load_full = [&](int k0, int set) {
                const long vcol = vls[k0 + p_d];
#pragma unroll
                for (int e = 0; e < D_ROWS_PER_T; ++e)
#pragma unroll
                    for (int s = 0; s < NS; ++s) rd[set][s][e] = drow[s][e][vcol];
#pragma unroll
                for (int e = 0; e < PHI_PER_T; ++e) {
                    const int idx = tid + kThreads * e;
                    const int g = min(g0 + idx % W, ngrid - 1);
                    if constexpr (OFF32)
                        rphi[set][e] = *reinterpret_cast<const double *>(
                            reinterpret_cast<const char *>(ao_rhs) +
                            ((unsigned)vls[k0 + idx / W] * (unsigned)(stride_grid * 8) + (unsigned)g * 8u));
                    else
                        rphi[set][e] = ao_rhs[(long)vls[k0 + idx / W] * stride_grid + g];
                }
            };
            auto load_tail = [&](int k0) {
                const int pv = k0 + p_d;
                const long vcol = vls[pv < segn ? pv : 0];
#pragma unroll
                for (int e = 0; e < D_ROWS_PER_T; ++e)
#pragma unroll
                    for (int s = 0; s < NS; ++s) rd[0][s][e] = pv < segn ? drow[s][e][vcol] : 0.0;
#pragma unroll
                for (int e = 0; e < PHI_PER_T; ++e) {
                    const int idx = tid + kThreads * e;
                    const int pr = k0 + idx / W;
                    const int g = min(g0 + idx % W, ngrid - 1);
                    rphi[0][e] = pr < segn ? ao_rhs[(long)vls[pr] * stride_grid + g] : 0.0;
                }
            };
            auto store_slice = [&](int buf, int set) {
#pragma unroll
                for (int e = 0; e < PHI_PER_T; ++e) {
                    const int idx = tid + kThreads * e;
                    Bs[buf][idx / W][idx % W] = rphi[set][e];
                }
#pragma unroll
                for (int e = 0; e < D_ROWS_PER_T; ++e)
#pragma unroll
                    for (int s = 0; s < NS; ++s) As(buf, s, p_d, tid / kKV + (kThreads / kKV) * e) = rd[set][s][e];
            };
            auto compute_slice = [&](int cur) {
#pragma unroll
                for (int kk = 0; kk < kKV / 4; ++kk) {
                    double fa[NS], fb[NT];
#pragma unroll
                    for (int s = 0; s < NS; ++s) fa[s] = As(cur, s, kk * 4 + lk, wave * 16 + lrow);
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt) fb[nt] = Bs[cur][kk * 4 + lk][nt * 16 + lrow];
#pragma unroll
                    for (int s = 0; s < NS; ++s)
#pragma unroll
                        for (int nt = 0; nt < NT; ++nt)
                            acc[s][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[s], fb[nt], acc[s][nt], 0, 0, 0);
                }
            };

            // Slice t lives in register set t % PF, then in LDS buffer t % 2.  Iteration t refills
            // set t % PF (slice t is already in LDS) with slice t+PF, computes slice t and stores
            // slice t+1 (loaded PF iterations earlier) into the other buffer.  In the main part the
            // refills are unconditional (past the end they reload the last slice, unused) so that
            // the compiler can count the loads in flight and wait only for the slice it stores;
            // the last nfull % PF slices were loaded by then.
            const int nfull = segn / kKV;
            if (nfull > 0) {
#pragma unroll
                for (int u = 0; u < PF; ++u) load_full(min(u, nfull - 1) * kKV, u);
                store_slice(0, 0);
            }
            __syncthreads();
            // main part: whole rounds of PF slices, straight-line (a store past the end writes
            // the reloaded last slice into the idle buffer, never computed)
            int t0 = 0;
            for (; t0 + PF <= nfull; t0 += PF) {
#pragma unroll
                for (int u = 0; u < PF; ++u) {
                    const int t = t0 + u;
                    load_full(min(t + PF, nfull - 1) * kKV, u);
                    compute_slice(t & 1);
                    store_slice((t + 1) & 1, (u + 1) % PF);
                    __syncthreads();
                }
            }
#pragma unroll
            for (int u = 0; u < PF; ++u) {  // the last nfull % PF slices
                const int t = t0 + u;
                if (t < nfull) {
                    compute_slice(t & 1);
                    if (t + 1 < nfull) store_slice((t + 1) & 1, (u + 1) % PF);
                    __syncthreads();
                }
            }
            if (nfull * kKV < segn) {
                load_tail(nfull * kKV);
                store_slice(0, 0);
                __syncthreads();
                compute_slice(0);
                __syncthreads();
            }
        }
        if (SYM && rg == 0) {
#pragma unroll
            for (int s = 0; s < NS; ++s)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) acc[s][nt] *= 2.0;
        }
    }

    // Epilogue: rho_partial[g] = sum_u lhs[u][g] * T[u][g] over this chunk's rows, for each
    // left-hand component (T is shared).
    double *red = As_raw;  // [4 waves][NS][W]; all slice reads finished at the last barrier
    for (int comp = 0; comp < task.ncomp; ++comp) {
        const double *__restrict__ lhs_c = ao_lhs + comp * task.lhs_comp_stride;
        double *__restrict__ part_c = partial + comp * task.part_comp_stride;
        if (comp > 0) __syncthreads();  // the previous component's reads of red are done
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int g = g0 + nt * 16 + lrow;
            const bool g_ok = g < ngrid;
            double lhs[4];
            bool row_ok[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = wave * 16 + lk + 4 * r;  // D[l/16 + 4r][l%16]
                row_ok[r] = i < nrows;
                lhs[r] = lhs_c[(long)ulist[i] * stride_grid + (g_ok ? g : ngrid - 1)];
            }
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                // rows past the chunk end (fallback rows) contribute nothing
                double part = 0.0;
#pragma unroll
                for (int r = 0; r < 4; ++r) part = row_ok[r] ? fma(acc[s][nt][r], lhs[r], part) : part;
                part += __shfl_xor(part, 16);
                part += __shfl_xor(part, 32);
                if (lk == 0) red[(wave * NS + s) * W + nt * 16 + lrow] = part;
            }
        }
        __syncthreads();
        for (int idx = tid; idx < NS * W; idx += kThreads) {
            const int s = idx / W, col = idx % W;
            const int g = g0 + col;
            if (g < ngrid) {
                double v = red[s * W + col];
#pragma unroll
                for (int w = 1; w < kThreads / 64; ++w) v += red[(w * NS + s) * W + col];
                part_c[((long)chunk * nset + set0 + s) * ngrid + g] = v;
            }
        }
    }
}


// Whether the 64x64 tiles (ti, tj) and (tj, ti) of the nao x nao matrix d (row stride ld) are
// not each other's transpose; the same answer on every thread of the workgroup (NT threads, a
// multiple of 64 dividing 4096, all of which must call it).  t: LDS scratch of 64 x 65 words.
// Two entries d[i][j], d[j][i] count as equal when their bits are, or when they differ by at most
// 2^-50 of the larger magnitude (4 units in the last place): a D formed as C C^T by a general
// product is symmetric only to rounding.  The symmetric form then contracts d[min][max] for both,
// which changes each term by at most that much; NaN (unless bitwise equal), infinities against
// finite values and every larger difference keep the general form.
template <int NT>
__device__ inline bool tile_pair_differs(const double *__restrict__ d, long ld, int nao, int ti, int tj,
                                         unsigned long long (*t)[65]) {
    // NT threads: 64 / (NT / 64) rows each, all loads of a phase issued before any is used (clamped
    // in-bounds addresses, values outside the matrix replaced afterwards)
    constexpr int DR = NT / 64, PER = 64 / DR;
    static_assert(DR * PER == 64 && DR * 64 == NT, "a multiple of 64 threads dividing 4096");
    const int c = threadIdx.x & 63, r0 = threadIdx.x >> 6;
    const int i0 = ti * 64, j0 = tj * 64;
    unsigned long long v[PER];
    __syncthreads();  // an earlier call's readers of t are done
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int i = i0 + r0 + q * DR, j = j0 + c;
        v[q] = (unsigned long long)__double_as_longlong(d[(long)min(i, nao - 1) * ld + min(j, nao - 1)]);
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int i = i0 + r0 + q * DR, j = j0 + c;
        t[r0 + q * DR][c] = i < nao && j < nao ? v[q] : 0ull;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; ++q) {  // row j0 + r, column i0 + c of d against t[c][r]
        const int j = j0 + r0 + q * DR, i = i0 + c;
        v[q] = (unsigned long long)__double_as_longlong(d[(long)min(j, nao - 1) * ld + min(i, nao - 1)]);
    }
    bool differ = false;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int r = r0 + q * DR, j = j0 + r, i = i0 + c;
        const double x = __longlong_as_double((long long)v[q]), y = __longlong_as_double((long long)t[c][r]);
        const bool close = fabs(x - y) <= 0x1p-50 * fmax(fabs(x), fabs(y));
        differ |= i < nao && j < nao && v[q] != t[c][r] && !close;
    }
    return __syncthreads_or(differ ? 1 : 0) != 0;
}

}  // namespace rho
}  // namespace cfgmm
