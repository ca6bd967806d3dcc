// rho.hip -- fused non0tab-masked AO-DM-AO contraction on FP64 MFMA (gfx950).
//
//   rho[s][g] += sum_{u,v in L_b} lhs[u][g] * D_s[u][v] * rhs[v][g],   b = g / blksize,
//   L_b = { x < nao : tab[x*ldtab + b] != 0 }   (ascending x)
//
// Replaces aodmao2rho_loop_parallel / _loop_2nd / _loop_1st and
// pack_dm_in_aodmao2rho (src/impl_aodmao2rho.rs:10-244).  The reference
// compacts L_b on the CPU per (u,v) task (get_indices_from_non0tab,
// src/impl_matmul_non0tab.rs:119-129), gathers the D sub-block and the phi
// strips into packed buffers, runs an MR x NR microkernel (D_sub . phi_rhs),
// multiplies by phi_lhs and reduces over u under a lock.  Here:
//
//   1. build_ao_lists: one workgroup per grid block compacts L_b on the device
//      (wave ballots + prefix counts) into lists[b][*], counts[b].
//   2. rho_chunk_kernel: one workgroup per (grid block, 64-row chunk of L_b)
//      for each group of up to four density matrices.  It streams 16-deep slices of the compacted v list through
//      double-buffered LDS -- the gathered D_s[u][v] tile (A operand) and the
//      gathered phi_rhs[v][g] tile (B operand) -- and accumulates
//      T_s[u][g] = sum_v D_s[u][v] phi_rhs[v][g] with v_mfma_f64_16x16x4_f64.
//      T never leaves registers: the epilogue multiplies by phi_lhs[u][g],
//      sums the four rows a lane holds, the four lane groups (cross-lane
//      shuffles) and the four waves (LDS), and writes one partial rho row per
//      chunk.  No ngrid x nao intermediate is written to HBM.
//   3. rho_reduce_kernel: rho[s][g] += sum over the chunks of g's block, in
//      increasing chunk order -- deterministic, no float atomics.
#include "internal.hpp"

namespace cfgmm {

// Per grid block b: lists[b*nao + i] = i-th x with tab[x*ldtab + b] != 0, counts[b] = |L_b|.
__global__ __launch_bounds__(256) void build_ao_lists(const uint8_t *__restrict__ tab, long ldtab, int nao, int nblk,
                                                      int *__restrict__ lists, int *__restrict__ counts) {
    const int b = blockIdx.x;
    if (b >= nblk) return;
    __shared__ int wave_cnt[256 / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int base = 0;
    for (int x0 = 0; x0 < nao; x0 += 256) {
        const int x = x0 + threadIdx.x;
        const bool on = x < nao && tab[(long)x * ldtab + b] != 0;
        const unsigned long long bal = __ballot(on);
        const int prefix = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) wave_cnt[wave] = __popcll(bal);
        __syncthreads();
        int off = base;
        for (int w = 0; w < wave; ++w) off += wave_cnt[w];
        if (on) lists[(long)b * nao + off + prefix] = x;
        int tot = 0;
        for (int w = 0; w < 256 / 64; ++w) tot += wave_cnt[w];
        base += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) counts[b] = base;
}

int launch_build_lists(hipStream_t stream, const uint8_t *tab, size_t ldtab, size_t nrows, size_t nblk, int *lists,
                       int *counts) {
    hipEvent_t t0 = trace_start(stream);
    hipLaunchKernelGGL(build_ao_lists, dim3((unsigned)nblk), dim3(256), 0, stream, tab, (long)ldtab, (int)nrows,
                       (int)nblk, lists, counts);
    trace_stop(stream, "build_ao_lists", t0);
    return hip_status(hipGetLastError(), "build_ao_lists launch");
}

// rho[s*stride_grid + g] += sum_{chunk < ceil(|L_b|/64)} partial[chunk][s][g], b = g / blksize,
// in increasing chunk order (deterministic).
__global__ void rho_reduce_kernel(double *__restrict__ rhos, const double *__restrict__ partial,
                                  const int *__restrict__ counts, long stride_grid, int ngrid, int nset,
                                  int blksize) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)nset * ngrid) return;
    const int s = (int)(idx / ngrid), g = (int)(idx % ngrid);
    const int nch = (counts[g / blksize] + 63) / 64;
    double v = 0.0;
    for (int ch = 0; ch < nch; ++ch) v += partial[((long)ch * nset + s) * ngrid + g];
    rhos[(long)s * stride_grid + g] += v;
}

int launch_rho_reduce(hipStream_t stream, double *rhos, const double *partial, const int *counts, size_t stride_grid,
                      size_t ngrid, size_t nset, size_t blksize) {
    const size_t total = nset * ngrid;
    hipEvent_t t0 = trace_start(stream);
    hipLaunchKernelGGL(rho_reduce_kernel, dim3((unsigned)div_up(total, 256)), dim3(256), 0, stream, rhos, partial,
                       counts, (long)stride_grid, (int)ngrid, (int)nset, (int)blksize);
    trace_stop(stream, "rho_reduce_kernel", t0);
    return hip_status(hipGetLastError(), "rho_reduce_kernel launch");
}

namespace {

constexpr int kThreads = 256;
constexpr int kUC = 64;  // rows of L_b per workgroup (4 waves x 16)
constexpr int kKV = 16;  // v slice depth per LDS stage
constexpr int kSeg = 2048;  // compacted-list entries staged in LDS at a time

template <int NS, int NT>
__global__ __launch_bounds__(kThreads, 2) void rho_chunk_kernel(
    double *__restrict__ partial, const double *__restrict__ dms, const double *__restrict__ ao_lhs,
    const double *__restrict__ ao_rhs, long stride_ao, long stride_grid, int nao, int ngrid, int nset, int set0,
    const int *__restrict__ lists, const int *__restrict__ counts, int blksize, int nsub) {
    constexpr int W = NT * 16;                          // grid points per workgroup
    constexpr int PHI_PER_T = kKV * W / kThreads;       // phi_rhs elements per thread per slice
    constexpr int D_ROWS_PER_T = kUC * kKV / kThreads;  // D rows per thread per slice and set (4)
    static_assert(PHI_PER_T * kThreads == kKV * W, "phi tile split");
    static_assert(D_ROWS_PER_T * kThreads == kUC * kKV, "D tile split");

    const int chunk = blockIdx.y;
    const int blk = blockIdx.x / nsub, sub = blockIdx.x % nsub;
    const int cnt = counts[blk];
    const int u_begin = chunk * kUC;
    if (u_begin >= cnt) return;
    const int *__restrict__ list = lists + (long)blk * nao;
    const int g0 = blk * blksize + sub * W;
    if (g0 >= ngrid) return;  // subtile past the end of a partial last block

    // LDS: double-buffered A (D_s[u][v] stored [v][u]) and B (phi_rhs[v][g]) slices, a
    // segment of the compacted v list, this chunk's u rows.  The cross-wave
    // reduction buffer of the epilogue reuses the A slices.
    constexpr int AS_ELEMS = 2 * NS * kKV * (kUC + 1);
    static_assert(AS_ELEMS >= (kThreads / 64) * NS * W, "epilogue buffer fits in the A slices");
    __shared__ double As_raw[AS_ELEMS];
    __shared__ double Bs[2][kKV][W];
    __shared__ int vls[kSeg];
    __shared__ int ulist[kUC];
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

    double rphi[PHI_PER_T], rd[NS][D_ROWS_PER_T];
    f64x4 acc[NS][NT];
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[s][nt] = f64x4{0.0, 0.0, 0.0, 0.0};

    for (int seg0 = 0; seg0 < cnt; seg0 += kSeg) {
        const int segn = min(kSeg, cnt - seg0);
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
        auto load_full = [&](int k0) {
            const long vcol = vls[k0 + p_d];
#pragma unroll
            for (int e = 0; e < D_ROWS_PER_T; ++e)
#pragma unroll
                for (int s = 0; s < NS; ++s) rd[s][e] = drow[s][e][vcol];
#pragma unroll
            for (int e = 0; e < PHI_PER_T; ++e) {
                const int idx = tid + kThreads * e;
                const int g = min(g0 + idx % W, ngrid - 1);
                rphi[e] = ao_rhs[(long)vls[k0 + idx / W] * stride_grid + g];
            }
        };
        auto load_tail = [&](int k0) {
            const int pv = k0 + p_d;
            const long vcol = vls[pv < segn ? pv : 0];
#pragma unroll
            for (int e = 0; e < D_ROWS_PER_T; ++e)
#pragma unroll
                for (int s = 0; s < NS; ++s) rd[s][e] = pv < segn ? drow[s][e][vcol] : 0.0;
#pragma unroll
            for (int e = 0; e < PHI_PER_T; ++e) {
                const int idx = tid + kThreads * e;
                const int pr = k0 + idx / W;
                const int g = min(g0 + idx % W, ngrid - 1);
                rphi[e] = pr < segn ? ao_rhs[(long)vls[pr] * stride_grid + g] : 0.0;
            }
        };
        auto store_slice = [&](int buf) {
#pragma unroll
            for (int e = 0; e < PHI_PER_T; ++e) {
                const int idx = tid + kThreads * e;
                Bs[buf][idx / W][idx % W] = rphi[e];
            }
#pragma unroll
            for (int e = 0; e < D_ROWS_PER_T; ++e)
#pragma unroll
                for (int s = 0; s < NS; ++s) As(buf, s, p_d, tid / kKV + (kThreads / kKV) * e) = rd[s][e];
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

        const int nfull = segn / kKV;
        if (nfull > 0) {
            load_full(0);
            store_slice(0);
        }
        __syncthreads();
        for (int t = 0; t < nfull; ++t) {
            const int cur = t & 1;
            if (t + 1 < nfull) load_full((t + 1) * kKV);
            compute_slice(cur);
            if (t + 1 < nfull) store_slice(cur ^ 1);
            __syncthreads();
        }
        if (nfull * kKV < segn) {
            load_tail(nfull * kKV);
            store_slice(0);
            __syncthreads();
            compute_slice(0);
            __syncthreads();
        }
    }

    // Epilogue: rho_partial[g] = sum_u lhs[u][g] * T[u][g] over this chunk's rows.
    double *red = As_raw;  // [4 waves][NS][W]; all slice reads finished at the last barrier
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
            lhs[r] = ao_lhs[(long)ulist[i] * stride_grid + (g_ok ? g : ngrid - 1)];
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
            partial[((long)chunk * nset + set0 + s) * ngrid + g] = v;
        }
    }
}

template <int NS>
int launch_ns(int nt, dim3 grid, hipStream_t stream, double *partial, const double *dms, const double *ao_lhs,
              const double *ao_rhs, long stride_ao, long stride_grid, int nao, int ngrid, int nset, int set0,
              const int *lists, const int *counts, int blksize, int nsub) {
    hipEvent_t t0 = trace_start(stream);
#define CFGMM_RHO_LAUNCH(NT_)                                                                                      \
    hipLaunchKernelGGL((rho_chunk_kernel<NS, NT_>), grid, dim3(kThreads), 0, stream, partial, dms, ao_lhs, ao_rhs, \
                       stride_ao, stride_grid, nao, ngrid, nset, set0, lists, counts, blksize, nsub)
    switch (nt) {
        case 1: CFGMM_RHO_LAUNCH(1); break;
        case 2: CFGMM_RHO_LAUNCH(2); break;
        case 3: CFGMM_RHO_LAUNCH(3); break;
        default: CFGMM_RHO_LAUNCH(4); break;
    }
#undef CFGMM_RHO_LAUNCH
    trace_stop(stream, "rho_chunk_kernel", t0);
    return hip_status(hipGetLastError(), "rho_chunk_kernel launch");
}

}  // namespace

// Columns per workgroup: the widest of 64/48/32/16 grid points that tiles a
// block exactly (blksize is a multiple of 16).
static int rho_tile_nt(size_t blksize) {
    const size_t q = blksize / 16;
    for (int nt = 4; nt >= 1; --nt)
        if (q % nt == 0) return nt;
    return 1;
}

int launch_rho(hipStream_t stream, double *rhos, const double *dms, const double *ao_lhs, const double *ao_rhs,
               size_t stride_ao, size_t stride_grid, size_t nao, size_t ngrid, size_t nset, const uint8_t *tab,
               size_t ldtab, size_t blksize) {
    if (nao == 0 || ngrid == 0 || nset == 0) return CFGMM_OK;
    if (nao > INT32_MAX || ngrid > INT32_MAX || nset > 65535) return fail(CFGMM_EINVAL, "rho: size out of range");
    const size_t nblk = div_up(ngrid, blksize);
    const size_t maxchunks = div_up(nao, kUC);
    const int nt = rho_tile_nt(blksize);
    const int nsub = (int)(blksize / (16 * nt));
    if (nblk * nsub > (size_t)INT32_MAX || maxchunks > 65535u) return fail(CFGMM_EINVAL, "rho: too many grid blocks");

    // workspace: [lists nblk*nao int][counts nblk int][pad][partial maxchunks*nset*ngrid double]
    const size_t lists_bytes = sizeof(int) * nblk * nao;
    const size_t counts_bytes = sizeof(int) * nblk;
    const size_t part_off = div_up(lists_bytes + counts_bytes, 256) * 256;
    const size_t part_bytes = sizeof(double) * maxchunks * nset * ngrid;
    void *ws = nullptr;
    int st = workspace_get(stream, part_off + part_bytes, &ws);
    if (st != CFGMM_OK) return st;
    int *lists = static_cast<int *>(ws);
    int *counts = lists + nblk * nao;
    double *partial = reinterpret_cast<double *>(static_cast<char *>(ws) + part_off);

    st = launch_build_lists(stream, tab, ldtab, nao, nblk, lists, counts);
    if (st != CFGMM_OK) return st;

    for (size_t set0 = 0; set0 < nset; set0 += 4) {
        const int ns = (int)std::min<size_t>(4, nset - set0);
        dim3 grid((unsigned)(nblk * nsub), (unsigned)maxchunks, 1);
        const long sa = (long)stride_ao, sg = (long)stride_grid;
        switch (ns) {
            case 1: st = launch_ns<1>(nt, grid, stream, partial, dms, ao_lhs, ao_rhs, sa, sg, (int)nao, (int)ngrid, (int)nset, (int)set0, lists, counts, (int)blksize, nsub); break;
            case 2: st = launch_ns<2>(nt, grid, stream, partial, dms, ao_lhs, ao_rhs, sa, sg, (int)nao, (int)ngrid, (int)nset, (int)set0, lists, counts, (int)blksize, nsub); break;
            case 3: st = launch_ns<3>(nt, grid, stream, partial, dms, ao_lhs, ao_rhs, sa, sg, (int)nao, (int)ngrid, (int)nset, (int)set0, lists, counts, (int)blksize, nsub); break;
            default: st = launch_ns<4>(nt, grid, stream, partial, dms, ao_lhs, ao_rhs, sa, sg, (int)nao, (int)ngrid, (int)nset, (int)set0, lists, counts, (int)blksize, nsub); break;
        }
        if (st != CFGMM_OK) return st;
    }
    return launch_rho_reduce(stream, rhos, partial, counts, stride_grid, ngrid, nset, blksize);
}

}  // namespace cfgmm
