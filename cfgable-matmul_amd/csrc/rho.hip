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

namespace {

constexpr int kThreads = 256;
constexpr int kUC = 64;  // rows of L_b per workgroup (4 waves x 16)
constexpr int kKV = 16;  // v slice depth per LDS stage

template <int NS, int NT>
__global__ __launch_bounds__(kThreads, 2) void rho_chunk_kernel(
    double *__restrict__ partial, const double *__restrict__ dms, const double *__restrict__ ao_lhs,
    const double *__restrict__ ao_rhs, long stride_ao, long stride_grid, int nao, int ngrid, int nset, int set0,
    const int *__restrict__ lists, const int *__restrict__ counts, int blksize, int nsub) {
    constexpr int W = NT * 16;                        // grid points per workgroup
    constexpr int PHI_PER_T = kKV * W / kThreads;     // phi_rhs elements per thread per slice
    constexpr int D_PER_T = NS * kKV * kUC / kThreads;  // D elements per thread per slice
    static_assert(PHI_PER_T * kThreads == kKV * W, "phi tile split");
    static_assert(D_PER_T * kThreads == NS * kKV * kUC, "D tile split");

    const int chunk = blockIdx.y;
    const int blk = blockIdx.x / nsub, sub = blockIdx.x % nsub;
    const int cnt = counts[blk];
    const int u_begin = chunk * kUC;
    if (u_begin >= cnt) return;
    const int *__restrict__ list = lists + (long)blk * nao;
    const int g0 = blk * blksize + sub * W;

    __shared__ double Bs[2][kKV][W];            // phi_rhs[v][g]
    __shared__ double As[2][NS][kKV][kUC + 1];  // D_s[u][v], stored [v][u]
    __shared__ double red[kThreads / 64][NS][W];
    __shared__ int ulist[kUC];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int lrow = lane & 15, lk = lane >> 4;
    if (tid < kUC) ulist[tid] = (u_begin + tid < cnt) ? list[u_begin + tid] : -1;
    __syncthreads();

    const long dstride = stride_ao * stride_ao;
    double rphi[PHI_PER_T], rd[D_PER_T];
    auto load_slice = [&](int k0) {
#pragma unroll
        for (int e = 0; e < PHI_PER_T; ++e) {
            const int idx = tid + kThreads * e;
            const int r = idx / W, col = idx % W;
            const int g = g0 + col;
            double v = 0.0;
            if (k0 + r < cnt && g < ngrid) v = ao_rhs[(long)list[k0 + r] * stride_grid + g];
            rphi[e] = v;
        }
#pragma unroll
        for (int e = 0; e < D_PER_T; ++e) {
            const int idx = tid + kThreads * e;
            const int s = idx / (kKV * kUC);
            const int rem = idx % (kKV * kUC);
            const int i = rem / kKV, p = rem % kKV;
            const int u = ulist[i];
            double v = 0.0;
            if (u >= 0 && k0 + p < cnt)
                v = dms[(long)(set0 + s) * dstride + (long)u * stride_ao + list[k0 + p]];
            rd[e] = v;
        }
    };
    auto store_slice = [&](int buf) {
#pragma unroll
        for (int e = 0; e < PHI_PER_T; ++e) {
            const int idx = tid + kThreads * e;
            Bs[buf][idx / W][idx % W] = rphi[e];
        }
#pragma unroll
        for (int e = 0; e < D_PER_T; ++e) {
            const int idx = tid + kThreads * e;
            const int s = idx / (kKV * kUC);
            const int rem = idx % (kKV * kUC);
            As[buf][s][rem % kKV][rem / kKV] = rd[e];
        }
    };

    f64x4 acc[NS][NT];
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[s][nt] = f64x4{0.0, 0.0, 0.0, 0.0};

    const int nslices = (cnt + kKV - 1) / kKV;
    load_slice(0);
    store_slice(0);
    __syncthreads();
    for (int t = 0; t < nslices; ++t) {
        const int cur = t & 1;
        if (t + 1 < nslices) load_slice((t + 1) * kKV);
#pragma unroll
        for (int kk = 0; kk < kKV / 4; ++kk) {
            double fa[NS], fb[NT];
#pragma unroll
            for (int s = 0; s < NS; ++s) fa[s] = As[cur][s][kk * 4 + lk][wave * 16 + lrow];
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) fb[nt] = Bs[cur][kk * 4 + lk][nt * 16 + lrow];
#pragma unroll
            for (int s = 0; s < NS; ++s)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                    acc[s][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[s], fb[nt], acc[s][nt], 0, 0, 0);
        }
        if (t + 1 < nslices) store_slice(cur ^ 1);
        __syncthreads();
    }

    // Epilogue: rho_partial[g] = sum_u lhs[u][g] * T[u][g] over this chunk's rows.
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        const int g = g0 + nt * 16 + lrow;
        double lhs[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int u = ulist[wave * 16 + lk + 4 * r];  // D[l/16 + 4r][l%16]
            lhs[r] = (u >= 0 && g < ngrid) ? ao_lhs[(long)u * stride_grid + g] : 0.0;
        }
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            double part = acc[s][nt][0] * lhs[0];
#pragma unroll
            for (int r = 1; r < 4; ++r) part = fma(acc[s][nt][r], lhs[r], part);
            part += __shfl_xor(part, 16);
            part += __shfl_xor(part, 32);
            if (lk == 0) red[wave][s][nt * 16 + lrow] = part;
        }
    }
    __syncthreads();
    for (int idx = tid; idx < NS * W; idx += kThreads) {
        const int s = idx / W, col = idx % W;
        const int g = g0 + col;
        if (g < ngrid) {
            double v = red[0][s][col];
#pragma unroll
            for (int w = 1; w < kThreads / 64; ++w) v += red[w][s][col];
            partial[((long)chunk * nset + set0 + s) * ngrid + g] = v;
        }
    }
}

// rho[s*stride_grid + g] += sum_{chunk < ceil(|L_b|/64)} partial[chunk][s][g], b = g / blksize.
__global__ void rho_reduce_kernel(double *__restrict__ rhos, const double *__restrict__ partial,
                                  const int *__restrict__ counts, long stride_grid, int ngrid, int nset,
                                  int blksize) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)nset * ngrid) return;
    const int s = (int)(idx / ngrid), g = (int)(idx % ngrid);
    const int nch = (counts[g / blksize] + kUC - 1) / kUC;
    double v = 0.0;
    for (int ch = 0; ch < nch; ++ch) v += partial[((long)ch * nset + s) * ngrid + g];
    rhos[(long)s * stride_grid + g] += v;
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
    const size_t total = nset * ngrid;
    hipEvent_t t0 = trace_start(stream);
    hipLaunchKernelGGL(rho_reduce_kernel, dim3((unsigned)div_up(total, 256)), dim3(256), 0, stream, rhos, partial,
                       counts, (long)stride_grid, (int)ngrid, (int)nset, (int)blksize);
    trace_stop(stream, "rho_reduce_kernel", t0);
    return hip_status(hipGetLastError(), "rho_reduce_kernel launch");
}

}  // namespace cfgmm
