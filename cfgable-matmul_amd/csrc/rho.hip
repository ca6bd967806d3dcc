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
//   2. rho_plan_kernel orders the (block, chunk) tasks by the AO range of their
//      rows, so that tasks sharing rows of D run together on one XCD (block-major
//      order, a tunable, reads every chunk's D rows from memory instead).
//   3. rho_chunk_kernel: one workgroup per (grid block, 64-row chunk of L_b)
//      for each group of up to four density matrices.  It streams 16-deep slices of the compacted v list through
//      double-buffered LDS -- the gathered D_s[u][v] tile (A operand) and the
//      gathered phi_rhs[v][g] tile (B operand) -- and accumulates
//      T_s[u][g] = sum_v D_s[u][v] phi_rhs[v][g] with v_mfma_f64_16x16x4_f64.
//      T never leaves registers: the epilogue multiplies by phi_lhs[u][g],
//      sums the four rows a lane holds, the four lane groups (cross-lane
//      shuffles) and the four waves (LDS), and writes one partial rho row per
//      chunk.  No ngrid x nao intermediate is written to HBM.
//   4. rho_reduce_kernel: rho[s][g] += sum over the chunks of g's block, in
//      increasing chunk order -- deterministic, no float atomics.
//
// When lhs and rhs are one array (the density of one set of orbitals), the other workgroups of
// the plan launch test every D_s for D_s == D_s^T (entries equal bitwise or within 2^-50 of the
// larger magnitude) beside the plan; if all are, each chunk contracts only the list positions
// from its own rows on, those past its rows twice: half the gathers and MFMAs.  The flag stays on the device (no host synchronisation); the chunk
// kernel reads it.
#include "internal.hpp"
#include "rho_body.hpp"

#include <atomic>
#include <map>
#include <mutex>
#include <string>

namespace cfgmm {

// Per grid block b: lists[b*nao + i] = i-th x with tab[x*ldtab + b] != 0, counts[b] = |L_b|.
__global__ __launch_bounds__(256) void build_ao_lists(const uint8_t *__restrict__ tab, long ldtab, int nao, int nblk,
                                                      int *__restrict__ lists, int *__restrict__ counts,
                                                      int *__restrict__ keys, int maxchunks) {
    const int b = blockIdx.x;
    if (b >= nblk) return;
    __shared__ int wave_cnt[256 / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // the column's bytes of kQ chunks of 256 rows loaded together (clamped addresses), then compacted
    // chunk by chunk: one load latency per kQ chunks instead of one per chunk
    constexpr int kQ = 16;
    int base = 0;
    for (int x00 = 0; x00 < nao; x00 += 256 * kQ) {
        uint8_t v[kQ];
#pragma unroll
        for (int q = 0; q < kQ; ++q) v[q] = tab[(long)min(x00 + q * 256 + (int)threadIdx.x, nao - 1) * ldtab + b];
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
            const int x0 = x00 + q * 256;
            if (x0 >= nao) break;  // uniform
            const int x = x0 + threadIdx.x;
            const bool on = x < nao && v[q] != 0;
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
    }
    if (threadIdx.x == 0) counts[b] = base;
    if (keys) {
        __syncthreads();  // the list entries written by this workgroup
        const int *__restrict__ lb = lists + (long)b * nao;
        for (int c = threadIdx.x; c < maxchunks; c += 256) {
            int k = -1;
            if (c * 64 < base) {
                const long mid = ((long)lb[c * 64] + lb[min(c * 64 + 64, base) - 1]) / 2;
                k = (int)(mid * kPlanBuckets / nao);
            }
            keys[(long)b * maxchunks + c] = k;
        }
    }
}

int launch_build_lists(hipStream_t stream, const uint8_t *tab, size_t ldtab, size_t nrows, size_t nblk, int *lists,
                       int *counts, int *keys, size_t maxchunks) {
    hipEvent_t t0 = trace_start(stream);
    hipLaunchKernelGGL(build_ao_lists, dim3((unsigned)nblk), dim3(256), 0, stream, tab, (long)ldtab, (int)nrows,
                       (int)nblk, lists, counts, keys, (int)maxchunks);
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
    const double *__restrict__ p = partial + (long)s * ngrid + g;
    const long step = (long)nset * ngrid;
    // eight loads in flight at a time (clamped indices, no branch around a load), summed in chunk order
    double v = 0.0;
    for (int c0 = 0; c0 < nch; c0 += 8) {
        double x[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) x[i] = p[min(c0 + i, nch - 1) * step];
#pragma unroll
        for (int i = 0; i < 8; ++i) v = c0 + i < nch ? v + x[i] : v;
    }
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

using rho::kThreads;
using rho::kUC;

constexpr int kPlanThreads = 1024;

// Orders the (block, chunk) tasks of one evaluation by the AO range their rows cover (the
// bucketed midpoint of the chunk's first and last row, keys from build_ao_lists): tasks that run
// together on an XCD then share rows of D in its L2, where block-major order reads each chunk's D
// rows from memory (the columns of a chunk are a third of each row, spread over all its cache
// lines).  perm[0 .. *ntask) = slot b*maxchunks + c of the non-empty chunks, bucket-ordered; the
// order inside a bucket follows atomics and varies from call to call, which the result does not
// depend on (each task writes its own partial slot, reduced in fixed chunk order).
// bounds[0..8]: the sequence cut into eight ranges of equal work in the symmetric form, one per
// XCD; a task's work is rows x (positions from its first row on + 64) (pwork[pos]).  The general
// form's work is even along the sequence: it takes equal counts.
// Workgroup 0 plans.  With dms, workgroups 1.. compare the 64x64 tile pairs (ti <= tj) of every
// D_s with their transposes and write `epoch` to *symflag at a difference: after this kernel,
// *symflag != epoch means every D_s is symmetric.
__global__ __launch_bounds__(kPlanThreads) void rho_plan_kernel(
    const int *__restrict__ keys, const int *__restrict__ counts, int nblk, int maxchunks, int *__restrict__ perm,
    float *__restrict__ pwork, int *__restrict__ ntask, int *__restrict__ bounds, const double *__restrict__ dms,
    long stride_ao, int nao, int npairs, int *__restrict__ symflag, unsigned epoch) {
    const int tid = threadIdx.x;
    if (blockIdx.x > 0) {
        __shared__ unsigned long long t[64][65];
        const int w = blockIdx.x - 1, s = w / npairs;
        const long k = w % npairs;
        // upper-triangle index k = tj (tj + 1) / 2 + ti, 0 <= ti <= tj
        long tj = (long)((sqrt(8.0 * (double)k + 1.0) - 1.0) / 2.0);
        while (tj * (tj + 1) / 2 > k) --tj;
        while ((tj + 1) * (tj + 2) / 2 <= k) ++tj;
        const int ti = (int)(k - tj * (tj + 1) / 2);
        if (rho::tile_pair_differs<kPlanThreads>(dms + (long)s * stride_ao * stride_ao, stride_ao, nao, ti, (int)tj, t) &&
            tid == 0)
            __hip_atomic_store(symflag, (int)epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    __shared__ int hist[kPlanBuckets];
    __shared__ double wsum[kPlanThreads];
    __shared__ int s_ntask;
    const long nslot = (long)nblk * maxchunks;
    for (int i = tid; i < kPlanBuckets; i += kPlanThreads) hist[i] = 0;
    __syncthreads();
    for (long sl = tid; sl < nslot; sl += kPlanThreads) {
        const int k = keys[sl];
        if (k >= 0) atomicAdd(&hist[k], 1);
    }
    __syncthreads();
    // exclusive scan of the bucket counts (Hillis-Steele over one bucket per thread)
    static_assert(kPlanBuckets == kPlanThreads, "one bucket per thread");
    const int mine = hist[tid];
    int incl = mine;
    for (int d = 1; d < kPlanBuckets; d <<= 1) {
        __syncthreads();
        const int add = tid >= d ? hist[tid - d] : 0;
        __syncthreads();
        incl += add;
        hist[tid] = incl;
    }
    __syncthreads();
    hist[tid] = incl - mine;
    if (tid == kPlanThreads - 1) {
        *ntask = incl;
        s_ntask = incl;
    }
    __syncthreads();
    for (long sl = tid; sl < nslot; sl += kPlanThreads) {
        const int k = keys[sl];
        if (k < 0) continue;
        const int pos = atomicAdd(&hist[k], 1);
        perm[pos] = (int)sl;
        const int b = (int)(sl / maxchunks), c = (int)(sl % maxchunks), cnt = counts[b];
        pwork[pos] = (float)min(kUC, cnt - c * kUC) * (float)(cnt - c * kUC + kUC);
    }
    __syncthreads();  // pwork complete (the global writes of this workgroup are visible to it)

    // equal-work XCD ranges: each thread takes a contiguous piece of the sequence
    const int n = s_ntask;
    const int per = (n + kPlanThreads - 1) / kPlanThreads;
    const int p0 = min(n, tid * per), p1 = min(n, p0 + per);
    double mine2 = 0.0;
    for (int pos = p0; pos < p1; ++pos) mine2 += pwork[pos];
    double inc2 = mine2;  // inclusive scan of the per-thread sums (Hillis-Steele)
    wsum[tid] = inc2;
    for (int d = 1; d < kPlanThreads; d <<= 1) {
        __syncthreads();
        const double a0 = tid >= d ? wsum[tid - d] : 0.0;
        __syncthreads();
        inc2 += a0;
        wsum[tid] = inc2;
    }
    __syncthreads();
    const double tot = wsum[kPlanThreads - 1];
    if (tid == 0) {
        bounds[0] = 0;
        for (int x = 1; x <= 8; ++x) bounds[x] = n;
    }
    __syncthreads();
    if (tot <= 0.0) return;
    // position pos goes to XCD floor(8 * (work before pos) / total); bounds[x] = first position of XCD >= x
    double before = inc2 - mine2;
    int prev = p0 > 0 && p0 < p1 ? (int)min(7.0, floor((before - pwork[p0 - 1]) * 8.0 / tot)) : 0;
    for (int pos = p0; pos < p1; ++pos) {
        const int xcd = (int)min(7.0, floor(before * 8.0 / tot));
        for (int x = prev + 1; x <= xcd; ++x) bounds[x] = pos;
        prev = xcd;
        before += pwork[pos];
    }
}

// *sym = 0 unless every D_s (s < nset) equals its transpose on [0, nao)^2 (tile_pair_differs's rule).  Workgroup
// (tj, ti, s) with ti <= tj compares the 64x64 tile (ti, tj) with the transpose of (tj, ti)
// through LDS; the others exit.
__global__ __launch_bounds__(256) void dm_symmetry_kernel(const double *__restrict__ dms, long stride_ao, int nao,
                                                          int *__restrict__ sym) {
    const int tj = blockIdx.x, ti = blockIdx.y, s = blockIdx.z;
    if (ti > tj) return;
    __shared__ unsigned long long t[64][65];
    if (rho::tile_pair_differs<256>(dms + (long)s * stride_ao * stride_ao, stride_ao, nao, ti, tj, t) &&
        threadIdx.x == 0)
        __hip_atomic_store(sym, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// OFF32: 32-bit byte offsets into ao_rhs (see rho_chunk_body)
template <int NS, int NT, bool OFF32>
__global__ __launch_bounds__(kThreads, 2) void rho_chunk_kernel(
    double *__restrict__ partial, const double *__restrict__ dms, const double *__restrict__ ao_lhs,
    const double *__restrict__ ao_rhs, long stride_ao, long stride_grid, int nao, int ngrid, int nset, int set0,
    const int *__restrict__ lists, const int *__restrict__ counts, int blksize, int nsub, int maxchunks,
    int ngroups, int ncomp, long lhs_comp_stride, long part_comp_stride, const int *__restrict__ perm,
    const int *__restrict__ ntask, const int *__restrict__ sym, unsigned epoch, const int *__restrict__ bounds) {
    __shared__ __attribute__((aligned(16))) unsigned char smem[rho::chunk_lds_bytes<NS, NT>()];
    const bool symmetric = sym && *sym != (int)epoch;
    bool first = true;
    // one (block, chunk, group of NS density matrices, subtile) task
    auto run = [&](int blk, int chunk, int grp, int sub) {
        rho::RhoChunk t;
        t.cnt = counts[blk];
        t.chunk = chunk;
        t.g0 = blk * blksize + sub * NT * 16;
        if (t.chunk * kUC >= t.cnt || t.g0 >= ngrid) return;  // empty chunk slot, subtile past the grid
        if (!first) __syncthreads();  // the previous task's epilogue is done with LDS
        first = false;
        t.partial = partial;
        t.dms = dms;
        t.ao_lhs = ao_lhs;
        t.ao_rhs = ao_rhs;
        t.stride_ao = stride_ao;
        t.stride_grid = stride_grid;
        t.ngrid = ngrid;
        t.nset = nset;
        t.set0 = set0 + grp * NS;
        t.list = lists + (long)blk * nao;
        t.ncomp = ncomp;
        t.lhs_comp_stride = lhs_comp_stride;
        t.part_comp_stride = part_comp_stride;
        if (symmetric)
            rho::rho_chunk_body<NS, NT, true, OFF32>(t, smem);
        else
            rho::rho_chunk_body<NS, NT, false, OFF32>(t, smem);
    };
    const int x = blockIdx.x % 8, k = blockIdx.x / 8;
    if (perm && symmetric) {
        // AO-ordered tasks of the symmetric form, whose work falls with the AO range: XCD x
        // (workgroup i runs on XCD i % 8) takes the planned positions [bounds[x], bounds[x+1]) --
        // equal work per XCD -- for every group and subtile, group-major; its workgroups stride
        // over them (several tasks each where short tasks are many)
        const int b0 = bounds[x], npos = bounds[x + 1] - b0;
        const long per = (long)npos * nsub, items = per * ngroups;
        for (long kk = k; kk < items; kk += gridDim.x / 8) {
            const int grp = (int)(kk / per);
            const long rj = kk % per;
            const int sl = perm[b0 + rj / nsub];
            run(sl / maxchunks, sl % maxchunks, grp, (int)(rj % nsub));
        }
    } else if (perm) {
        // AO-ordered tasks (group of NS density matrices, planned chunk j, subtile), the
        // sequence cut into eight contiguous parts of equal count, one per XCD (the grid is a
        // multiple of 8 with at least as many workgroups as tasks)
        const long per = (long)*ntask * nsub, total = per * ngroups;
        const long part = (total + 7) / 8, tt = x * part + k;
        if (k >= part || tt >= total) return;
        const long rj = tt % per;
        const int sl = perm[rj / nsub];
        run(sl / maxchunks, sl % maxchunks, (int)(tt / per), (int)(rj % nsub));
    } else {
        // block-major (block, subtile, group, chunk): the tasks of one block, which share its
        // phi tile, run back to back on one XCD; empty chunk slots exit
        const int ntasks = gridDim.x, q = ntasks / 8, r = ntasks % 8;
        const int L = x * q + min(x, r) + k;
        const int chunk = L % maxchunks;
        const int rest = L / maxchunks;
        const int grp = rest % ngroups;
        const int rest2 = rest / ngroups;
        run(rest2 / nsub, chunk, grp, rest2 % nsub);
    }
}

template <int NS, bool OFF32>
int launch_ns(int nt, dim3 grid, hipStream_t stream, double *partial, const double *dms, const double *ao_lhs,
              const double *ao_rhs, long stride_ao, long stride_grid, int nao, int ngrid, int nset, int set0,
              const int *lists, const int *counts, int blksize, int nsub, int maxchunks, int ngroups, int ncomp,
              long lhs_cs, long part_cs, const int *perm, const int *ntask, const int *sym, unsigned epoch,
              const int *bounds) {
    hipEvent_t t0 = trace_start(stream);
#define CFGMM_RHO_LAUNCH(NT_)                                                                                      \
    hipLaunchKernelGGL((rho_chunk_kernel<NS, NT_, OFF32>), grid, dim3(kThreads), 0, stream, partial, dms, ao_lhs, ao_rhs, \
                       stride_ao, stride_grid, nao, ngrid, nset, set0, lists, counts, blksize, nsub, maxchunks, \
                       ngroups, ncomp, lhs_cs, part_cs, perm, ntask, sym, epoch, bounds)
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

int launch_dm_symmetry(hipStream_t stream, const double *dms, size_t stride_ao, size_t nao, size_t nset, int *flag) {
    if (nao == 0 || nset == 0) return CFGMM_OK;
    const size_t nt64 = div_up(nao, 64);
    if (nt64 > 65535 || nset > 65535) return fail(CFGMM_EINVAL, "dm symmetry: matrix too large");
    hipEvent_t t0 = trace_start(stream);
    hipLaunchKernelGGL(dm_symmetry_kernel, dim3((unsigned)nt64, (unsigned)nt64, (unsigned)nset), dim3(256), 0, stream,
                       dms, (long)stride_ao, (int)nao, flag);
    trace_stop(stream, "dm_symmetry_kernel", t0);
    return hip_status(hipGetLastError(), "dm_symmetry_kernel launch");
}

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
    return launch_rho_multi(stream, rhos, 0, dms, ao_lhs, 0, 1, ao_rhs, stride_ao, stride_grid, nao, ngrid, nset, tab,
                            ldtab, blksize);
}

// ncomp evaluations sharing D, rhs and the mask: rho_c = rhos + c*rho_cs, lhs_c = ao_lhs + c*lhs_cs.
int launch_rho_multi(hipStream_t stream, double *rhos, size_t rho_cs, const double *dms, const double *ao_lhs,
                     size_t lhs_cs, size_t ncomp, const double *ao_rhs, size_t stride_ao, size_t stride_grid,
                     size_t nao, size_t ngrid, size_t nset, const uint8_t *tab, size_t ldtab, size_t blksize,
                     const RhoScratch *scratch) {
    if (nao == 0 || ngrid == 0 || nset == 0 || ncomp == 0) return CFGMM_OK;
    if (ncomp > 16) return fail(CFGMM_EINVAL, "rho: at most 16 left-hand components");
    if (nao > INT32_MAX || ngrid > INT32_MAX || nset > 65535) return fail(CFGMM_EINVAL, "rho: size out of range");
    const size_t nblk = div_up(ngrid, blksize);
    const size_t maxchunks = div_up(nao, kUC);
    const int nt = rho_tile_nt(blksize);
    const int nsub = (int)(blksize / (16 * nt));

    // workspace: [lists nblk*nao int][counts nblk int][perm nblk*maxchunks int][ntask int][sym int]
    //            [bounds 9 int][keys nblk*maxchunks int][pwork nblk*maxchunks float][pad]
    //            [partial ncomp*maxchunks*nset*ngrid double]
    const size_t nslot = nblk * maxchunks;
    const size_t head_bytes = sizeof(int) * (nblk * nao + nblk + nslot + 2 + 9 + nslot) + sizeof(float) * nslot;
    const size_t part_off = div_up(head_bytes, 256) * 256;
    const size_t part_cs = maxchunks * nset * ngrid;  // doubles per component
    const size_t part_bytes = sizeof(double) * part_cs * ncomp;
    void *ws = nullptr;
    int st = CFGMM_OK;
    if (scratch) {
        if (part_off + part_bytes > scratch->bytes)
            return fail(CFGMM_EINVAL, "rho: this problem needs " + std::to_string(part_off + part_bytes) +
                                          " bytes of scratch; the allocation-free entry point has " +
                                          std::to_string(scratch->bytes));
        ws = scratch->ptr;
    } else {
        st = workspace_get(stream, part_off + part_bytes, &ws);
        if (st != CFGMM_OK) return st;
    }
    int *lists = static_cast<int *>(ws);
    int *counts = lists + nblk * nao;
    int *perm = counts + nblk;
    int *ntask = perm + nslot;
    int *bounds = ntask + 2;
    // the symmetry flag lives in a word of its own, beside the workspace (the scratch): inside it,
    // where each problem size puts other data, it could hold a value that equals this call's epoch
    int *sym = scratch ? scratch->flag : workspace_flag(stream);
    int *keys = bounds + 9;
    float *pwork = reinterpret_cast<float *>(keys + nslot);
    double *partial = reinterpret_cast<double *>(static_cast<char *>(ws) + part_off);

    const bool ordered = tune_get(CFGMM_TUNE_RHO_ORDER) == 0;
    st = launch_build_lists(stream, tab, ldtab, nao, nblk, lists, counts, ordered ? keys : nullptr, maxchunks);
    if (st != CFGMM_OK) return st;
    // the symmetric form: one set of orbitals on both sides and every D_s symmetric, tested on the
    // device by the plan launch's other workgroups (a flag against this call's epoch, unique in the
    // process: no clearing)
    const size_t nt64 = div_up(nao, 64), npairs = nt64 * (nt64 + 1) / 2;
    const bool try_sym = ordered && ncomp == 1 && ao_lhs == ao_rhs && tune_get(CFGMM_TUNE_RHO_SYMMETRIC) == 0 &&
                         npairs * nset + 1 <= (size_t)INT32_MAX;
    const unsigned epoch = try_sym ? (scratch ? scratch->epoch : workspace_epoch()) : 0u;
    if (ordered) {
        hipEvent_t t0 = trace_start(stream);
        hipLaunchKernelGGL(rho_plan_kernel, dim3((unsigned)(1 + (try_sym ? npairs * nset : 0))), dim3(kPlanThreads),
                           0, stream, keys, counts, (int)nblk, (int)maxchunks, perm, pwork, ntask, bounds, dms,
                           (long)stride_ao, (int)nao, (int)npairs, sym, epoch);
        trace_stop(stream, "rho_plan_kernel", t0);
        st = hip_status(hipGetLastError(), "rho_plan_kernel launch");
        if (st != CFGMM_OK) return st;
    }
    const int *symp = try_sym ? sym : nullptr;

    // Density matrices per workgroup: up to max_sets.  One launch covers every full group of NS
    // sets (the group is a task coordinate); a smaller remainder group gets a launch of its own.
    const int NS = (int)std::min<size_t>((size_t)tune_get(CFGMM_TUNE_RHO_MAX_SETS), nset);
    const size_t full = nset / NS, rem = nset % NS;
    const long sa = (long)stride_ao, sg = (long)stride_grid;
    // every byte offset into ao_rhs below 2^32: its slice loads take 32-bit offsets
    const bool off32 = (uint64_t)nao * stride_grid * 8 <= (uint64_t)UINT32_MAX;
    auto launch_groups = [&](int ns, size_t set0, size_t ngroups) -> int {
        if (nblk * nsub * maxchunks * ngroups + 8 > (size_t)INT32_MAX)
            return fail(CFGMM_EINVAL, "rho: too many workgroups for one launch");
        const size_t slots = nblk * nsub * ngroups * maxchunks;
        const dim3 grid((unsigned)(ordered ? div_up(slots, 8) * 8 : slots), 1, 1);
        const int *pp = ordered ? perm : nullptr;
#define CFGMM_RHO_NS(NS_)                                                                          \
    (off32 ? launch_ns<NS_, true>(nt, grid, stream, partial, dms, ao_lhs, ao_rhs, sa, sg, (int)nao, (int)ngrid, (int)nset, (int)set0, lists, counts, (int)blksize, nsub, (int)maxchunks, (int)ngroups, (int)ncomp, (long)lhs_cs, (long)part_cs, pp, ntask, symp, epoch, bounds) \
           : launch_ns<NS_, false>(nt, grid, stream, partial, dms, ao_lhs, ao_rhs, sa, sg, (int)nao, (int)ngrid, (int)nset, (int)set0, lists, counts, (int)blksize, nsub, (int)maxchunks, (int)ngroups, (int)ncomp, (long)lhs_cs, (long)part_cs, pp, ntask, symp, epoch, bounds))
        switch (ns) {
            case 1: return CFGMM_RHO_NS(1);
            case 2: return CFGMM_RHO_NS(2);
            case 3: return CFGMM_RHO_NS(3);
            default: return CFGMM_RHO_NS(4);
#undef CFGMM_RHO_NS
        }
    };
    st = launch_groups(NS, 0, full);
    if (st != CFGMM_OK) return st;
    if (rem) {
        st = launch_groups((int)rem, full * NS, 1);
        if (st != CFGMM_OK) return st;
    }
    for (size_t c = 0; c < ncomp; ++c) {
        st = launch_rho_reduce(stream, rhos + c * rho_cs, partial + c * part_cs, counts, stride_grid, ngrid, nset,
                               blksize);
        if (st != CFGMM_OK) return st;
    }
    return CFGMM_OK;
}

// ---------------------------------------------------------------------------------------------
// Allocation-free evaluation (cfgmm_aodmao2rho_dev_noalloc): the AO lists, plan and chunk
// partials of launch_rho live in `scratch`, a file-scope device array, instead of the stream's
// workspace.  Each GPU holds its own copy at its own address; in host code the name `scratch`
// only identifies the array, so its device address comes from hipGetSymbolAddress (once per
// device, with that device current) and only that address is passed to kernels.  Every call on
// a GPU shares the array: calls that overlap in time on one GPU race.  Nothing in it needs
// clearing: each evaluation rewrites the lists, plan and partials it reads, and the symmetry
// flag is tagged with a per-call epoch.
constexpr size_t kRhoScratchBytes = CFGMM_RHO_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kRhoScratchBytes];

static std::mutex g_scratch_mutex;
static std::map<int, void *> g_scratch_addr;  // device -> address of its `scratch`
static std::atomic<unsigned> g_scratch_epoch{0};

static int scratch_address(void **out) {
    int dev = 0;
    CFGMM_HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    auto it = g_scratch_addr.find(dev);
    if (it == g_scratch_addr.end()) {
        void *p = nullptr;
        CFGMM_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(CFGMM_EHIP, "hipGetSymbolAddress(scratch) returned no address");
        it = g_scratch_addr.emplace(dev, p).first;
    }
    *out = it->second;
    return CFGMM_OK;
}

int launch_rho_noalloc(hipStream_t stream, double *rhos, const double *dms, const double *ao_lhs,
                       const double *ao_rhs, size_t stride_ao, size_t stride_grid, size_t nao, size_t ngrid,
                       size_t nset, const uint8_t *tab, size_t ldtab, size_t blksize) {
    void *p = nullptr;
    const int st = scratch_address(&p);
    if (st != CFGMM_OK) return st;
    const TraceOff no_trace;  // library tracing would create events
    // the first 256 bytes of `scratch` hold the symmetry flag (zero at load, then only epochs)
    unsigned epoch;
    do epoch = ++g_scratch_epoch;
    while (epoch == 0);
    const RhoScratch sc{static_cast<unsigned char *>(p) + 256, kRhoScratchBytes - 256, epoch, static_cast<int *>(p)};
    return launch_rho_multi(stream, rhos, 0, dms, ao_lhs, 0, 1, ao_rhs, stride_ao, stride_grid, nao, ngrid, nset,
                            tab, ldtab, blksize, &sc);
}

}  // namespace cfgmm
