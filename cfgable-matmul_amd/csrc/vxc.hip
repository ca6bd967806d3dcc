// vxc.hip -- the masked AO-grid-AO contraction that follows rho in a DFT step (FP64 MFMA).
//
//   V[u*ldv + v] += sum_b sum_{g in block b, g < ngrid} [u in L_b][v in L_b] lhs[u][g] w[g] rhs[v][g],
//   L_b = { x < nao : tab[x*ldtab + b] != 0 }.
//
// No counterpart in the reference (it has the density side, aodmao2rho); the mask, strides and
// block semantics are those of src/impl_aodmao2rho.rs:171-244, read the other way round: the
// contraction runs over grid points instead of AOs.  Deterministic, in two launches:
//   1. vxc_block_kernel: for each grid block b, the compacted block product
//        S_b[i][j] = sum_{g in b} lhs[L_b[i]][g] w[g] rhs[L_b[j]][g],   i, j < |L_b|,
//      one 64x64 tile per workgroup (FP64 MFMA over the block's points, K <= blksize), stored
//      in a workspace of sum_b kpad_b^2 doubles (kpad_b = |L_b| rounded up to 64).
//   2. vxc_reduce_kernel: each 64x64 tile of V sums, over the blocks in increasing order, the
//      entries S_b[pos_b(u)][pos_b(v)] of the (u, v) active in block b (pos_b(x) = index of x
//      in L_b, or -1), then adds the sum into V.  No float atomics: the result does not depend
//      on scheduling.
#include "internal.hpp"

#include <vector>

namespace cfgmm {
namespace {

constexpr int kThreads = 256;
constexpr int kT = 64;   // output tile edge (rows of S_b / of V)
constexpr int kKB = 16;  // grid points per LDS slice

// pos[b*nao + x] = index of x in L_b, or -1.
__global__ void vxc_positions_kernel(const int *__restrict__ lists, const int *__restrict__ counts, int nao, int nblk,
                                     int *__restrict__ pos) {
    const int b = blockIdx.y;
    if (b >= nblk) return;
    int *__restrict__ pb = pos + (long)b * nao;
    for (int x = blockIdx.x * blockDim.x + threadIdx.x; x < nao; x += gridDim.x * blockDim.x) pb[x] = -1;
    __syncthreads();  // the whole row belongs to this workgroup when gridDim.x == 1
    const int n = counts[b];
    const int *__restrict__ lb = lists + (long)b * nao;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) pb[lb[i]] = i;
}

// One 64x64 tile (ti, tj) of S_b.  Task: x = block, y = ti, z = tj; w unused.
__global__ __launch_bounds__(kThreads, 2) void vxc_block_kernel(
    const int4 *__restrict__ tasks, const long *__restrict__ s_off, const int *__restrict__ kpads,
    const int *__restrict__ lists, const int *__restrict__ counts, const double *__restrict__ lhs,
    const double *__restrict__ rhs, const double *__restrict__ w, long stride_grid, int nao, int ngrid, int blksize,
    double *__restrict__ S) {
    __shared__ double As[kKB][kT + 1];  // lhs*w [point][row]
    __shared__ double Bs[kKB][kT + 1];  // rhs   [point][col]
    __shared__ int ui[kT], vi[kT];
    const int4 t = tasks[blockIdx.x];
    const int b = t.x, ti = t.y, tj = t.z;
    const int n = counts[b], kpad = kpads[b];
    const int *__restrict__ lb = lists + (long)b * nao;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lrow = lane & 15, lk = lane >> 4;
    if (tid < kT) {
        const int i = ti * kT + tid, j = tj * kT + tid;
        ui[tid] = i < n ? lb[i] : -1;
        vi[tid] = j < n ? lb[j] : -1;
    }
    __syncthreads();
    const int g0 = b * blksize, npts = min(blksize, ngrid - g0);
    // wave w computes rows 32*(w/2) .. +31 and columns 32*(w%2) .. +31 of the tile
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    f64x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c) acc[a][c] = f64x4{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < npts; k0 += kKB) {
        // stage 64 rows x 16 points of lhs*w and of rhs (rows outside L_b and points past the
        // block are zero)
        for (int idx = tid; idx < kT * kKB; idx += kThreads) {
            const int r = idx / kKB, p = idx % kKB;
            const int g = g0 + k0 + p;
            const bool pok = k0 + p < npts;
            const int u = ui[r], v = vi[r];
            const double wg = pok ? (w ? w[g] : 1.0) : 0.0;
            As[p][r] = (pok && u >= 0) ? lhs[(long)u * stride_grid + g] * wg : 0.0;
            Bs[p][r] = (pok && v >= 0) ? rhs[(long)v * stride_grid + g] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kKB / 4; ++kk) {
            double fa[2], fb[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) fa[a] = As[kk * 4 + lk][wr + a * 16 + lrow];
#pragma unroll
            for (int c = 0; c < 2; ++c) fb[c] = Bs[kk * 4 + lk][wc + c * 16 + lrow];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int c = 0; c < 2; ++c)
                    acc[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[a], fb[c], acc[a][c], 0, 0, 0);
        }
        __syncthreads();
    }
    double *__restrict__ Sb = S + s_off[b];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = ti * kT + wr + a * 16 + lk + 4 * r;  // D[l/16 + 4r][l%16]
                const int j = tj * kT + wc + c * 16 + lrow;
                Sb[(long)i * kpad + j] = acc[a][c][r];  // i, j < kpad: the padding is written too
            }
}

// One 64x64 tile of V: V[u][v] += sum over blocks, in increasing order, of S_b[pos_b(u)][pos_b(v)].
__global__ __launch_bounds__(kThreads) void vxc_reduce_kernel(const int *__restrict__ pos,
                                                               const long *__restrict__ s_off,
                                                               const int *__restrict__ kpads, int nao, int nblk,
                                                               const double *__restrict__ S, double *__restrict__ V,
                                                               long ldv) {
    __shared__ int pu[kT], pv[kT];
    const int u0 = blockIdx.y * kT, v0 = blockIdx.x * kT;
    const int tid = threadIdx.x;
    // thread -> column v0 + tid%64, rows u0 + tid/64 + 4e (e < 16)
    const int cv = tid % kT, r0 = tid / kT;
    double acc[kT / 4];
#pragma unroll
    for (int e = 0; e < kT / 4; ++e) acc[e] = 0.0;
    for (int b = 0; b < nblk; ++b) {
        __syncthreads();
        if (tid < kT) pu[tid] = u0 + tid < nao ? pos[(long)b * nao + u0 + tid] : -1;
        else if (tid < 2 * kT) pv[tid - kT] = v0 + tid - kT < nao ? pos[(long)b * nao + v0 + tid - kT] : -1;
        __syncthreads();
        const int jv = pv[cv];
        if (jv < 0) continue;
        const double *__restrict__ Sb = S + s_off[b];
        const int kpad = kpads[b];
#pragma unroll
        for (int e = 0; e < kT / 4; ++e) {
            const int iu = pu[r0 + 4 * e];
            if (iu >= 0) acc[e] += Sb[(long)iu * kpad + jv];
        }
    }
    const int v = v0 + cv;
    if (v >= nao) return;
#pragma unroll
    for (int e = 0; e < kT / 4; ++e) {
        const int u = u0 + r0 + 4 * e;
        if (u < nao) V[(long)u * ldv + v] += acc[e];
    }
}

}  // namespace

int launch_vxc(hipStream_t stream, double *V, size_t ldv, const double *ao_lhs, const double *ao_rhs, const double *w,
               size_t stride_grid, size_t nao, size_t ngrid, const uint8_t *tab, size_t ldtab, size_t blksize) {
    if (nao == 0 || ngrid == 0) return CFGMM_OK;
    if (nao > INT32_MAX || ngrid > INT32_MAX) return fail(CFGMM_EINVAL, "vxc: size out of range");
    const size_t nblk = div_up(ngrid, blksize);
    auto al = [](size_t x) { return div_up(x, 256) * 256; };
    // workspace: lists | counts | pos ; then the tables and S (sizes need the counts)
    const size_t o_counts = al(sizeof(int) * nblk * nao);
    const size_t o_pos = o_counts + al(sizeof(int) * nblk);
    const size_t head = o_pos + al(sizeof(int) * nblk * nao);
    void *ws = nullptr;
    int st = workspace_get(stream, head, &ws);
    if (st != CFGMM_OK) return st;
    int *lists = static_cast<int *>(ws);
    int *counts = reinterpret_cast<int *>(static_cast<char *>(ws) + o_counts);
    st = launch_build_lists(stream, tab, ldtab, nao, nblk, lists, counts);
    if (st != CFGMM_OK) return st;
    std::vector<int> hc(nblk);
    CFGMM_HIP_TRY(hipMemcpyAsync(hc.data(), counts, sizeof(int) * nblk, hipMemcpyDeviceToHost, stream));
    CFGMM_HIP_TRY(hipStreamSynchronize(stream));
    std::vector<long> soff(nblk);
    std::vector<int> kp(nblk);
    std::vector<int4> tasks;
    size_t s_elems = 0;
    for (size_t b = 0; b < nblk; ++b) {
        kp[b] = (int)(div_up((size_t)hc[b], kT) * kT);
        soff[b] = (long)s_elems;
        s_elems += (size_t)kp[b] * kp[b];
        const int nt = kp[b] / kT;
        for (int i = 0; i < nt; ++i)
            for (int j = 0; j < nt; ++j) tasks.push_back(int4{(int)b, i, j, 0});
    }
    if (tasks.size() > (size_t)INT32_MAX) return fail(CFGMM_EINVAL, "vxc: too many tiles");
    const size_t o_soff = head;
    const size_t o_kp = o_soff + al(sizeof(long) * nblk);
    const size_t o_tasks = o_kp + al(sizeof(int) * nblk);
    const size_t o_S = o_tasks + al(sizeof(int4) * std::max<size_t>(tasks.size(), 1));
    st = workspace_get(stream, o_S + sizeof(double) * std::max<size_t>(s_elems, 1), &ws);  // may move: re-derive
    if (st != CFGMM_OK) return st;
    char *base = static_cast<char *>(ws);
    // a grown workspace no longer holds the lists: build them again in the new buffer
    lists = reinterpret_cast<int *>(base);
    counts = reinterpret_cast<int *>(base + o_counts);
    int *pos = reinterpret_cast<int *>(base + o_pos);
    long *d_soff = reinterpret_cast<long *>(base + o_soff);
    int *d_kp = reinterpret_cast<int *>(base + o_kp);
    int4 *d_tasks = reinterpret_cast<int4 *>(base + o_tasks);
    double *S = reinterpret_cast<double *>(base + o_S);
    st = launch_build_lists(stream, tab, ldtab, nao, nblk, lists, counts);
    if (st != CFGMM_OK) return st;
    CFGMM_HIP_TRY(hipMemcpyAsync(d_soff, soff.data(), sizeof(long) * nblk, hipMemcpyHostToDevice, stream));
    CFGMM_HIP_TRY(hipMemcpyAsync(d_kp, kp.data(), sizeof(int) * nblk, hipMemcpyHostToDevice, stream));
    if (!tasks.empty())
        CFGMM_HIP_TRY(hipMemcpyAsync(d_tasks, tasks.data(), sizeof(int4) * tasks.size(), hipMemcpyHostToDevice, stream));
    hipEvent_t t0 = trace_start(stream);
    hipLaunchKernelGGL(vxc_positions_kernel, dim3(1, (unsigned)nblk), dim3(kThreads), 0, stream, lists, counts,
                       (int)nao, (int)nblk, pos);
    if (!tasks.empty())
        hipLaunchKernelGGL(vxc_block_kernel, dim3((unsigned)tasks.size()), dim3(kThreads), 0, stream, d_tasks, d_soff,
                           d_kp, lists, counts, ao_lhs, ao_rhs, w, (long)stride_grid, (int)nao, (int)ngrid,
                           (int)blksize, S);
    trace_stop(stream, "vxc_block_kernel", t0);
    st = hip_status(hipGetLastError(), "vxc_block_kernel launch");
    if (st != CFGMM_OK) return st;
    const unsigned tiles = (unsigned)div_up(nao, kT);
    t0 = trace_start(stream);
    hipLaunchKernelGGL(vxc_reduce_kernel, dim3(tiles, tiles), dim3(kThreads), 0, stream, pos, d_soff, d_kp, (int)nao,
                       (int)nblk, S, V, (long)ldv);
    trace_stop(stream, "vxc_reduce_kernel", t0);
    st = hip_status(hipGetLastError(), "vxc_reduce_kernel launch");
    if (st != CFGMM_OK) return st;
    // the host tables above must outlive their copies
    return hip_status(hipStreamSynchronize(stream), "vxc: sync");
}

}  // namespace cfgmm
