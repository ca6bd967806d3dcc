// vxc.hip -- the masked AO-grid-AO contraction that follows rho in a DFT step (FP64 MFMA).
//
//   V[u*ldv + v] += sum_b sum_{g in block b, g < ngrid} [u in L_b][v in L_b] lhs[u][g] w[g] rhs[v][g],
//   L_b = { x < nao : tab[x*ldtab + b] != 0 }.
//
// No counterpart in the reference (it has the density side, aodmao2rho); the mask, strides and
// block semantics are those of src/impl_aodmao2rho.rs:171-244, read the other way round: the
// contraction runs over grid points instead of AOs.  Deterministic, in two main launches:
//   1. vxc_block_kernel: for each grid block b, the compacted block product
//        S_b[i][j] = sum_{g in b} lhs[L_b[i]][g] w[g] rhs[L_b[j]][g],   i, j < |L_b|,
//      by FP64 MFMA, each workgroup a strip of up to twelve 64x64 tiles of one row of tiles (the
//      row operand in registers, the next tile's column operand loaded during the current
//      tile's MFMAs); S_b lives in a workspace of sum_b kpad_b^2 doubles (kpad_b = |L_b|
//      rounded up to 64).  The strips of one block run back to back on one XCD.
//      With lhs and rhs one array S_b and V's increment X are symmetric: only the tiles
//      (ti, tj >= ti) of S_b are computed and stored.
//   2. vxc_reduce_kernel: each wave owns one row u of V over a 256-column tile and streams, for
//      the blocks in increasing order where u is active, the contiguous segment of row
//      pos_b(u) of S_b whose columns L_b[k] fall in the tile, adding it into an LDS copy of
//      the row (pos_b(x) = index of x in L_b).  Every element of S_b is read once, by full
//      lanes; no float atomics, so the result does not depend on scheduling.  Symmetric form:
//      the row takes only columns v >= u (list positions from pos_b(u) on, all in stored tiles),
//      adds them into V and writes those v > u into X; vxc_mirror_kernel then adds X^T into the
//      lower triangle of V through LDS tiles.
#include "internal.hpp"

#include <vector>

namespace cfgmm {
namespace {

constexpr int kThreads = 256;
constexpr int kT = 64;      // tile edge of S_b
constexpr int kKC = 64;     // grid points staged per LDS pass of the block kernel
constexpr int kStrip = 12;  // S_b tiles per block-kernel workgroup (at most)
static_assert(kStrip * kT % kThreads == 0, "column indexes of a strip: whole passes of the workgroup");
constexpr int kRW = 4;      // V rows per reduce workgroup (one per wave)
constexpr int kRC = 256;    // V columns per reduce tile
constexpr int kPH = 4;      // active blocks whose segments are loaded together in the reduction
constexpr int kSeg = kRC / 64;  // 64-wide pieces of one S_b row segment

// pos[b*nao + x] = index of x in L_b, or -1.
__global__ void vxc_positions_kernel(const int *__restrict__ lists, const int *__restrict__ counts, int nao, int nblk,
                                     int *__restrict__ pos) {
    const int b = blockIdx.x;  // one workgroup per block: it owns the whole row of pos
    if (b >= nblk) return;
    int *__restrict__ pb = pos + (long)b * nao;
    for (int x = threadIdx.x; x < nao; x += blockDim.x) pb[x] = -1;
    __syncthreads();
    const int n = counts[b];
    const int *__restrict__ lb = lists + (long)b * nao;
    for (int i = threadIdx.x; i < n; i += blockDim.x) pb[lb[i]] = i;
}

// kr[b*(nct+1) + c] = #{ i < |L_b| : L_b[i] < c*kRC },  c = 0..nct  (binary search of the sorted list).
__global__ void vxc_ranges_kernel(const int *__restrict__ lists, const int *__restrict__ counts, int nao, int nblk,
                                  int nct, int *__restrict__ kr) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)nblk * (nct + 1)) return;
    const int b = (int)(idx / (nct + 1)), c = (int)(idx % (nct + 1));
    const int *__restrict__ lb = lists + (long)b * nao;
    const long key = (long)c * kRC;
    int lo = 0, hi = counts[b];
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (lb[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    kr[idx] = lo;
}

// A strip of S_b tiles: task {b, ti, tj0, ntj} -> tiles (ti, tj0 .. tj0+ntj-1).  Wave w
// computes rows 32*(w/2) .. +31 and columns 32*(w%2) .. +31 of each tile.
//
// Blocks of 4m <= 64 points: on gfx950 every VALU instruction takes issue cycles from the MFMA
// pipe of its SIMD, so the per-tile loop keeps VALU work out: row bases are uniform (scalar
// unit) with one per-lane offset for all loads; LDS and store addresses are per-strip bases
// plus immediates; the first MFMA of a tile takes a zero accumulator; rows outside L_b are
// zeroed by uniform branches.  The waves keep their A fragments (lhs*w) in registers for the
// whole strip, and tile t+1's column rows are loaded during tile t's MFMAs.  Other blocks
// (longer, or a partial last block of 4m+1..4m+3 points) take 32-point passes with both
// operands through LDS.
__global__ __launch_bounds__(kThreads, 3) void vxc_block_kernel(
    const int4 *__restrict__ tasks, const long *__restrict__ s_off, const int *__restrict__ kpads,
    const int *__restrict__ lists, const int *__restrict__ counts, const double *__restrict__ lhs,
    const double *__restrict__ rhs, const double *__restrict__ w, long stride_grid, int nao, int ngrid, int blksize,
    double *__restrict__ S) {
    // One 64-point x 64-row tile [point][row] (33 KB, so that three workgroups fit a CU); the
    // general path splits it into two 32-point tiles, A and B of a pass.
    __shared__ double smem[kKC * (kT + 1)];
    double(*__restrict__ T1)[kT + 1] = reinterpret_cast<double(*)[kT + 1]>(smem);
    __shared__ int ui[kT], vi[kStrip * kT];
    // workgroup i takes task x*q + min(x, r) + i/8 (x = i%8): each XCD a contiguous run of the
    // block-major task order
    const int ntasks = gridDim.x, xcd = blockIdx.x % 8, q = ntasks / 8, rr = ntasks % 8;
    const int4 t = tasks[xcd * q + min(xcd, rr) + blockIdx.x / 8];
    const int b = t.x, ti = t.y, tj0 = t.z, ntj = t.w;
    const int n = counts[b], kpad = kpads[b];
    const int *__restrict__ lb = lists + (long)b * nao;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lrow = lane & 15, lk = lane >> 4;
    if (tid < kT) {
        const int i = ti * kT + tid;
        ui[tid] = i < n ? lb[i] : -1;
    }
#pragma unroll
    for (int e = 0; e < kStrip * kT / kThreads; ++e) {
        const int jj = e * kThreads + tid, j = tj0 * kT + jj;
        vi[jj] = (jj < ntj * kT && j < n) ? lb[j] : -1;
    }
    __syncthreads();
    const int g0 = b * blksize, npts = min(blksize, ngrid - g0);
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    double *__restrict__ Sb = S + s_off[b];

    if (npts <= kKC && npts % 4 == 0) {
        const int nkk = npts / 4;
        // wave w stages operand rows 16w .. 16w+15 (lane = point); lanes past the block read
        // its last point (in bounds) and those points are never used: the MFMA steps stop at
        // npts
        const unsigned loff = (unsigned)min(lane, npts - 1) * 8u;
        const double wg = w ? w[g0 + min(lane, npts - 1)] : 1.0;
        auto load16 = [&](const double *__restrict__ src, const int *idx, double *r, int *rows) {
#pragma unroll
            for (int q4 = 0; q4 < 16; ++q4) {
                rows[q4] = __builtin_amdgcn_readfirstlane(idx[q4]);
                const char *base = (const char *)(src + (long)max(rows[q4], 0) * stride_grid + g0);
                r[q4] = *(const double *)(base + loff);
            }
        };
        auto put16 = [&](const double *r, const int *rows, bool scale) {
#pragma unroll
            for (int q4 = 0; q4 < 16; ++q4) {
                double x = 0.0;
                if (rows[q4] >= 0) x = scale ? r[q4] * wg : r[q4];
                T1[lane][wave * 16 + q4] = x;
            }
        };
        double rg[16];
        int rows[16];
        load16(lhs, ui + wave * 16, rg, rows);
        put16(rg, rows, true);  // A through the tile first
        load16(rhs, vi + wave * 16, rg, rows);
        __syncthreads();
        double fa[kKC / 4][2];
#pragma unroll
        for (int kk = 0; kk < kKC / 4; ++kk)
#pragma unroll
            for (int a = 0; a < 2; ++a) fa[kk][a] = T1[4 * kk + lk][wr + a * 16 + lrow];
        __syncthreads();  // every wave holds its A fragments before the tile is reused
        put16(rg, rows, false);
        __syncthreads();
        // per-lane store offsets (bytes) of D[l/16 + 4r][l%16] for the a-blocks; +128 B per c
        unsigned soff[2][4];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int r = 0; r < 4; ++r) soff[a][r] = (unsigned)(((wr + a * 16 + lk + 4 * r) * kpad + wc + lrow) * 8);
        char *const srow = (char *)(Sb + (long)ti * kT * kpad);
        // kq = 4-point steps to run; the full-block call is inlined with kq = 16 (no guards)
        auto mfma_tile = [&](int kq, f64x4 (&acc)[2][2]) {
#pragma unroll
            for (int kk = 0; kk < kKC / 4; ++kk) {
                if (kk < kq) {
                    double fb[2];
#pragma unroll
                    for (int c = 0; c < 2; ++c) fb[c] = T1[4 * kk + lk][wc + c * 16 + lrow];
#pragma unroll
                    for (int a = 0; a < 2; ++a)
#pragma unroll
                        for (int c = 0; c < 2; ++c)
                            acc[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                                fa[kk][a], fb[c], kk == 0 ? f64x4{0.0, 0.0, 0.0, 0.0} : acc[a][c], 0, 0, 0);
                }
            }
        };
        for (int jt = 0; jt < ntj; ++jt) {
            if (jt + 1 < ntj) load16(rhs, vi + (jt + 1) * kT + wave * 16, rg, rows);  // during the MFMAs
            f64x4 acc[2][2];
            if (nkk == kKC / 4) mfma_tile(kKC / 4, acc);
            else mfma_tile(nkk, acc);
            char *const st = srow + (long)(tj0 + jt) * kT * 8;
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int c = 0; c < 2; ++c)
#pragma unroll
                    for (int r = 0; r < 4; ++r) *(double *)(st + soff[a][r] + c * 128) = acc[a][c][r];
            if (jt + 1 < ntj) {
                __syncthreads();  // every wave has read tile jt
                put16(rg, rows, false);
                __syncthreads();
            }
        }
        return;
    }

    // general blocks: step = (tile jt, 32-point pass p), both operands through LDS (A in the
    // tile's first 32 point rows, B in the last 32); rows outside L_b and points past the
    // block stage as zero
    constexpr int kHP = kKC / 2;
    double(*__restrict__ As)[kT + 1] = T1;
    double(*__restrict__ Bs)[kT + 1] = T1 + kHP;
    const int npass = (npts + kHP - 1) / kHP, nsteps = ntj * npass;
    f64x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c) acc[a][c] = f64x4{0.0, 0.0, 0.0, 0.0};
    for (int step = 0; step < nsteps; ++step) {
        const int jt = step / npass, pass = step % npass, k0 = pass * kHP;
        // lanes 0..31 stage point k0 + lane of A, lanes 32..63 the same point of B
        const int pt = lane & (kHP - 1);
        const bool isb = lane >= kHP;
        const bool pok = k0 + pt < npts;
        const int g = pok ? g0 + k0 + pt : g0;
        const double wg = isb ? 1.0 : (w ? w[g] : 1.0);
        if (step > 0) __syncthreads();  // the previous step has finished reading both halves
#pragma unroll
        for (int q4 = 0; q4 < kT / 4; ++q4) {
            const int x = isb ? vi[jt * kT + wave + 4 * q4] : ui[wave + 4 * q4];
            const double *__restrict__ src = isb ? rhs : lhs;
            T1[lane][wave + 4 * q4] = (pok && x >= 0) ? src[(long)x * stride_grid + g] * wg : 0.0;
        }
        __syncthreads();
        const int kend = min(kHP, npts - k0);  // zero-padded up to a multiple of 4
        for (int kk = 0; kk < kend; kk += 4) {
            double fa[2], fb[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) fa[a] = As[kk + lk][wr + a * 16 + lrow];
#pragma unroll
            for (int c = 0; c < 2; ++c) fb[c] = Bs[kk + lk][wc + c * 16 + lrow];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int c = 0; c < 2; ++c)
                    acc[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[a], fb[c], acc[a][c], 0, 0, 0);
        }
        if (pass == npass - 1) {
            const int tj = tj0 + jt;
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int c = 0; c < 2; ++c) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int i = ti * kT + wr + a * 16 + lk + 4 * r;  // D[l/16 + 4r][l%16]
                        const int j = tj * kT + wc + c * 16 + lrow;
                        Sb[(long)i * kpad + j] = acc[a][c][r];  // i, j < kpad: the padding is written too
                    }
                    acc[a][c] = f64x4{0.0, 0.0, 0.0, 0.0};
                }
        }
    }
}

// Rows 4*blockIdx.x .. +3 of V (one per wave) over columns v0 .. v0+255, v0 = 256 * blockIdx.y.
__global__ __launch_bounds__(kThreads) void vxc_reduce_kernel(const int *__restrict__ pos,
                                                               const int *__restrict__ lists,
                                                               const int *__restrict__ kr,
                                                               const long *__restrict__ s_off,
                                                               const int *__restrict__ kpads, int nao, int nblk,
                                                               int nct, const double *__restrict__ S,
                                                               double *__restrict__ V, long ldv, int sym,
                                                               double *__restrict__ X) {
    __shared__ double Vt[kRW][kRC];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ct = blockIdx.y, v0 = ct * kRC, u = blockIdx.x * kRW + wave;
    double *__restrict__ row = Vt[wave];  // private to this wave: no barriers
    if (u >= nao) return;
    if (sym && v0 + kRC <= u) return;  // below the diagonal: vxc_mirror_kernel
#pragma unroll
    for (int c = 0; c < kSeg; ++c) row[c * 64 + lane] = 0.0;
    for (int c0 = 0; c0 < nblk; c0 += 64) {
        // lane l describes block c0 + l: the row's position in L_b and the segment [klo, khi)
        // of L_b whose columns lie in this tile (clamped indexes: every load is in bounds)
        const bool ok = c0 + lane < nblk;
        const int bq = ok ? c0 + lane : 0;
        int iu = pos[(long)bq * nao + u];
        const int khi = kr[(long)bq * (nct + 1) + ct + 1];
        // symmetric form: columns v >= u only, the list positions from u's own on
        const int klo = sym ? max(kr[(long)bq * (nct + 1) + ct], iu) : kr[(long)bq * (nct + 1) + ct];
        const long so = s_off[bq];
        const int kp = kpads[bq];
        iu = ok ? iu : -1;
        unsigned long long act = __ballot(iu >= 0 && khi > klo);
        while (act) {  // kPH active blocks per step, in increasing order
            int ls[kPH];
#pragma unroll
            for (int h = 0; h < kPH; ++h) {
                ls[h] = act ? __builtin_ctzll(act) : -1;
                act &= act - 1;  // (0 stays 0)
            }
            double xs[kPH][kSeg];
            int xc[kPH][kSeg];
#pragma unroll
            for (int h = 0; h < kPH; ++h) {
#pragma unroll
                for (int c = 0; c < kSeg; ++c) {
                    xs[h][c] = 0.0;
                    xc[h][c] = -1;
                }
                const int l = ls[h];
                if (l < 0) continue;
                const int iuh = __builtin_amdgcn_readlane(iu, l);
                const int klh = __builtin_amdgcn_readlane(klo, l), khh = __builtin_amdgcn_readlane(khi, l);
                const int kph = __builtin_amdgcn_readlane(kp, l);
                const long soh = (long)(((unsigned long)(unsigned)__builtin_amdgcn_readlane((int)(so >> 32), l) << 32) |
                                        (unsigned)__builtin_amdgcn_readlane((int)so, l));
                const double *__restrict__ src = S + soh + (long)iuh * kph;
                const int *__restrict__ lst = lists + (long)(c0 + l) * nao;
#pragma unroll
                for (int c = 0; c < kSeg; ++c) {
                    if (klh + 64 * c < khh) {
                        const int k = klh + 64 * c + lane;
                        if (k < khh) {  // raw loads only: the wait comes at the first use below
                            xs[h][c] = src[k];
                            xc[h][c] = lst[k];
                        }
                    }
                }
            }
#pragma unroll
            for (int h = 0; h < kPH; ++h)
#pragma unroll
                for (int c = 0; c < kSeg; ++c)
                    if (xc[h][c] >= 0)  // ds_add_f64: only this wave touches the row, in program order
                        __hip_atomic_fetch_add(&row[xc[h][c] - v0], xs[h][c], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WAVEFRONT);
        }
    }
    double *__restrict__ vrow = V + (long)u * ldv;
#pragma unroll
    for (int c = 0; c < kSeg; ++c) {
        const int v = v0 + c * 64 + lane;
        if (v < nao && (!sym || v >= u)) vrow[v] += row[c * 64 + lane];
        if (sym && v < nao && v > u) X[(long)u * nao + v] = row[c * 64 + lane];
    }
}

// Symmetric form: V[v][u] += X[u][v] for u < v.  Workgroup (J, I), J <= I: the tile of rows
// u in J, columns v in I of X through LDS into rows v in I, columns u in J of V.
__global__ __launch_bounds__(kThreads) void vxc_mirror_kernel(const double *__restrict__ X, double *__restrict__ V,
                                                              long ldv, int nao) {
    const int J = blockIdx.x, I = blockIdx.y;
    if (J > I) return;
    __shared__ double t[kT][kT + 1];
    const int c = threadIdx.x & 63, r0 = threadIdx.x >> 6;
    for (int r = r0; r < kT; r += kThreads / 64) {
        const int u = J * kT + r, v = I * kT + c;
        t[r][c] = (u < nao && v < nao && v > u) ? X[(long)u * nao + v] : 0.0;
    }
    __syncthreads();
    for (int r = r0; r < kT; r += kThreads / 64) {
        const int v = I * kT + r, u = J * kT + c;
        if (u < nao && v < nao && v > u) V[(long)v * ldv + u] += t[c][r];
    }
}

}  // namespace

int launch_vxc(hipStream_t stream, double *V, size_t ldv, const double *ao_lhs, const double *ao_rhs, const double *w,
               size_t stride_grid, size_t nao, size_t ngrid, const uint8_t *tab, size_t ldtab, size_t blksize) {
    if (nao == 0 || ngrid == 0) return CFGMM_OK;
    if (nao > INT32_MAX || ngrid > INT32_MAX) return fail(CFGMM_EINVAL, "vxc: size out of range");
    const size_t nblk = div_up(ngrid, blksize), nct = div_up(nao, kRC);
    auto al = [](size_t x) { return div_up(x, 256) * 256; };
    // workspace: lists | counts | pos | kr ; then the tables and S (sizes need the counts)
    const size_t o_counts = al(sizeof(int) * nblk * nao);
    const size_t o_pos = o_counts + al(sizeof(int) * nblk);
    const size_t o_kr = o_pos + al(sizeof(int) * nblk * nao);
    const size_t head = o_kr + al(sizeof(int) * nblk * (nct + 1));
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
    // S_b offsets and the strip tasks, block-major (block, tile row, strip); one array on both
    // sides makes S_b symmetric
    const int sym = ao_lhs == ao_rhs && tune_get(CFGMM_TUNE_RHO_SYMMETRIC) == 0 ? 1 : 0;
    std::vector<long> soff(nblk);
    std::vector<int> kp(nblk);
    std::vector<int4> tasks;
    size_t s_elems = 0;
    for (size_t b = 0; b < nblk; ++b) {
        kp[b] = (int)(div_up((size_t)hc[b], kT) * kT);
        soff[b] = (long)s_elems;
        s_elems += (size_t)kp[b] * kp[b];
        const int nt = kp[b] / kT;
        for (int i = 0; i < nt; ++i) {
            // tiles j0 .. nt-1 of row i (the symmetric form: j0 = i), strips of near-equal length
            const int first = sym ? i : 0, len = nt - first;
            const int ns = (len + kStrip - 1) / kStrip;
            for (int q = 0; q < ns; ++q) {
                const int j0 = first + q * len / ns, j1 = first + (q + 1) * len / ns;
                tasks.push_back(int4{(int)b, i, j0, j1 - j0});
            }
        }
    }
    if (tasks.size() > (size_t)INT32_MAX) return fail(CFGMM_EINVAL, "vxc: too many tiles");
    const size_t o_soff = head;
    const size_t o_kp = o_soff + al(sizeof(long) * nblk);
    const size_t o_tasks = o_kp + al(sizeof(int) * nblk);
    const size_t o_S = o_tasks + al(sizeof(int4) * std::max<size_t>(tasks.size(), 1));
    const size_t o_X = o_S + al(sizeof(double) * std::max<size_t>(s_elems, 1));  // symmetric form: X, nao^2
    void *const ws_lists = ws;
    st = workspace_get(stream, o_X + (sym ? sizeof(double) * nao * nao : 0), &ws);  // may move: re-derive
    if (st != CFGMM_OK) return st;
    char *base = static_cast<char *>(ws);
    lists = reinterpret_cast<int *>(base);
    counts = reinterpret_cast<int *>(base + o_counts);
    int *pos = reinterpret_cast<int *>(base + o_pos);
    int *kr = reinterpret_cast<int *>(base + o_kr);
    long *d_soff = reinterpret_cast<long *>(base + o_soff);
    int *d_kp = reinterpret_cast<int *>(base + o_kp);
    int4 *d_tasks = reinterpret_cast<int4 *>(base + o_tasks);
    double *S = reinterpret_cast<double *>(base + o_S);
    double *X = sym ? reinterpret_cast<double *>(base + o_X) : nullptr;
    if (ws != ws_lists) {  // a grown workspace no longer holds the lists: build them again
        st = launch_build_lists(stream, tab, ldtab, nao, nblk, lists, counts);
        if (st != CFGMM_OK) return st;
    }
    CFGMM_HIP_TRY(hipMemcpyAsync(d_soff, soff.data(), sizeof(long) * nblk, hipMemcpyHostToDevice, stream));
    CFGMM_HIP_TRY(hipMemcpyAsync(d_kp, kp.data(), sizeof(int) * nblk, hipMemcpyHostToDevice, stream));
    if (!tasks.empty())
        CFGMM_HIP_TRY(hipMemcpyAsync(d_tasks, tasks.data(), sizeof(int4) * tasks.size(), hipMemcpyHostToDevice, stream));
    hipEvent_t t0 = trace_start(stream);
    hipLaunchKernelGGL(vxc_positions_kernel, dim3((unsigned)nblk), dim3(kThreads), 0, stream, lists, counts,
                       (int)nao, (int)nblk, pos);
    const size_t nkr = nblk * (nct + 1);
    hipLaunchKernelGGL(vxc_ranges_kernel, dim3((unsigned)div_up(nkr, kThreads)), dim3(kThreads), 0, stream, lists,
                       counts, (int)nao, (int)nblk, (int)nct, kr);
    if (!tasks.empty())
        hipLaunchKernelGGL(vxc_block_kernel, dim3((unsigned)tasks.size()), dim3(kThreads), 0, stream, d_tasks, d_soff,
                           d_kp, lists, counts, ao_lhs, ao_rhs, w, (long)stride_grid, (int)nao, (int)ngrid,
                           (int)blksize, S);
    trace_stop(stream, "vxc_block_kernel", t0);
    st = hip_status(hipGetLastError(), "vxc_block_kernel launch");
    if (st != CFGMM_OK) return st;
    t0 = trace_start(stream);
    hipLaunchKernelGGL(vxc_reduce_kernel, dim3((unsigned)div_up(nao, kRW), (unsigned)nct), dim3(kThreads), 0, stream,
                       pos, lists, kr, d_soff, d_kp, (int)nao, (int)nblk, (int)nct, S, V, (long)ldv, sym, X);
    if (sym) {
        const unsigned nt64 = (unsigned)div_up(nao, kT);
        hipLaunchKernelGGL(vxc_mirror_kernel, dim3(nt64, nt64), dim3(kThreads), 0, stream, X, V, (long)ldv, (int)nao);
    }
    trace_stop(stream, "vxc_reduce_kernel", t0);
    st = hip_status(hipGetLastError(), "vxc_reduce_kernel launch");
    if (st != CFGMM_OK) return st;
    // the host tables above must outlive their copies
    return hip_status(hipStreamSynchronize(stream), "vxc: sync");
}

}  // namespace cfgmm
