// gto.hip -- AO tabulation on the grid (the step before rho) and the per-AO block mask.
//
//   phi[(ao_off+c)*stride_grid + g] = dx^lx dy^ly dz^lz sum_k coefs[k] exp(-exps[k] r^2),
//   d = coords[g] - center,  c = the component (lx, ly, lz) of the shell in PySCF's Cartesian
//   order (lx = l..0, then ly = l-lx..0);
//   tab[(ao_off+c)*ldtab + b] = [max over block b and the shell's components of |phi| > cutoff],
//   phi zeroed on the blocks where that is 0.
//
// No counterpart in the reference, which takes phi and the mask as inputs (README.md:49 describes
// the mask as PySCF-non0tab-like; src/impl_aodmao2rho.rs:171-244 reads it per AO row).  One wave
// per (grid block, group of shells): lane = grid point, so every phi row segment is written by
// one coalesced store; the block maximum is a wave reduction (deterministic).  Blocks of up to
// 64 points keep the values in registers between the maximum and the store; longer blocks
// evaluate twice.
#include "internal.hpp"

#include <vector>

namespace cfgmm {
namespace {

constexpr int kShellsPerWg = 8;

template <int L>
__device__ inline void shell_values(double dx, double dy, double dz, double rad, double *v) {
    double px[L + 1], py[L + 1], pz[L + 1];
    px[0] = py[0] = pz[0] = 1.0;
#pragma unroll
    for (int i = 1; i <= L; ++i) {
        px[i] = px[i - 1] * dx;
        py[i] = py[i - 1] * dy;
        pz[i] = pz[i - 1] * dz;
    }
    int c = 0;
#pragma unroll
    for (int lx = L; lx >= 0; --lx)
#pragma unroll
        for (int ly = L - lx; ly >= 0; --ly) v[c++] = px[lx] * py[ly] * pz[L - lx - ly] * rad;
}

__device__ inline double wave_max(double m) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    return m;
}

template <int L>
__device__ void tabulate_shell(const cfgmm_gto_shell &sh, const double *__restrict__ exps,
                               const double *__restrict__ coefs, const double *__restrict__ coords, int g0, int npts,
                               int b, double cutoff, double *__restrict__ phi, long stride_grid,
                               uint8_t *__restrict__ tab, long ldtab) {
    constexpr int NF = (L + 1) * (L + 2) / 2;
    const int lane = threadIdx.x;
    auto values = [&](int p, double *v) {  // point g0 + p < ngrid
        const double *__restrict__ r = coords + 3 * (long)(g0 + p);
        const double dx = r[0] - sh.center[0], dy = r[1] - sh.center[1], dz = r[2] - sh.center[2];
        const double r2 = dx * dx + dy * dy + dz * dz;
        double rad = 0.0;
        for (int k = 0; k < sh.nprim; ++k) rad += coefs[sh.prim_off + k] * exp(-exps[sh.prim_off + k] * r2);
        shell_values<L>(dx, dy, dz, rad, v);
    };
    double *__restrict__ row0 = phi + (long)sh.ao_off * stride_grid + g0;
    bool on;
    if (npts <= 64) {
        double v[NF];
        const bool ok = lane < npts;
        if (ok) values(lane, v);
        double m = 0.0;
#pragma unroll
        for (int c = 0; c < NF; ++c) m = ok ? fmax(m, fabs(v[c])) : m;
        on = wave_max(m) > cutoff;
        if (ok) {
#pragma unroll
            for (int c = 0; c < NF; ++c) row0[(long)c * stride_grid + lane] = on ? v[c] : 0.0;
        }
    } else {
        double m = 0.0;
        for (int p = lane; p < npts; p += 64) {
            double v[NF];
            values(p, v);
#pragma unroll
            for (int c = 0; c < NF; ++c) m = fmax(m, fabs(v[c]));
        }
        on = wave_max(m) > cutoff;
        for (int p = lane; p < npts; p += 64) {
            double v[NF];
            if (on) values(p, v);
#pragma unroll
            for (int c = 0; c < NF; ++c) row0[(long)c * stride_grid + p] = on ? v[c] : 0.0;
        }
    }
    if (lane < NF) tab[(long)(sh.ao_off + lane) * ldtab + b] = on ? 1 : 0;
}

__global__ __launch_bounds__(64) void gto_eval_kernel(const cfgmm_gto_shell *__restrict__ shells, int nshell,
                                                      const double *__restrict__ exps,
                                                      const double *__restrict__ coefs,
                                                      const double *__restrict__ coords, int ngrid, int blksize,
                                                      double cutoff, double *__restrict__ phi, long stride_grid,
                                                      uint8_t *__restrict__ tab, long ldtab) {
    const int b = blockIdx.x;
    const int g0 = b * blksize, npts = min(blksize, ngrid - g0);
    for (int si = 0; si < kShellsPerWg; ++si) {
        const int s = blockIdx.y * kShellsPerWg + si;
        if (s >= nshell) return;
        const cfgmm_gto_shell sh = shells[s];
        switch (sh.l) {
            case 0: tabulate_shell<0>(sh, exps, coefs, coords, g0, npts, b, cutoff, phi, stride_grid, tab, ldtab); break;
            case 1: tabulate_shell<1>(sh, exps, coefs, coords, g0, npts, b, cutoff, phi, stride_grid, tab, ldtab); break;
            case 2: tabulate_shell<2>(sh, exps, coefs, coords, g0, npts, b, cutoff, phi, stride_grid, tab, ldtab); break;
            case 3: tabulate_shell<3>(sh, exps, coefs, coords, g0, npts, b, cutoff, phi, stride_grid, tab, ldtab); break;
            default: tabulate_shell<4>(sh, exps, coefs, coords, g0, npts, b, cutoff, phi, stride_grid, tab, ldtab); break;
        }
    }
}

}  // namespace

int launch_gto(hipStream_t stream, double *phi, size_t stride_grid, uint8_t *tab, size_t ldtab,
               const cfgmm_gto_shell *shells, size_t nshell, const double *exps, const double *coefs,
               size_t nprim_total, const double *coords, size_t ngrid, size_t blksize, double cutoff) {
    if (nshell == 0 || ngrid == 0) return CFGMM_OK;
    const size_t nblk = div_up(ngrid, blksize), ngrp = div_up(nshell, kShellsPerWg);
    if (ngrp > 65535) return fail(CFGMM_EINVAL, "gto: too many shells for one launch");
    auto al = [](size_t x) { return div_up(x, 256) * 256; };
    const size_t o_exps = al(sizeof(cfgmm_gto_shell) * nshell);
    const size_t o_coefs = o_exps + al(sizeof(double) * std::max<size_t>(nprim_total, 1));
    void *ws = nullptr;
    int st = workspace_get(stream, o_coefs + sizeof(double) * std::max<size_t>(nprim_total, 1), &ws);
    if (st != CFGMM_OK) return st;
    char *base = static_cast<char *>(ws);
    auto *d_shells = reinterpret_cast<cfgmm_gto_shell *>(base);
    auto *d_exps = reinterpret_cast<double *>(base + o_exps);
    auto *d_coefs = reinterpret_cast<double *>(base + o_coefs);
    CFGMM_HIP_TRY(hipMemcpyAsync(d_shells, shells, sizeof(cfgmm_gto_shell) * nshell, hipMemcpyHostToDevice, stream));
    if (nprim_total) {
        CFGMM_HIP_TRY(hipMemcpyAsync(d_exps, exps, sizeof(double) * nprim_total, hipMemcpyHostToDevice, stream));
        CFGMM_HIP_TRY(hipMemcpyAsync(d_coefs, coefs, sizeof(double) * nprim_total, hipMemcpyHostToDevice, stream));
    }
    hipEvent_t t0 = trace_start(stream);
    hipLaunchKernelGGL(gto_eval_kernel, dim3((unsigned)nblk, (unsigned)ngrp), dim3(64), 0, stream, d_shells,
                       (int)nshell, d_exps, d_coefs, coords, (int)ngrid, (int)blksize, cutoff, phi,
                       (long)stride_grid, tab, (long)ldtab);
    trace_stop(stream, "gto_eval_kernel", t0);
    st = hip_status(hipGetLastError(), "gto_eval_kernel launch");
    if (st != CFGMM_OK) return st;
    return hip_status(hipStreamSynchronize(stream), "gto: sync");  // the host tables outlive their copies
}

}  // namespace cfgmm
