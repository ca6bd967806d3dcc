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
// one coalesced store.  The host packs each group's shells and primitives into one record of
// fixed size, so a workgroup brings all it needs into LDS with one load next to that of its grid
// coordinates.  Screening: from a sphere around the block's points, an upper bound of |phi| per
// shell (sum over primitives of |c| max_{r >= rmin} r^l exp(-a r^2)); a shell whose bound is under
// half the cutoff is written as zeros without evaluating it.  The bound only skips work whose
// result is known (every value under the cutoff, so a zero block), so the output is the same as
// evaluating everything.  Blocks of up to 64 points keep the values in registers between the
// block maximum (a ballot) and the store; longer blocks evaluate twice.
#include "internal.hpp"

#include <cstring>
#include <vector>

namespace cfgmm {
namespace {

constexpr int kShellsPerWg = 8;
constexpr int kPrimCap = 64;  // primitives of one group held in its record

// One workgroup's shells.  prim_off of a shell indexes pe / pc, or, for a group of one shell with
// more than kPrimCap primitives (global != 0), the exps / coefs tables in device memory.
struct alignas(16) ShellGroup {
    int ns, global, pad[2];
    cfgmm_gto_shell sh[kShellsPerWg];
    double pe[kPrimCap], pc[kPrimCap];
};
constexpr int kGroupWords = sizeof(ShellGroup) / 16;
static_assert(sizeof(ShellGroup) % 16 == 0 && kGroupWords > 64 && kGroupWords <= 128, "record load");

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

template <int Op>
__device__ inline double wave_reduce(double x) {  // Op 0: min, 1: max; every lane gets the result
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double y = __shfl_xor(x, off);
        x = Op == 0 ? fmin(x, y) : fmax(x, y);
    }
    return x;
}

// Upper bound of |phi| over the shell's components and the points at distance >= rmin from its
// centre: |dx^lx dy^ly dz^lz| <= r^l, and r^l exp(-a r^2) peaks at r = sqrt(l / 2a).  +inf (no
// screening) for a non-positive exponent.
__device__ inline double shell_bound(int l, int np, const double *__restrict__ pe, const double *__restrict__ pc,
                                     double rmin) {
    double bound = 0.0;
    for (int k = 0; k < np; ++k) {
        const double a = pe[k];
        if (!(a > 0.0)) return INFINITY;
        const double r = l ? fmax(rmin, sqrt(l / (2.0 * a))) : rmin;
        double rl = 1.0;
        for (int i = 0; i < l; ++i) rl *= r;
        bound += fabs(pc[k]) * rl * exp(-a * r * r);
    }
    return bound;
}

template <int L>
__device__ inline void shell_point(const cfgmm_gto_shell &sh, const double *__restrict__ pe,
                                   const double *__restrict__ pc, double px, double py, double pz, double *v) {
    const double dx = px - sh.center[0], dy = py - sh.center[1], dz = pz - sh.center[2];
    const double r2 = dx * dx + dy * dy + dz * dz;
    double rad = 0.0;
    const int np = __builtin_amdgcn_readfirstlane(sh.nprim), po = __builtin_amdgcn_readfirstlane(sh.prim_off);
    for (int k = 0; k < np; ++k) rad += pc[po + k] * exp(-pe[po + k] * r2);
    shell_values<L>(dx, dy, dz, rad, v);
}

// One shell on one block.  kLong: blocks longer than 64 points, whose maximum needs every point
// before the first store, evaluated twice (maximum, then the stores).
template <int L, bool kLong>
__device__ void tabulate_shell(const cfgmm_gto_shell &sh, const double *__restrict__ pe,
                               const double *__restrict__ pc, const double *__restrict__ coords, double px,
                               double py, double pz, int g0, int npts, int b, double cutoff,
                               double *__restrict__ phi, long stride_grid, uint8_t *__restrict__ tab, long ldtab) {
    constexpr int NF = (L + 1) * (L + 2) / 2;
    const int lane = threadIdx.x;
    const int ao = __builtin_amdgcn_readfirstlane(sh.ao_off);
    double *__restrict__ row0 = phi + (long)ao * stride_grid + g0;
    bool on;
    if (!kLong) {
        double v[NF];
        const bool ok = lane < npts;
        shell_point<L>(sh, pe, pc, px, py, pz, v);
        double m = 0.0;
#pragma unroll
        for (int c = 0; c < NF; ++c) m = fmax(m, fabs(v[c]));
        on = __ballot(ok && m > cutoff) != 0;
        if (ok) {
#pragma unroll
            for (int c = 0; c < NF; ++c) row0[(long)c * stride_grid + lane] = on ? v[c] : 0.0;
        }
    } else {
        auto at = [&](int p, double *v) {  // point g0 + p < ngrid
            const double *__restrict__ r = coords + 3 * (long)(g0 + p);
            shell_point<L>(sh, pe, pc, r[0], r[1], r[2], v);
        };
        double m = 0.0;
        for (int p = lane; p < npts; p += 64) {
            double v[NF];
            at(p, v);
#pragma unroll
            for (int c = 0; c < NF; ++c) m = fmax(m, fabs(v[c]));
        }
        on = __ballot(m > cutoff) != 0;
        for (int p = lane; p < npts; p += 64) {
            double v[NF];
            if (on) at(p, v);
#pragma unroll
            for (int c = 0; c < NF; ++c) row0[(long)c * stride_grid + p] = on ? v[c] : 0.0;
        }
    }
    if (lane < NF) tab[(long)(ao + lane) * ldtab + b] = on ? 1 : 0;
}

// Grid (blocks, groups), one wave: the group's record into LDS, the lane's point into registers,
// the screening bounds of the group's shells on lanes 0..ns-1, then the shells one after another.
template <bool kLong>
__global__ __launch_bounds__(64) void gto_eval_kernel(const ShellGroup *__restrict__ groups,
                                                      const double *__restrict__ exps,
                                                      const double *__restrict__ coefs,
                                                      const double *__restrict__ coords, int ngrid, int blksize,
                                                      double cutoff, double *__restrict__ phi, long stride_grid,
                                                      uint8_t *__restrict__ tab, long ldtab) {
    __shared__ ShellGroup grp;
    const int b = blockIdx.x, lane = threadIdx.x;
    const int g0 = b * blksize, npts = min(blksize, ngrid - g0);
    {
        const int4 *__restrict__ src = reinterpret_cast<const int4 *>(groups + blockIdx.y);
        int4 *dst = reinterpret_cast<int4 *>(&grp);
        const bool two = lane + 64 < kGroupWords;
        const int4 w0 = src[lane];
        const int4 w1 = src[two ? lane + 64 : lane];
        dst[lane] = w0;
        if (two) dst[lane + 64] = w1;
    }
    // the lane's point (short blocks; clamped into the block: lanes past it compute, and store nothing)
    double px = 0.0, py = 0.0, pz = 0.0;
    if (!kLong) {
        const double *__restrict__ r = coords + 3 * (long)(g0 + (lane < npts ? lane : 0));
        px = r[0];
        py = r[1];
        pz = r[2];
    }
    // a sphere holding the block's points: centre of their bounding box, largest distance from it
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (!kLong) {
        lo[0] = hi[0] = px;
        lo[1] = hi[1] = py;
        lo[2] = hi[2] = pz;
    } else {
        for (int p = lane; p < npts; p += 64) {
            const double *__restrict__ r = coords + 3 * (long)(g0 + p);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                lo[d] = fmin(lo[d], r[d]);
                hi[d] = fmax(hi[d], r[d]);
            }
        }
    }
    double cen[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) cen[d] = 0.5 * (wave_reduce<0>(lo[d]) + wave_reduce<1>(hi[d]));
    double far2 = 0.0;
    if (!kLong) {
        const double dx = px - cen[0], dy = py - cen[1], dz = pz - cen[2];
        far2 = dx * dx + dy * dy + dz * dz;
    } else {
        for (int p = lane; p < npts; p += 64) {
            const double *__restrict__ r = coords + 3 * (long)(g0 + p);
            const double dx = r[0] - cen[0], dy = r[1] - cen[1], dz = r[2] - cen[2];
            far2 = fmax(far2, dx * dx + dy * dy + dz * dz);
        }
    }
    const double far = sqrt(wave_reduce<1>(far2));
    __syncthreads();
    const int ns = __builtin_amdgcn_readfirstlane(grp.ns);
    const bool global = __builtin_amdgcn_readfirstlane(grp.global) != 0;
    // screening: shell `lane` of the group is skipped when its bound is under cutoff / 2 (the
    // margin covers the rounding of the bound and of the values)
    bool skip = false;
    if (lane < ns && cutoff >= 1e-300) {
        const cfgmm_gto_shell &sh = grp.sh[lane];
        const double dx = sh.center[0] - cen[0], dy = sh.center[1] - cen[1], dz = sh.center[2] - cen[2];
        const double dist = sqrt(dx * dx + dy * dy + dz * dz);
        const double rmin = fmax(0.0, dist - far - 1e-12 * (dist + far + 1.0));
        const double *pe = global ? exps + sh.prim_off : grp.pe + sh.prim_off;
        const double *pc = global ? coefs + sh.prim_off : grp.pc + sh.prim_off;
        skip = shell_bound(sh.l, sh.nprim, pe, pc, rmin) < 0.5 * cutoff;
    }
    const unsigned long long skipped = __ballot(skip);
    for (int i = 0; i < ns; ++i) {
        const cfgmm_gto_shell &sh = grp.sh[i];
        const int l = __builtin_amdgcn_readfirstlane(sh.l);
        if ((skipped >> i) & 1) {
            const int ao = __builtin_amdgcn_readfirstlane(sh.ao_off), nf = (l + 1) * (l + 2) / 2;
            double *__restrict__ row0 = phi + (long)ao * stride_grid + g0;
            for (int c = 0; c < nf; ++c)
                for (int p = lane; p < npts; p += 64) row0[(long)c * stride_grid + p] = 0.0;
            if (lane < nf) tab[(long)(ao + lane) * ldtab + b] = 0;
            continue;
        }
#define CFGMM_GTO_SHELL(L_)                                                                                       \
    (global ? tabulate_shell<L_, kLong>(sh, exps, coefs, coords, px, py, pz, g0, npts, b, cutoff, phi, stride_grid, \
                                        tab, ldtab)                                                                \
            : tabulate_shell<L_, kLong>(sh, grp.pe, grp.pc, coords, px, py, pz, g0, npts, b, cutoff, phi,          \
                                        stride_grid, tab, ldtab))
        switch (l) {
            case 0: CFGMM_GTO_SHELL(0); break;
            case 1: CFGMM_GTO_SHELL(1); break;
            case 2: CFGMM_GTO_SHELL(2); break;
            case 3: CFGMM_GTO_SHELL(3); break;
            default: CFGMM_GTO_SHELL(4); break;
        }
#undef CFGMM_GTO_SHELL
    }
}

}  // namespace

int launch_gto(hipStream_t stream, double *phi, size_t stride_grid, uint8_t *tab, size_t ldtab,
               const cfgmm_gto_shell *shells, size_t nshell, const double *exps, const double *coefs,
               size_t nprim_total, const double *coords, size_t ngrid, size_t blksize, double cutoff) {
    if (nshell == 0 || ngrid == 0) return CFGMM_OK;
    // Groups of up to kShellsPerWg consecutive shells and kPrimCap primitives; a shell with more
    // primitives forms a group of its own that reads them from the device copies of the tables.
    std::vector<ShellGroup> groups;
    bool any_global = false;
    for (size_t s = 0; s < nshell;) {
        ShellGroup g{};
        int np = 0;
        while (s < nshell && g.ns < kShellsPerWg) {
            const cfgmm_gto_shell &sh = shells[s];
            if (sh.nprim > kPrimCap) {
                if (g.ns == 0) {
                    g.sh[g.ns++] = sh;
                    g.global = 1;
                    any_global = true;
                    ++s;
                }
                break;
            }
            if (np + sh.nprim > kPrimCap) break;
            g.sh[g.ns] = sh;
            g.sh[g.ns].prim_off = np;
            for (int k = 0; k < sh.nprim; ++k) {
                g.pe[np + k] = exps[sh.prim_off + k];
                g.pc[np + k] = coefs[sh.prim_off + k];
            }
            np += sh.nprim;
            ++g.ns;
            ++s;
        }
        groups.push_back(g);
    }
    const size_t nblk = div_up(ngrid, blksize), ngrp = groups.size();
    if (ngrp > 65535) return fail(CFGMM_EINVAL, "gto: too many shell groups for one launch");
    auto al = [](size_t x) { return div_up(x, 256) * 256; };
    const size_t ntab = any_global ? nprim_total : 0;
    const size_t o_exps = al(sizeof(ShellGroup) * ngrp), o_coefs = o_exps + al(sizeof(double) * ntab);
    const size_t bytes = o_coefs + sizeof(double) * std::max<size_t>(ntab, 1);
    void *ws = nullptr, *hs = nullptr;
    int st = workspace_get(stream, bytes, &ws);
    if (st == CFGMM_OK) st = host_staging_get(stream, bytes, &hs);
    if (st != CFGMM_OK) return st;
    // one copy of the records (and the primitive tables when a group reads them) from pinned memory
    char *h = static_cast<char *>(hs);
    std::memcpy(h, groups.data(), sizeof(ShellGroup) * ngrp);
    if (ntab) {
        std::memcpy(h + o_exps, exps, sizeof(double) * ntab);
        std::memcpy(h + o_coefs, coefs, sizeof(double) * ntab);
    }
    char *base = static_cast<char *>(ws);
    auto *d_groups = reinterpret_cast<ShellGroup *>(base);
    auto *d_exps = reinterpret_cast<double *>(base + o_exps);
    auto *d_coefs = reinterpret_cast<double *>(base + o_coefs);
    CFGMM_HIP_TRY(hipMemcpyAsync(base, h, ntab ? bytes : sizeof(ShellGroup) * ngrp, hipMemcpyHostToDevice, stream));
    st = host_staging_release(stream);  // the next call waits for this copy before refilling
    if (st != CFGMM_OK) return st;
    hipEvent_t t0 = trace_start(stream);
    const dim3 grid((unsigned)nblk, (unsigned)ngrp);
    if (blksize <= 64)
        hipLaunchKernelGGL(gto_eval_kernel<false>, grid, dim3(64), 0, stream, d_groups, d_exps, d_coefs, coords,
                           (int)ngrid, (int)blksize, cutoff, phi, (long)stride_grid, tab, (long)ldtab);
    else
        hipLaunchKernelGGL(gto_eval_kernel<true>, grid, dim3(64), 0, stream, d_groups, d_exps, d_coefs, coords,
                           (int)ngrid, (int)blksize, cutoff, phi, (long)stride_grid, tab, (long)ldtab);
    trace_stop(stream, "gto_eval_kernel", t0);
    return hip_status(hipGetLastError(), "gto_eval_kernel launch");
}

}  // namespace cfgmm
