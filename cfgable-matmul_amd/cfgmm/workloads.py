"""Seeded synthetic workloads for the two hot paths (numpy PCG64 streams).

* DGEMM: row-major f64 operands, i.i.d. uniform[-1, 1).
* Masked AO-DM-AO rho: a random "molecule" -- shells of 1/3/5/7 AOs (s/p/d/f)
  at uniform points in a cube, grid blocks of jittered points, Gaussian-damped
  AO values phi[u][g] = N(0,1) * exp(-0.3 r^2), a per-(grid block, shell)
  non0tab bit set where the block-to-shell minimum distance is below the
  quantile cutoff that gives the requested density, phi zeroed under off bits,
  and PSD density matrices D_s = C_s C_s^T / nocc.

Shell-count rule: 700 shells at 40/30/20/10 % s/p/d/f give 2100 AOs on
average, not the 2228 asked for; shell types are drawn with those
probabilities and then s shells are promoted to p (+2 AOs each) in draw order
(or p demoted to s when over) until the AO count hits the target exactly.

Also: the reference tests' own closed-form generators (sin/cos operands, the
periodic 30 % and 25 % masks), reproduced exactly
(src/impl_matmul.rs:258-259, src/impl_matmul_non0tab.rs:343-357,
src/impl_aodmao2rho.rs:269-296).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

SHELL_SIZES = np.array([1, 3, 5, 7])
SHELL_PROBS = np.array([0.4, 0.3, 0.2, 0.1])


# ---------------------------------------------------------------------------
# DGEMM
# ---------------------------------------------------------------------------
@dataclass
class GemmProblem:
    m: int
    n: int
    k: int
    transa: bool
    transb: bool
    a: np.ndarray
    b: np.ndarray
    lda: int
    ldb: int
    ldc: int

    @property
    def flops(self) -> float:
        return 2.0 * self.m * self.n * self.k

    def dense(self):
        """(op(A), op(B)) as 2-D numpy views."""
        A = self.a.reshape(-1, self.lda)
        B = self.b.reshape(-1, self.ldb)
        A = A[: self.k, : self.m].T if self.transa else A[: self.m, : self.k]
        B = B[: self.n, : self.k].T if self.transb else B[: self.k, : self.n]
        return A, B


def gemm_problem(m, n, k, transa=False, transb=False, seed=0, lda=None, ldb=None, ldc=None) -> GemmProblem:
    rng = np.random.default_rng(seed)
    lda = lda or (m if transa else k)
    ldb = ldb or (k if transb else n)
    ldc = ldc or n
    a_rows = k if transa else m
    b_rows = n if transb else k
    a = rng.uniform(-1.0, 1.0, size=a_rows * lda)
    b = rng.uniform(-1.0, 1.0, size=b_rows * ldb)
    return GemmProblem(m, n, k, transa, transb, a, b, lda, ldb, ldc)


def dgemm_small_nt(seed=0) -> GemmProblem:
    """C(1031x977) = A(1031x1543) . B(977x1543)^T, lda = ldb = 1543, ldc = 977."""
    return gemm_problem(1031, 977, 1543, transa=False, transb=True, seed=seed)


def dgemm_large_nn(seed=0) -> GemmProblem:
    """(3527x6581) . (6581x9583), NN -- the README's benchmark shape (README.md:11-17)."""
    return gemm_problem(3527, 9583, 6581, transa=False, transb=False, seed=seed)


# ---------------------------------------------------------------------------
# masked rho
# ---------------------------------------------------------------------------
@dataclass
class RhoProblem:
    nao: int
    ngrid: int
    nset: int
    blksize: int
    phi: np.ndarray          # (nao, ngrid) float64, zero under off bits
    dms: np.ndarray          # (nset, nao, nao) float64
    tab: np.ndarray          # (nao, nblk) uint8, per-AO expansion of the per-shell bits
    shell_tab: np.ndarray    # (nshell, nblk) bool
    ao_loc: np.ndarray       # (nshell + 1,) AO offsets of the shells
    meta: dict = field(default_factory=dict)

    @property
    def nblk(self) -> int:
        return self.tab.shape[1]

    def counts(self) -> np.ndarray:
        """|L_b| per grid block."""
        return self.tab.astype(bool).sum(axis=0)

    @property
    def flops(self) -> float:
        """Masked algorithmic flops: sum_s sum_b npts_b * (2 n_b^2 + 2 n_b)."""
        n_b = self.counts().astype(np.float64)
        npts = np.full(self.nblk, float(self.blksize))
        npts[-1] = self.ngrid - self.blksize * (self.nblk - 1)
        return float(self.nset * np.sum(npts * (2.0 * n_b * n_b + 2.0 * n_b)))

    @property
    def flops_symmetric(self) -> float:
        """Masked flops of the upper-triangle form (D symmetric, lhs == rhs): each 64-row chunk c of
        a block's list takes the positions from its own first row on, sum_s sum_b npts_b *
        sum_c (2 rows_c (n_b - 64 c) + 2 rows_c)."""
        npts = np.full(self.nblk, float(self.blksize))
        npts[-1] = self.ngrid - self.blksize * (self.nblk - 1)
        tot = 0.0
        for nb, npt in zip(self.counts().astype(np.int64), npts):
            c = np.arange((nb + 63) // 64)
            rows = np.minimum(64, nb - 64 * c).astype(np.float64)
            tot += npt * float(np.sum(2.0 * rows * (nb - 64 * c) + 2.0 * rows))
        return float(self.nset) * tot

    @property
    def dense_flops(self) -> float:
        return float(self.nset) * self.ngrid * (2.0 * self.nao * self.nao + 2.0 * self.nao)

    def skipped_fraction(self) -> float:
        n_b = self.counts().astype(np.float64)
        return float(1.0 - np.sum(n_b * n_b) / (self.nblk * float(self.nao) ** 2))

    def compulsory_bytes(self) -> float:
        """phi nonzero blocks read once + D + rho read/write + mask."""
        n_b = self.counts().astype(np.float64)
        phi = 8.0 * self.blksize * float(n_b.sum())
        return phi + 8.0 * self.nset * self.nao ** 2 + 16.0 * self.nset * self.ngrid + float(self.tab.size)


def _shell_types(rng, nshell, nao_target):
    types = rng.choice(4, size=nshell, p=SHELL_PROBS)
    sizes = SHELL_SIZES[types].copy()
    if nao_target is not None:
        total = int(sizes.sum())
        if (nao_target - total) % 2 != 0:
            raise ValueError("target AO count must have the parity of the shell count")
        if not nshell <= nao_target <= 7 * nshell:
            raise ValueError("target AO count out of reach of the shell count")
        i = 0
        while total != nao_target and i < nshell:  # first pass: s <-> p in draw order
            if total < nao_target and sizes[i] == 1:
                sizes[i] = 3
                total += 2
            elif total > nao_target and sizes[i] == 3:
                sizes[i] = 1
                total -= 2
            i += 1
        i = 0
        while total != nao_target:  # only if still off: any one-step promotion / demotion, cyclically
            if total < nao_target and sizes[i] < 7:
                sizes[i] += 2
                total += 2
            elif total > nao_target and sizes[i] > 1:
                sizes[i] -= 2
                total -= 2
            i = (i + 1) % nshell
    return sizes


def rho_problem(seed=0, nshell=700, nao=2228, ngrid=16384, nset=3, blksize=64, density=0.3, box=25.0,
                nocc=200, dm_logspan=None) -> RhoProblem:
    """The masked AO-DM-AO workload (defaults: 700 shells, 2228 AOs, 256 blocks x 64 points, 3 PSD DMs, 30 %).

    ``dm_logspan=(lo, hi)`` replaces each D entry's magnitude by 10**U(lo, hi),
    keeping its sign and the symmetry of D (entries drawn for i <= j, mirrored).
    """
    rng = np.random.default_rng(seed)
    sizes = _shell_types(rng, nshell, nao)
    nao = int(sizes.sum())
    ao_loc = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    shell_of_ao = np.repeat(np.arange(nshell), sizes)
    centers = rng.uniform(0.0, box, size=(nshell, 3))

    nblk = (ngrid + blksize - 1) // blksize
    blk_centers = rng.uniform(0.0, box, size=(nblk, 3))
    blk_of_pt = np.minimum(np.arange(ngrid) // blksize, nblk - 1)
    points = blk_centers[blk_of_pt] + rng.uniform(-0.5, 0.5, size=(ngrid, 3))

    # per (block, shell) minimum distance
    mind = np.empty((nshell, nblk))
    for b in range(nblk):
        pts = points[b * blksize: min((b + 1) * blksize, ngrid)]
        d2 = ((pts[:, None, :] - centers[None, :, :]) ** 2).sum(-1)
        mind[:, b] = np.sqrt(d2.min(axis=0))
    cutoff = np.quantile(mind, density)
    shell_tab = mind < cutoff
    tab = shell_tab[shell_of_ao].astype(np.uint8)

    # phi[u][g] = N(0,1) exp(-0.3 r^2), zero under off bits.  r^2 summed over x, y, z in that
    # order (as a sum over a length-3 axis adds), and the damping applied per shell's rows.
    r2 = (centers[:, 0, None] - points[None, :, 0]) ** 2
    r2 += (centers[:, 1, None] - points[None, :, 1]) ** 2
    r2 += (centers[:, 2, None] - points[None, :, 2]) ** 2  # (nshell, ngrid)
    damp = np.exp(-0.3 * r2)
    phi = rng.standard_normal((nao, ngrid))
    for sh in range(nshell):
        phi[ao_loc[sh]:ao_loc[sh + 1]] *= damp[sh]
    on = np.repeat(tab.astype(bool), blksize, axis=1)[:, :ngrid]
    phi[~on] = 0.0

    dms = np.empty((nset, nao, nao))
    for s in range(nset):
        C = rng.standard_normal((nao, nocc))
        dms[s] = C @ C.T / nocc
        if dm_logspan is not None:
            lo, hi = dm_logspan
            mag = 10.0 ** rng.uniform(lo, hi, size=(nao, nao))
            mag = np.triu(mag) + np.triu(mag, 1).T
            dms[s] = np.sign(dms[s]) * mag
    meta = dict(seed=seed, nshell=nshell, density=density, cutoff=float(cutoff), box=box, nocc=nocc,
                dm_logspan=dm_logspan)
    return RhoProblem(nao, ngrid, nset, blksize, np.ascontiguousarray(phi), np.ascontiguousarray(dms), tab, shell_tab,
                      ao_loc, meta)


def rho_reference_dense(p: RhoProblem, phi_lhs=None, phi_rhs=None) -> np.ndarray:
    """rho[s][g] = sum_{u,v} lhs[u,g] D_s[u,v] rhs[v,g] with lhs/rhs zeroed under off bits (numpy float64)."""
    on = np.repeat(p.tab.astype(bool), p.blksize, axis=1)[:, : p.ngrid]
    L = np.where(on, p.phi if phi_lhs is None else phi_lhs, 0.0)
    R = np.where(on, p.phi if phi_rhs is None else phi_rhs, 0.0)
    out = np.empty((p.nset, p.ngrid))
    for s in range(p.nset):
        out[s] = np.einsum("ug,ug->g", L, p.dms[s] @ R)
    return out


# ---------------------------------------------------------------------------
# the reference tests' closed-form generators
# ---------------------------------------------------------------------------
MASK_30 = np.array([True, False, True, False, False, True, False, False, False, False])  # src/impl_matmul_non0tab.rs:343
MASK_25 = np.array([True, False, True, False, False, False, False, False, False, False,
                    False, True, True, False, False, False, False, False, True, False])  # src/impl_aodmao2rho.rs:277-280


def sin_cos_operands(na: int, nb: int):
    """a[x] = sin(x), b[x] = cos(x) over flat indices (src/impl_matmul.rs:258-259)."""
    return np.sin(np.arange(na, dtype=np.float64)), np.cos(np.arange(nb, dtype=np.float64))


def periodic_mask(nrows: int, nblk: int, pattern: np.ndarray) -> np.ndarray:
    """tab[i] = pattern[i % len(pattern)] over the flat (row, blk) index."""
    idx = np.arange(nrows * nblk)
    return pattern[idx % len(pattern)].reshape(nrows, nblk).astype(np.uint8)


def zero_under_mask(x: np.ndarray, tab: np.ndarray, blksize: int, ncols: int) -> None:
    """x[p][j] = 0 where !tab[p][j / blksize] (src/impl_matmul_non0tab.rs:346-357)."""
    on = np.repeat(tab.astype(bool), blksize, axis=1)[:, :ncols]
    x.reshape(tab.shape[0], -1)[:, :ncols][~on] = 0.0


# ---------------------------------------------------------------------------
# many independent rho problems (batched launch, serving mix)
# ---------------------------------------------------------------------------
@dataclass
class Molecule:
    """A generated molecule on a 64-block grid, sliced by the problems that use it."""
    seed: int
    prob: RhoProblem  # nset = 4 density matrices, ngrid = 64 * nblk_max


@dataclass
class PoolProblem:
    mol: int      # index into the molecule pool
    ngrid: int    # leading grid points of the molecule used (multiple of 64)
    nset: int     # leading density matrices used

    def flops(self, pool) -> float:
        p = pool[self.mol].prob
        nblk = self.ngrid // p.blksize
        n_b = p.counts()[:nblk].astype(np.float64)
        return float(self.nset * np.sum(p.blksize * (2.0 * n_b * n_b + 2.0 * n_b)))


def molecule_pool(n=32, seed=100, nao_range=(300, 2228), max_blocks=64, blksize=64, density_range=(0.1, 0.6),
                  nset=4):
    """n molecules, nao ~ U{nao_range} (shells ~ nao/3), density ~ U(density_range), max_blocks 64-point blocks."""
    rng = np.random.default_rng(seed)
    pool = []
    for i in range(n):
        nao = int(rng.integers(nao_range[0], nao_range[1] + 1))
        nshell = max(1, nao // 3)
        if (nao - nshell) % 2:
            nao += 1
        dens = float(rng.uniform(*density_range))
        p = rho_problem(seed=seed * 1000 + i, nshell=nshell, nao=nao, ngrid=max_blocks * blksize, nset=nset,
                        blksize=blksize, density=dens, nocc=min(200, nao))
        pool.append(Molecule(seed * 1000 + i, p))
    return pool


def batched_problems(pool, n=512, seed=7, max_blocks=64, blksize=64, nset_range=(1, 4)):
    """n problems: a molecule from the pool, ngrid = blksize * U{1..max_blocks}, nset ~ U{nset_range}."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        m = int(rng.integers(0, len(pool)))
        ngrid = blksize * int(rng.integers(1, max_blocks + 1))
        nset = int(rng.integers(nset_range[0], nset_range[1] + 1))
        out.append(PoolProblem(m, ngrid, nset))
    return out


# ---------------------------------------------------------------------------
# serving mix
# ---------------------------------------------------------------------------
def splitmix64_uniform(seed: int, n: int) -> np.ndarray:
    """The serving loop's device generator: x[i] = (splitmix64((seed << 32) + i) >> 11) * 2^-52 - 1."""
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) << np.uint64(32)) + np.arange(n, dtype=np.uint64)
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -52 - 1.0


def serving_gemm_operands(m: int, n: int, k: int, seed: int):
    """(A, B) of a served GEMM request, as the device generates them (row-major, NN)."""
    return (splitmix64_uniform(2 * seed, m * k).reshape(m, k), splitmix64_uniform(2 * seed + 1, k * n).reshape(k, n))


def result_checksum(x: np.ndarray) -> int:
    """The serving loop's result hash: sum_i splitmix64(bits(x[i]) ^ splitmix64(i)) mod 2^64."""
    def mix(z):
        with np.errstate(over="ignore"):
            z = z + np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            return z ^ (z >> np.uint64(31))
    bits = np.ascontiguousarray(x, dtype=np.float64).ravel().view(np.uint64)
    with np.errstate(over="ignore"):
        return int(np.sum(mix(bits ^ mix(np.arange(bits.size, dtype=np.uint64))), dtype=np.uint64))


def tile_bytes_estimate(p: "RhoProblem", nslices: int = 6) -> float:
    """Bytes of the split-precision resident tiles of a problem (kpad_b = |L_b| rounded up to 64)."""
    kp = np.ceil(p.counts() / 64.0) * 64.0
    return float(np.sum(p.nset * nslices * kp * kp + p.blksize * kp * (nslices + 8)))


def serving_pool(n=8, seed=500, nao_range=(500, 4000), blocks_range=(64, 1024), density_range=(0.1, 0.6), nset=4,
                 blksize=64, max_tile_bytes=None, threads=None, stratify=True):
    """Molecules for the serving mix: nao ~ U{nao_range} (shells ~ nao/3), ngrid = blksize * nblk
    (4096 .. 65536 points at the defaults), mask density ~ U(density_range), nset PSD DMs.  With
    stratify, molecule k draws nblk from the k-th of n equal parts of blocks_range (so the pool
    spans the whole range); otherwise nblk ~ U{blocks_range}.  With max_tile_bytes, a molecule
    whose 6-slice resident tiles would exceed it is redrawn; by default every draw is kept.  The
    draws are sequential; the molecules of kept draws are generated on a thread pool (each depends
    only on its own seed)."""
    import os
    from concurrent.futures import ThreadPoolExecutor

    rng = np.random.default_rng(seed)

    def make(i, nao, nshell, nblk, dens):
        return rho_problem(seed=seed * 1000 + i, nshell=nshell, nao=nao, ngrid=nblk * blksize, nset=nset,
                           blksize=blksize, density=dens, nocc=min(200, nao))
    pool = []
    i = 0
    while len(pool) < n:
        draws = []
        while len(pool) + len(draws) < n:
            nao = int(rng.integers(nao_range[0], nao_range[1] + 1))
            nshell = max(1, nao // 3)
            if (nao - nshell) % 2:
                nao += 1
            lo, hi = blocks_range
            if stratify:
                kk, span = len(pool) + len(draws), hi - lo + 1
                lo, hi = blocks_range[0] + span * kk // n, blocks_range[0] + span * (kk + 1) // n - 1
            nblk = int(rng.integers(lo, hi + 1))
            dens = float(rng.uniform(*density_range))
            i += 1
            if max_tile_bytes is not None:
                # cheap pre-check before generating: blocks of about dens * nao AOs
                kp = np.ceil(dens * nao / 64.0) * 64.0
                if nblk * (nset * 6 * kp * kp + blksize * kp * 14) > max_tile_bytes:
                    continue
            draws.append((i, nao, nshell, nblk, dens))
        with ThreadPoolExecutor(threads or min(16, os.cpu_count() or 1)) as ex:
            probs = list(ex.map(lambda d: make(*d), draws))
        for d, p in zip(draws, probs):
            if max_tile_bytes is not None and tile_bytes_estimate(p) > max_tile_bytes:
                continue
            pool.append(Molecule(seed * 1000 + d[0], p))
    return pool


def independent_problems(n=512, seed=4000, nao_range=(300, 2228), blocks_range=(1, 64), nset_range=(1, 4),
                         density_range=(0.1, 0.6), blksize=64, threads=None):
    """n independent masked rho problems, each its own molecule, D, phi and mask: problem i draws
    nao ~ U{nao_range} (shells ~ nao/3, nao raised by one where the shell count's parity needs it),
    ngrid = blksize * U{blocks_range}, nset ~ U{nset_range} and density ~ U(density_range) from
    default_rng((seed, i)), and is then generated by rho_problem(seed=seed + i).  Problems are
    generated on a thread pool; each depends only on its own seeds."""
    import os
    from concurrent.futures import ThreadPoolExecutor

    def one(i):
        r = np.random.default_rng((seed, i))
        nao = int(r.integers(nao_range[0], nao_range[1] + 1))
        nshell = max(1, nao // 3)
        if (nao - nshell) % 2:
            nao += 1
        nblk = int(r.integers(blocks_range[0], blocks_range[1] + 1))
        nset = int(r.integers(nset_range[0], nset_range[1] + 1))
        dens = float(r.uniform(*density_range))
        return rho_problem(seed=seed + i, nshell=nshell, nao=nao, ngrid=nblk * blksize, nset=nset, blksize=blksize,
                           density=dens, nocc=min(200, nao))
    with ThreadPoolExecutor(threads or min(16, os.cpu_count() or 1)) as ex:
        return list(ex.map(one, range(n)))


@dataclass
class ServedRequest:
    kind: int            # 0 rho, 1 GEMM
    arrival_s: float
    mol: int = -1        # rho: pool index
    ngrid: int = 0
    nset: int = 0
    m: int = 0
    n: int = 0
    k: int = 0
    seed: int = 0

    def flops(self, pool) -> float:
        if self.kind == 1:
            return 2.0 * self.m * self.n * self.k
        return PoolProblem(self.mol, self.ngrid, self.nset).flops(pool)


def serving_requests(pool, n, rate, seed=11, rho_fraction=0.7, gemm_range=(256, 4096), min_blocks=64):
    """n requests with Poisson arrivals at `rate` per second: rho (a pool molecule, leading
    blksize*U{min(min_blocks, nblk)..nblk} points -- 4096 and up at the defaults -- and U{1..nset}
    DMs) with probability rho_fraction, else a DGEMM with m, n, k = round(exp U(ln lo, ln hi))."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    out = []
    lo, hi = np.log(gemm_range[0]), np.log(gemm_range[1])
    for i in range(n):
        if rng.uniform() < rho_fraction:
            mi = int(rng.integers(0, len(pool)))
            p = pool[mi].prob
            nb = int(rng.integers(min(min_blocks, p.nblk), p.nblk + 1))
            out.append(ServedRequest(0, float(t[i]), mol=mi, ngrid=nb * p.blksize,
                                     nset=int(rng.integers(1, p.nset + 1))))
        else:
            m, nn, k = (int(round(np.exp(rng.uniform(lo, hi)))) for _ in range(3))
            out.append(ServedRequest(1, float(t[i]), m=m, n=nn, k=k, seed=int(rng.integers(1, 2 ** 30))))
    return out


def gto_workload(ngrid=16384, blksize=64, nshell=700, box=25.0, seed=11):
    """A Cartesian GTO basis on a grid like the rho workload's: nshell shells uniform in a box,
    40 % s / 40 % p / 20 % d, 1-3 primitives (exponents 10^U(-0.7, 0.5), coefficients
    U(0.3, 1.2)), grid points jittered by U(-0.5, 0.5) around ngrid/blksize random block centres."""
    rng = np.random.default_rng(seed)
    ls = rng.choice(3, size=nshell, p=[0.4, 0.4, 0.2])
    nprim = rng.integers(1, 4, size=nshell)
    poff = np.concatenate([[0], np.cumsum(nprim)[:-1]])
    nf = (ls + 1) * (ls + 2) // 2
    aoff = np.concatenate([[0], np.cumsum(nf)[:-1]])
    centers = rng.uniform(0.0, box, (nshell, 3))
    exps = 10.0 ** rng.uniform(-0.7, 0.5, int(nprim.sum()))
    coefs = rng.uniform(0.3, 1.2, int(nprim.sum()))
    nblk = -(-ngrid // blksize)
    blk_c = rng.uniform(0.0, box, (nblk, 3))
    coords = np.ascontiguousarray(blk_c[np.arange(ngrid) // blksize] + rng.uniform(-0.5, 0.5, (ngrid, 3)))
    return dict(centers=centers, ls=ls, nprim=nprim, poff=poff, aoff=aoff, nao=int(nf.sum()), exps=exps,
                coefs=coefs, coords=coords, ngrid=ngrid, blksize=blksize, nblk=nblk, cutoff=1e-8)
