"""CPU oracle for AO tabulation on a grid -- TEST INFRASTRUCTURE ONLY.

No counterpart in the reference crate, which takes phi and the block mask as inputs
(README.md:49: the mask is "similar to PySCF non0tab"; src/impl_aodmao2rho.rs:171-244 reads
it per AO row).  This is a numpy restatement of Cartesian Gaussian-type orbitals:

    phi[ao_off + c][g] = dx^lx dy^ly dz^lz * sum_k coefs[k] exp(-exps[k] r^2),  d = coords[g] - R,

components c in PySCF's Cartesian order (lx = l..0, then ly = l-lx..0, lz = l-lx-ly), and the
per-AO block mask tab[ao][b] = [max over the points of block b and the shell's components of
|phi| > cutoff], phi zeroed where it is 0.  Pinned in tests/test_gto_oracle.py by closed forms
(Gaussian normalisation integrals by Gauss-Hermite quadrature), a known-answer component
order, the trace identity xx+yy+zz = r^2 s, translation invariance and a brute-force mask.
"""
from __future__ import annotations

import numpy as np


def cart_components(l: int):
    """(lx, ly, lz) of a Cartesian shell in PySCF order."""
    return [(lx, ly, l - lx - ly) for lx in range(l, -1, -1) for ly in range(l - lx, -1, -1)]


def eval_gto(shells, exps, coefs, coords, nao):
    """shells: iterable of (center(3), l, nprim, prim_off, ao_off); coords (ngrid, 3).
    Returns phi (nao, ngrid), zero rows for AOs owned by no shell."""
    coords = np.asarray(coords, dtype=np.float64)
    phi = np.zeros((nao, coords.shape[0]))
    for center, l, nprim, po, ao in shells:
        d = coords - np.asarray(center, dtype=np.float64)[None, :]
        r2 = np.einsum("gi,gi->g", d, d)
        rad = np.zeros_like(r2)
        for k in range(po, po + nprim):
            rad += coefs[k] * np.exp(-exps[k] * r2)
        for c, (lx, ly, lz) in enumerate(cart_components(l)):
            phi[ao + c] = d[:, 0] ** lx * d[:, 1] ** ly * d[:, 2] ** lz * rad
    return phi


def block_mask(phi, shells, ngrid, blksize, cutoff):
    """(phi with off blocks zeroed, tab (nao, nblk) uint8) for the shells' AO rows."""
    nblk = -(-ngrid // blksize)
    out = phi.copy()
    tab = np.zeros((phi.shape[0], nblk), np.uint8)
    for _, l, _, _, ao in shells:
        nf = (l + 1) * (l + 2) // 2
        for b in range(nblk):
            sl = slice(b * blksize, min((b + 1) * blksize, ngrid))
            on = np.max(np.abs(phi[ao:ao + nf, sl])) > cutoff
            tab[ao:ao + nf, b] = 1 if on else 0
            if not on:
                out[ao:ao + nf, sl] = 0.0
    return out, tab


def random_basis(rng, nshell, box=6.0, lmax=4, max_prim=3):
    """A random contracted Cartesian basis: (shells, exps, coefs, nao)."""
    shells, exps, coefs = [], [], []
    ao = 0
    for _ in range(nshell):
        l = int(rng.integers(0, lmax + 1))
        n = int(rng.integers(1, max_prim + 1))
        center = rng.uniform(0.0, box, 3)
        shells.append((center, l, n, len(exps), ao))
        exps.extend(10.0 ** rng.uniform(-1.0, 1.0, n))
        coefs.extend(rng.uniform(0.2, 1.5, n))
        ao += (l + 1) * (l + 2) // 2
    return shells, np.array(exps), np.array(coefs), ao
