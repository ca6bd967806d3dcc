#!/usr/bin/env python3
"""Writes the golden fixtures in this folder (small .npz files: inputs and expected outputs).

The reference crate asserts nothing in its tests (results are printed beside
a dense OpenBLAS product), so these fixtures come from independent arithmetic:
numpy products in x87 extended precision (np.longdouble, 64-bit mantissa) over
inputs built with the reference tests' closed-form generators
(src/impl_matmul.rs:258-259 sin/cos operands, src/impl_matmul_non0tab.rs:343-357
30 % periodic mask, src/impl_aodmao2rho.rs:269-296 25 % periodic mask) and one
small generated molecule.  Expected outputs are rounded to float64 once.

Every output accumulates into a nonzero starting value (the reference adds
into `c` / `rhos`), leading dimensions are padded, and the masked cases leave
nonzero values under off mask bits: the mask, not the data, decides.

Run from anywhere:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "cfgable-matmul_amd"))

from cfgmm import workloads as W  # noqa: E402

LD = np.longdouble


def op_view(flat, rows, cols, ld, trans):
    """op(X) of a row-major flat buffer: X is rows x cols (trans: stored cols x rows)."""
    if trans:
        return flat[: cols * ld].reshape(cols, ld)[:, :rows].T
    return flat[: rows * ld].reshape(rows, ld)[:, :cols]


def gemm_case(transa, transb, m=37, n=29, k=53, pad=3):
    lda = (m if transa else k) + pad
    ldb = (k if transb else n) + pad
    ldc = n + pad
    na = (k if transa else m) * lda
    nb = (n if transb else k) * ldb
    a, b = W.sin_cos_operands(na, nb)
    c0 = 0.5 * np.sin(0.1 * np.arange(m * ldc))
    A = op_view(a, m, k, lda, transa).astype(LD)
    B = op_view(b, k, n, ldb, transb).astype(LD)
    c = c0.copy().reshape(m, ldc)
    c[:, :n] = (c[:, :n].astype(LD) + A @ B).astype(np.float64)
    return dict(a=a, b=b, c0=c0, c=c.ravel(), dims=np.array([m, n, k, lda, ldb, ldc, transa, transb]))


def masked_case(transa, m=53, k=112, n=200, blksize=48, pad=2):
    nblk = (n + blksize - 1) // blksize
    ldtab = nblk + 1
    lda = (m if transa else k) + pad
    ldb = n + pad
    ldc = n + pad
    a, b = W.sin_cos_operands((k if transa else m) * lda, k * ldb)
    tab = W.periodic_mask(k, ldtab, W.MASK_30)
    c0 = 0.25 * np.cos(0.3 * np.arange(m * ldc))
    A = op_view(a, m, k, lda, transa).astype(LD)
    B = op_view(b, k, n, ldb, False).astype(LD).copy()
    on = np.repeat(tab[:, :nblk].astype(bool), blksize, axis=1)[:, :n]
    B[~on] = 0  # the mask decides; b itself keeps its nonzero values
    c = c0.copy().reshape(m, ldc)
    c[:, :n] = (c[:, :n].astype(LD) + A @ B).astype(np.float64)
    return dict(a=a, b=b, c0=c0, c=c.ravel(), tab=tab.ravel(),
                dims=np.array([m, n, k, lda, ldb, ldc, transa, 0, ldtab, blksize]))


def rho_expected(rhos0, dms, lhs, rhs, stride_ao, stride_grid, nao, ngrid, nset, tab, ldtab, blksize):
    nblk = (ngrid + blksize - 1) // blksize
    T = tab.reshape(nao, ldtab)[:, :nblk].astype(bool)
    on = np.repeat(T, blksize, axis=1)[:, :ngrid]
    L = np.where(on, lhs.reshape(nao, stride_grid)[:, :ngrid], 0.0).astype(LD)
    R = np.where(on, rhs.reshape(nao, stride_grid)[:, :ngrid], 0.0).astype(LD)
    out = rhos0.copy().reshape(nset, stride_grid)
    for s in range(nset):
        D = dms[s * stride_ao * stride_ao:(s + 1) * stride_ao * stride_ao].reshape(stride_ao, stride_ao)
        D = D[:nao, :nao].astype(LD)
        out[s, :ngrid] = (out[s, :ngrid].astype(LD) + np.sum(L * (D @ R), axis=0)).astype(np.float64)
    return out.ravel()


def rho_sincos_case(nao=100, ngrid=200, nset=2, blksize=64):
    stride_ao, stride_grid = nao + 2, ngrid + 5
    nblk = (ngrid + blksize - 1) // blksize
    ldtab = nblk + 2
    lhs, rhs = W.sin_cos_operands(nao * stride_grid, nao * stride_grid)
    dms = np.sin(0.37 * np.arange(nset * stride_ao * stride_ao) + 1.0)
    tab = W.periodic_mask(nao, ldtab, W.MASK_25).ravel()
    rhos0 = 0.1 * np.cos(np.arange(nset * stride_grid))
    rhos = rho_expected(rhos0, dms, lhs, rhs, stride_ao, stride_grid, nao, ngrid, nset, tab, ldtab, blksize)
    return dict(dms=dms, lhs=lhs, rhs=rhs, tab=tab, rhos0=rhos0, rhos=rhos,
                dims=np.array([stride_ao, stride_grid, nao, ngrid, nset, ldtab, blksize]))


def rho_molecule_case():
    p = W.rho_problem(seed=11, nshell=40, nao=120, ngrid=512, nset=3, density=0.3, box=12.0, nocc=60)
    rhos0 = np.zeros(p.nset * p.ngrid)
    tab = np.ascontiguousarray(p.tab).ravel()
    rhos = rho_expected(rhos0, p.dms.ravel(), p.phi.ravel(), p.phi.ravel(), p.nao, p.ngrid, p.nao, p.ngrid, p.nset,
                        tab, p.nblk, p.blksize)
    return dict(dms=p.dms.ravel(), lhs=p.phi.ravel(), rhs=p.phi.ravel(), tab=tab, rhos0=rhos0, rhos=rhos,
                dims=np.array([p.nao, p.ngrid, p.nao, p.ngrid, p.nset, p.nblk, p.blksize]))


CASES = {
    "gemm_nn": lambda: gemm_case(0, 0),
    "gemm_nt": lambda: gemm_case(0, 1),
    "gemm_tn": lambda: gemm_case(1, 0),
    "gemm_tt": lambda: gemm_case(1, 1),
    "masked_n": lambda: masked_case(0),
    "masked_t": lambda: masked_case(1),
    "rho_sincos": rho_sincos_case,
    "rho_molecule": rho_molecule_case,
}


def main() -> None:
    for name, make in CASES.items():
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **make())
        print(f"wrote {path} ({os.path.getsize(path) / 1024:.0f} KiB)")


if __name__ == "__main__":
    main()
