"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

Inputs come from the reference tests' closed-form generators (sin/cos
operands, periodic 30 % / 25 % masks; src/impl_matmul.rs:258-259,
src/impl_matmul_non0tab.rs:343-357, src/impl_aodmao2rho.rs:269-296) and one
small generated molecule; expected outputs were computed in extended
precision and rounded once.  The CPU oracle and the HIP path are both held
to them: plain FP64 to 1e-12 relative Frobenius error, split precision
(6 int8 slices) to 1e-10.
"""
from __future__ import annotations

import glob
import os

import numpy as np
import pytest

from helpers import rel_fro

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def load(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_fixtures_present():
    names = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))
    assert names == ["gemm_nn", "gemm_nt", "gemm_tn", "gemm_tt", "masked_n", "masked_t", "rho_molecule",
                     "rho_sincos"]


def test_fixture_masks_are_authoritative():
    """The masked fixtures hold nonzero data under off bits, so ignoring the mask fails them."""
    f = load("masked_n")
    m, n, k, lda, ldb, ldc, transa, _, ldtab, blksize = (int(x) for x in f["dims"])
    tab = f["tab"].reshape(k, ldtab)
    nblk = -(-n // blksize)
    off = ~np.repeat(tab[:, :nblk].astype(bool), blksize, axis=1)[:, :n]
    assert np.count_nonzero(f["b"].reshape(k, ldb)[:, :n][off]) > 0
    g = load("rho_sincos")
    sa, sg, nao, ngrid, nset, ldtab, blksize = (int(x) for x in g["dims"])
    tab = g["tab"].reshape(nao, ldtab)
    off = ~np.repeat(tab[:, : -(-ngrid // blksize)].astype(bool), blksize, axis=1)[:, :ngrid]
    assert np.count_nonzero(g["lhs"].reshape(nao, sg)[:, :ngrid][off]) > 0


# ---------------------------------------------------------------------------
# CPU: the oracle against the fixtures
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", ["gemm_nn", "gemm_nt", "gemm_tn", "gemm_tt"])
def test_oracle_gemm_golden(name):
    import oracle
    f = load(name)
    m, n, k, lda, ldb, ldc, ta, tb = (int(x) for x in f["dims"])
    c = f["c0"].copy()
    oracle.matmul(c, f["a"], f["b"], m, n, k, lda, ldb, ldc, ta, tb)
    assert rel_fro(c, f["c"]) <= 1e-14
    pad = np.ones(c.reshape(m, ldc)[:, n:].shape, bool)
    assert np.array_equal(c.reshape(m, ldc)[:, n:][pad], f["c0"].reshape(m, ldc)[:, n:][pad])


@pytest.mark.parametrize("name", ["masked_n", "masked_t"])
def test_oracle_masked_golden(name):
    import oracle
    f = load(name)
    m, n, k, lda, ldb, ldc, ta, tb, ldtab, blksize = (int(x) for x in f["dims"])
    c = f["c0"].copy()
    assert oracle.matmul_non0tab(c, f["a"], f["b"], m, n, k, lda, ldb, ldc, ta, tb, f["tab"], ldtab, blksize) == 0
    assert rel_fro(c, f["c"]) <= 1e-14


@pytest.mark.parametrize("name", ["rho_sincos", "rho_molecule"])
def test_oracle_rho_golden(name):
    import oracle
    f = load(name)
    sa, sg, nao, ngrid, nset, ldtab, blksize = (int(x) for x in f["dims"])
    r = f["rhos0"].copy()
    assert oracle.aodmao2rho(r, f["dms"], f["lhs"], f["rhs"], sa, sg, nao, ngrid, nset, f["tab"], ldtab, blksize) == 0
    assert rel_fro(r, f["rhos"]) <= 1e-14


# ---------------------------------------------------------------------------
# GPU: the HIP path through the C ABI against the same fixtures
# ---------------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("name", ["gemm_nn", "gemm_nt", "gemm_tn", "gemm_tt"])
def test_gpu_gemm_golden(name):
    import cfgmm
    f = load(name)
    m, n, k, lda, ldb, ldc, ta, tb = (int(x) for x in f["dims"])
    c = f["c0"].copy()
    cfgmm.MatmulLoops.matmul_loop_macro_mb(c, f["a"], f["b"], m, n, k, lda, ldb, ldc, bool(ta), bool(tb))
    assert rel_fro(c, f["c"]) <= 1e-12
    assert np.array_equal(c.reshape(m, ldc)[:, n:], f["c0"].reshape(m, ldc)[:, n:])  # padding untouched


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["masked_n", "masked_t"])
def test_gpu_masked_golden(name):
    import cfgmm
    f = load(name)
    m, n, k, lda, ldb, ldc, ta, tb, ldtab, blksize = (int(x) for x in f["dims"])
    c = f["c0"].copy()
    cfgmm.MatmulLoops.matmul_loop_parallel_mnk_pack_a_non0tab(blksize, c, f["a"], f["b"], m, n, k, lda, ldb, ldc,
                                                              bool(ta), False, f["tab"], ldtab)
    assert rel_fro(c, f["c"]) <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["rho_sincos", "rho_molecule"])
def test_gpu_rho_golden(name):
    import cfgmm
    f = load(name)
    sa, sg, nao, ngrid, nset, ldtab, blksize = (int(x) for x in f["dims"])
    r = f["rhos0"].copy()
    cfgmm.MatmulLoops.aodmao2rho_loop_parallel(blksize, r, f["dms"], f["lhs"], f["rhs"], [sa, sg], [nao, ngrid], nset,
                                               f["tab"], ldtab)
    assert rel_fro(r, f["rhos"]) <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["rho_sincos", "rho_molecule"])
def test_gpu_rho_split_golden(name):
    """Split precision, 6 int8 slices, on device buffers; rho accumulates into the fixture's start value."""
    import cfgmm
    f = load(name)
    sa, sg, nao, ngrid, nset, ldtab, blksize = (int(x) for x in f["dims"])
    s = cfgmm.Stream()
    bufs = {k: cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(f[k])) for k in ("rhos0", "dms", "lhs", "rhs", "tab")}
    cfgmm.aodmao2rho_dev(s, bufs["rhos0"], bufs["dms"], bufs["lhs"], bufs["rhs"], sa, sg, nao, ngrid, nset,
                         bufs["tab"], ldtab, blksize, cfgmm.PREC_OZAKI_INT8, 6)
    s.synchronize()
    r = bufs["rhos0"].to_array(np.float64, f["rhos0"].shape)
    for b in bufs.values():
        b.free()
    s.destroy()
    assert rel_fro(r, f["rhos"]) <= 1e-10
