"""GPU parity of the fused masked AO-DM-AO rho kernel against the CPU oracle.

Tolerance: relative Frobenius error <= 1e-12 (plain FP64), plus size-independent
properties at the full workload size (PSD => rho >= 0, linearity in D)."""
import numpy as np
import pytest

import cfgmm
import oracle
from cfgmm import workloads as W
from helpers import rel_fro

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _oracle_rho(p, rhos0=None, blksize=None):
    blksize = blksize or p.blksize
    rhos = np.zeros(p.nset * p.ngrid) if rhos0 is None else rhos0.copy()
    assert oracle.aodmao2rho(rhos, p.dms.ravel(), p.phi.ravel(), p.phi.ravel(), p.nao, p.ngrid, p.nao, p.ngrid,
                             p.nset, np.ascontiguousarray(p.tab), p.nblk, blksize) == 0
    return rhos


@pytest.fixture(scope="module")
def workload():
    return W.rho_problem(seed=0)  # 700 shells, nao 2228, ngrid 16384 (256 x 64), 3 PSD DMs, 30 %


def test_rho_full_workload_vs_oracle(workload):
    p = workload
    ref = _oracle_rho(p)
    rhos = np.zeros(p.nset * p.ngrid)
    cfgmm.MatmulLoops().aodmao2rho_loop_parallel(64, rhos, p.dms.ravel(), p.phi.ravel(), p.phi.ravel(),
                                                 [p.nao, p.ngrid], [p.nao, p.ngrid], p.nset,
                                                 np.ascontiguousarray(p.tab), p.nblk)
    err = rel_fro(rhos, ref)
    print(f"rho nao={p.nao} ngrid={p.ngrid} nset={p.nset} rel_fro = {err:.3e}")
    assert err <= TOL
    assert np.all(rhos >= -1e-12 * np.max(np.abs(rhos)))  # PSD D with lhs == rhs


def test_rho_full_workload_dev_api_linear(workload):
    """Device-pointer API; rho(D1 + D2) == rho(D1) + rho(D2) at full size."""
    p = workload
    dms = np.concatenate([p.dms[0].ravel(), p.dms[1].ravel(), (p.dms[0] + p.dms[1]).ravel()])
    s = cfgmm.Stream()
    d_r = cfgmm.DeviceBuffer(8 * 3 * p.ngrid)
    d_r.zero()
    d_d = cfgmm.DeviceBuffer.from_array(dms)
    d_phi = cfgmm.DeviceBuffer.from_array(p.phi)
    d_tab = cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(p.tab))
    cfgmm.aodmao2rho_dev(s, d_r, d_d, d_phi, d_phi, p.nao, p.ngrid, p.nao, p.ngrid, 3, d_tab, p.nblk, p.blksize)
    s.synchronize()
    r = d_r.to_array(np.float64, (3, p.ngrid))
    assert rel_fro(r[2], r[0] + r[1]) <= TOL
    s.destroy()


@pytest.mark.parametrize("blksize", [16, 48, 64, 128])
@pytest.mark.parametrize("nset", [1, 2, 5])
def test_rho_blksize_nset_ragged(blksize, nset):
    """Odd ngrid (partial last block), nset > 4 (two set groups), rho accumulated into nonzero input,
    lhs != rhs and both nonzero under off bits (mask authority)."""
    rng = np.random.default_rng(blksize * 10 + nset)
    nao, ngrid = 301, 1001
    nblk = (ngrid + blksize - 1) // blksize
    lhs = rng.standard_normal(nao * ngrid)
    rhs = rng.standard_normal(nao * ngrid)
    dms = rng.standard_normal(nset * nao * nao)
    tab = W.periodic_mask(nao, nblk, W.MASK_25)
    rhos = rng.standard_normal(nset * ngrid)
    ref = rhos.copy()
    assert oracle.aodmao2rho(ref, dms, lhs, rhs, nao, ngrid, nao, ngrid, nset, tab, nblk, blksize) == 0
    cfgmm.MatmulLoops().aodmao2rho_loop_parallel(blksize, rhos, dms, lhs, rhs, [nao, ngrid], [nao, ngrid], nset,
                                                 tab, nblk)
    assert rel_fro(rhos, ref) <= TOL


def test_rho_strided_padding_untouched():
    rng = np.random.default_rng(3)
    nao, ngrid, nset, blksize = 140, 300, 2, 48
    sa, sg = nao + 3, ngrid + 5
    nblk = (ngrid + blksize - 1) // blksize
    ldtab = nblk + 2
    phi = rng.standard_normal(nao * sg)
    dms = rng.standard_normal(nset * sa * sa)
    tab = W.periodic_mask(nao, ldtab, W.MASK_30)
    rhos = rng.standard_normal(nset * sg)
    ref = rhos.copy()
    assert oracle.aodmao2rho(ref, dms, phi, phi, sa, sg, nao, ngrid, nset, tab, ldtab, blksize) == 0
    got = rhos.copy()
    cfgmm.MatmulLoops().aodmao2rho_loop_parallel(blksize, got, dms, phi, phi, [sa, sg], [nao, ngrid], nset, tab, ldtab)
    assert rel_fro(got.reshape(nset, sg)[:, :ngrid], ref.reshape(nset, sg)[:, :ngrid]) <= TOL
    assert np.array_equal(got.reshape(nset, sg)[:, ngrid:], rhos.reshape(nset, sg)[:, ngrid:])


def test_rho_empty_and_full_mask():
    rng = np.random.default_rng(4)
    nao, ngrid, blksize = 90, 250, 64
    nblk = (ngrid + blksize - 1) // blksize
    phi = rng.standard_normal(nao * ngrid)
    dms = rng.standard_normal(nao * nao)
    rhos = rng.standard_normal(ngrid)
    r0 = rhos.copy()
    cfgmm.MatmulLoops().aodmao2rho_loop_parallel(blksize, rhos, dms, phi, phi, [nao, ngrid], [nao, ngrid], 1,
                                                 np.zeros((nao, nblk), np.uint8), nblk)
    assert np.array_equal(rhos, r0)
    cfgmm.MatmulLoops().aodmao2rho_loop_parallel(blksize, rhos, dms, phi, phi, [nao, ngrid], [nao, ngrid], 1,
                                                 np.ones((nao, nblk), np.uint8), nblk)
    L = phi.reshape(nao, ngrid)
    assert rel_fro(rhos, r0 + np.einsum("ug,ug->g", L, dms.reshape(nao, nao) @ L)) <= TOL


def test_rho_reference_test_generators():
    """src/impl_aodmao2rho.rs:264-345 at reduced ngrid: sin/cos phi, D = sin(0.38x+1.2), 25 % mask,
    BLKSIZE 48, nao 2228, odd ngrid."""
    nao, ngrid, nset = 2228, 8193, 1
    lhs, rhs = W.sin_cos_operands(nao * ngrid, nao * ngrid)
    dms = np.sin(np.arange(nset * nao * nao, dtype=np.float64) * 0.38 + 1.2)
    nblk = (ngrid + 47) // 48
    tab = W.periodic_mask(nao, nblk, W.MASK_25)
    W.zero_under_mask(lhs, tab, 48, ngrid)
    W.zero_under_mask(rhs, tab, 48, ngrid)
    ref = np.zeros(ngrid)
    assert oracle.aodmao2rho(ref, dms, lhs, rhs, nao, ngrid, nao, ngrid, nset, tab, nblk, 48) == 0
    rhos = np.zeros(ngrid)
    cfgmm.aodmao2rho_anyway(rhos, dms, lhs, rhs, [nao, ngrid], [nao, ngrid], nset, tab, nblk)
    assert rel_fro(rhos, ref) <= TOL


def _reference_harness(pattern):
    """The inputs of src/impl_aodmao2rho.rs:264-425: nao 2228, ngrid 131073 (odd: a partial last
    block), nset 1, lhs[x] = sin(x), rhs[x] = cos(x), D[x] = sin(0.38 x + 1.2), BLKSIZE 48, both
    operands zeroed under off bits of the periodic mask (None: every bit on)."""
    nao, ngrid, nset = 2228, 131073, 1
    lhs, rhs = W.sin_cos_operands(nao * ngrid, nao * ngrid)
    dms = np.sin(np.arange(nset * nao * nao, dtype=np.float64) * 0.38 + 1.2)
    nblk = (ngrid + 47) // 48
    if pattern is None:
        tab = np.ones((nao, nblk), np.uint8)
    else:
        tab = W.periodic_mask(nao, nblk, pattern)
        W.zero_under_mask(lhs, tab, 48, ngrid)
        W.zero_under_mask(rhs, tab, 48, ngrid)
    return nao, ngrid, nset, lhs, rhs, dms, tab, nblk


@pytest.mark.parametrize("case", ["masked_25", "all_dense"])
def test_rho_reference_harness_full_size(case):
    """test_aodmao2rho_1 (25 % periodic mask) and test_aodmao2rho_all_dense at the reference's
    own sizes, through aodmao2rho_anyway (BLKSIZE 48), against the oracle.

    The all-dense sin/cos sum cancels: |rho| <= 4 out of S_g = sum_uv |lhs_ug D_uv rhs_vg| ~ 1.3e6,
    so two correct FP64 summation orders differ by ~1e-12 relative (oracle vs numpy: 7.9e-13 at
    ngrid 4097).  What bounds a summation there is the componentwise form |err_g| <= c eps S_g
    (oracle vs numpy: c = 0.027); the test holds the GPU to c = 4, about 9e-16 S_g, while one
    dropped or doubled product term is about S_g / nao^2 = 2e-7 S_g.  The masked case is well
    conditioned and is held to the 1e-12 normwise bar as well."""
    nao, ngrid, nset, lhs, rhs, dms, tab, nblk = _reference_harness(W.MASK_25 if case == "masked_25" else None)
    ref = np.zeros(ngrid)
    assert oracle.aodmao2rho(ref, dms, lhs, rhs, nao, ngrid, nao, ngrid, nset, tab, nblk, 48) == 0
    rhos = np.zeros(ngrid)
    cfgmm.aodmao2rho_anyway(rhos, dms, lhs, rhs, [nao, ngrid], [nao, ngrid], nset, tab, nblk)
    err = rel_fro(rhos, ref)
    L, R = lhs.reshape(nao, ngrid), rhs.reshape(nao, ngrid)
    S = np.einsum("ug,ug->g", np.abs(L), np.abs(dms.reshape(nao, nao)) @ np.abs(R))
    comp = float(np.max(np.abs(rhos - ref) / (np.finfo(np.float64).eps * S)))
    print(f"rho reference harness {case} nao={nao} ngrid={ngrid} rel_fro = {err:.3e}, "
          f"max |err| / (eps S_g) = {comp:.3f}", flush=True)
    assert comp <= 4.0
    if case == "masked_25":
        assert err <= TOL


def test_rho_deterministic(workload):
    p = workload
    outs = []
    for _ in range(2):
        rhos = np.zeros(p.nset * p.ngrid)
        cfgmm.MatmulLoops().aodmao2rho_loop_parallel(64, rhos, p.dms.ravel(), p.phi.ravel(), p.phi.ravel(),
                                                     [p.nao, p.ngrid], [p.nao, p.ngrid], p.nset,
                                                     np.ascontiguousarray(p.tab), p.nblk)
        outs.append(rhos)
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("blksize", [64, 96])
def test_rho_task_order_bitwise(blksize):
    """AO-ordered and block-major task orders (CFGMM_TUNE_RHO_ORDER) with 1 and 3 density
    matrices per workgroup give bitwise-equal results in the general (non-symmetric) form, repeated AO-ordered calls too (the order
    inside a plan bucket varies), and match the oracle."""
    p = W.rho_problem(seed=6, nshell=200, nao=640, ngrid=blksize * 23 + 17, nset=3, blksize=blksize, density=0.3)
    s = cfgmm.Stream()
    d_d = cfgmm.DeviceBuffer.from_array(p.dms)
    d_phi = cfgmm.DeviceBuffer.from_array(p.phi)
    d_tab = cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(p.tab))
    d_r = cfgmm.DeviceBuffer(8 * p.nset * p.ngrid)
    old = (cfgmm.tune_get(cfgmm.TUNE_RHO_ORDER), cfgmm.tune_get(cfgmm.TUNE_RHO_MAX_SETS),
           cfgmm.tune_get(cfgmm.TUNE_RHO_SYMMETRIC))
    outs = []
    try:
        cfgmm.tune_set(cfgmm.TUNE_RHO_SYMMETRIC, 1)  # the general form in both orders
        for order, ms in ((0, 1), (0, 1), (1, 1), (0, 3), (1, 3)):
            cfgmm.tune_set(cfgmm.TUNE_RHO_ORDER, order)
            cfgmm.tune_set(cfgmm.TUNE_RHO_MAX_SETS, ms)
            d_r.zero()
            cfgmm.aodmao2rho_dev(s, d_r, d_d, d_phi, d_phi, p.nao, p.ngrid, p.nao, p.ngrid, p.nset, d_tab, p.nblk,
                                 p.blksize)
            s.synchronize()
            outs.append(d_r.to_array(np.float64, (p.nset * p.ngrid,)))
    finally:
        cfgmm.tune_set(cfgmm.TUNE_RHO_ORDER, old[0])
        cfgmm.tune_set(cfgmm.TUNE_RHO_MAX_SETS, old[1])
        cfgmm.tune_set(cfgmm.TUNE_RHO_SYMMETRIC, old[2])
    for b in (d_d, d_phi, d_tab, d_r):
        b.free()
    s.destroy()
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    ref = np.zeros(p.nset * p.ngrid)
    assert oracle.aodmao2rho(ref, p.dms.ravel(), p.phi.ravel(), p.phi.ravel(), p.nao, p.ngrid, p.nao, p.ngrid,
                             p.nset, np.ascontiguousarray(p.tab), p.nblk, p.blksize) == 0
    assert rel_fro(outs[0], ref) <= TOL


def test_rho_max_sets_tunable_bitwise():
    """Density matrices per workgroup (1..4) change the launch shape only: results are bitwise equal."""
    p = W.rho_problem(seed=4, nshell=180, nao=560, ngrid=2048, nset=4, density=0.35)
    s = cfgmm.Stream()
    d_d = cfgmm.DeviceBuffer.from_array(p.dms)
    d_phi = cfgmm.DeviceBuffer.from_array(p.phi)
    d_tab = cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(p.tab))
    d_r = cfgmm.DeviceBuffer(8 * p.nset * p.ngrid)
    old = cfgmm.tune_get(cfgmm.TUNE_RHO_MAX_SETS)
    outs = []
    try:
        for ms in (1, 2, 3, 4):
            cfgmm.tune_set(cfgmm.TUNE_RHO_MAX_SETS, ms)
            d_r.zero()
            cfgmm.aodmao2rho_dev(s, d_r, d_d, d_phi, d_phi, p.nao, p.ngrid, p.nao, p.ngrid, p.nset, d_tab, p.nblk,
                                 p.blksize)
            s.synchronize()
            outs.append(d_r.to_array(np.float64, (p.nset, p.ngrid)))
    finally:
        cfgmm.tune_set(cfgmm.TUNE_RHO_MAX_SETS, old)
    for b in (d_d, d_phi, d_tab, d_r):
        b.free()
    s.destroy()
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


def test_rho_multi_component_bitwise_and_oracle():
    """Four left-hand components (phi and three gradient-like factors) over one D.rhs: each component
    equals a separate evaluation bitwise and the oracle to 1e-12; rho accumulates into start values."""
    p = W.rho_problem(seed=12, nshell=150, nao=480, ngrid=1600, nset=2, density=0.35)
    rng = np.random.default_rng(3)
    comps = [p.phi] + [p.phi * rng.uniform(-2, 2, size=p.phi.shape) for _ in range(3)]
    lhs = np.ascontiguousarray(np.stack(comps))  # (4, nao, ngrid)
    r0 = rng.standard_normal((4, p.nset, p.ngrid))
    s = cfgmm.Stream()
    d_d = cfgmm.DeviceBuffer.from_array(p.dms)
    d_l = cfgmm.DeviceBuffer.from_array(lhs)
    d_rhs = cfgmm.DeviceBuffer.from_array(p.phi)
    d_t = cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(p.tab))
    d_r = cfgmm.DeviceBuffer.from_array(r0)
    cfgmm.aodmao2rho_multi_dev(s, d_r, p.nset * p.ngrid, d_d, d_l, p.nao * p.ngrid, 4, d_rhs, p.nao, p.ngrid, p.nao,
                               p.ngrid, p.nset, d_t, p.nblk, p.blksize)
    s.synchronize()
    multi = d_r.to_array(np.float64, r0.shape)
    singles = []
    for c in range(4):
        d_one = cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(r0[c]))
        d_lc = cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(lhs[c]))
        cfgmm.aodmao2rho_dev(s, d_one, d_d, d_lc, d_rhs, p.nao, p.ngrid, p.nao, p.ngrid, p.nset, d_t, p.nblk, p.blksize)
        s.synchronize()
        singles.append(d_one.to_array(np.float64, (p.nset, p.ngrid)))
        d_one.free()
        d_lc.free()
    for b in (d_d, d_l, d_rhs, d_t, d_r):
        b.free()
    s.destroy()
    for c in range(4):
        assert np.array_equal(multi[c], singles[c])
    ref = r0[2].ravel().copy()
    assert oracle.aodmao2rho(ref, p.dms.ravel(), lhs[2].ravel(), p.phi.ravel(), p.nao, p.ngrid, p.nao, p.ngrid,
                             p.nset, np.ascontiguousarray(p.tab), p.nblk, p.blksize) == 0
    assert rel_fro(multi[2].ravel() - r0[2].ravel(), ref - r0[2].ravel()) <= 1e-12


def test_rho_long_lists_span_segments():
    """Blocks with more active AOs than one LDS segment of the compacted list holds (~2000 of 3000,
    staged 1536 at a time): both task orders against the oracle, bitwise equal to each other."""
    rng = np.random.default_rng(9)
    nao, ngrid, blksize, nset = 3000, 256 + 40, 64, 2
    nblk = (ngrid + blksize - 1) // blksize
    tab = (rng.random((nao, nblk)) < 0.65).astype(np.uint8)
    phi = rng.standard_normal((nao, ngrid))
    C = rng.standard_normal((nset, nao, 40))
    dms = np.einsum("sik,sjk->sij", C, C) / 40
    dms = (dms + dms.transpose(0, 2, 1)) / 2  # bitwise symmetric
    assert tab.sum(axis=0).max() > 1536
    ref = np.zeros(nset * ngrid)
    assert oracle.aodmao2rho(ref, dms.ravel(), phi.ravel(), phi.ravel(), nao, ngrid, nao, ngrid, nset, tab, nblk,
                             blksize) == 0
    old = (cfgmm.tune_get(cfgmm.TUNE_RHO_ORDER), cfgmm.tune_get(cfgmm.TUNE_RHO_SYMMETRIC))
    outs = []
    try:
        for order, sym_off in ((0, 1), (1, 1), (0, 0)):
            cfgmm.tune_set(cfgmm.TUNE_RHO_ORDER, order)
            cfgmm.tune_set(cfgmm.TUNE_RHO_SYMMETRIC, sym_off)
            rhos = np.zeros(nset * ngrid)
            cfgmm.MatmulLoops().aodmao2rho_loop_parallel(blksize, rhos, dms.ravel(), phi.ravel(), phi.ravel(),
                                                         [nao, ngrid], [nao, ngrid], nset, tab, nblk)
            outs.append(rhos)
    finally:
        cfgmm.tune_set(cfgmm.TUNE_RHO_ORDER, old[0])
        cfgmm.tune_set(cfgmm.TUNE_RHO_SYMMETRIC, old[1])
    assert np.array_equal(outs[0], outs[1])
    assert rel_fro(outs[0], ref) <= TOL
    assert rel_fro(outs[2], ref) <= TOL  # symmetric form, ranges past the chunk spanning segments
    assert not np.array_equal(outs[2], outs[0])  # it did take the symmetric form (other rounding)


def _rho_dev(p, dms, lhs, rhs, same_buffer=True, blksize=None):
    blksize = blksize or p.blksize
    s = cfgmm.Stream()
    d_d = cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(dms))
    d_l = cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(lhs))
    d_rr = d_l if same_buffer else cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(rhs))
    d_t = cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(p.tab))
    d_r = cfgmm.DeviceBuffer(8 * p.nset * p.ngrid)
    d_r.zero()
    cfgmm.aodmao2rho_dev(s, d_r, d_d, d_l, d_rr, p.nao, p.ngrid, p.nao, p.ngrid, p.nset, d_t, p.nblk, blksize)
    s.synchronize()
    out = d_r.to_array(np.float64, (p.nset * p.ngrid,))
    for b in {id(x): x for x in (d_d, d_l, d_rr, d_t, d_r)}.values():
        b.free()
    s.destroy()
    return out


def _oracle(p, dms, lhs, rhs):
    ref = np.zeros(p.nset * p.ngrid)
    assert oracle.aodmao2rho(ref, np.ascontiguousarray(dms).ravel(), lhs.ravel(), rhs.ravel(), p.nao, p.ngrid, p.nao,
                             p.ngrid, p.nset, np.ascontiguousarray(p.tab), p.nblk, p.blksize) == 0
    return ref


def test_rho_phi_offsets_beyond_32_bits():
    """phi rows 2^21 + 9 points apart (byte offsets into phi past 2^32; only the used points of each
    row uploaded), so the slice loads take 64-bit offsets: equal to the oracle on the compact arrays
    and bitwise to the compact evaluation, whose offsets fit in 32 bits."""
    p = W.rho_problem(seed=17, nshell=100, nao=300, ngrid=64 * 9 + 5, nset=2, density=0.4)
    big = (1 << 21) + 9
    assert (p.nao - 1) * big * 8 > 2 ** 32
    lib = cfgmm.lib()
    s = cfgmm.Stream()
    d_phi = cfgmm.DeviceBuffer(8 * ((p.nao - 1) * big + p.ngrid))
    phi = np.ascontiguousarray(p.phi)
    for x in range(p.nao):
        cfgmm.check(lib.cfgmm_memcpy_htod(d_phi.ptr + 8 * x * big, phi[x].ctypes.data, 8 * p.ngrid))
    d_d = cfgmm.DeviceBuffer.from_array(p.dms)
    d_t = cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(p.tab))
    d_r = cfgmm.DeviceBuffer(8 * ((p.nset - 1) * big + p.ngrid))
    d_r.zero()
    cfgmm.aodmao2rho_dev(s, d_r, d_d, d_phi, d_phi, p.nao, big, p.nao, p.ngrid, p.nset, d_t, p.nblk, p.blksize)
    s.synchronize()
    got = np.empty((p.nset, p.ngrid))
    for q in range(p.nset):
        cfgmm.check(lib.cfgmm_memcpy_dtoh(got[q].ctypes.data, d_r.ptr + 8 * q * big, 8 * p.ngrid))
    for b in (d_phi, d_d, d_t, d_r):
        b.free()
    s.destroy()
    assert rel_fro(got, _oracle_rho(p).reshape(p.nset, p.ngrid)) <= TOL
    assert np.array_equal(got, _rho_dev(p, p.dms, p.phi, p.phi).reshape(p.nset, p.ngrid))


def test_rho_symmetric_form_on_fresh_streams():
    """Each call's symmetry flag is a word of its own, zeroed with the workspace, tested against an
    epoch unique in the process: a new stream's workspace (possibly memory an earlier one used)
    never carries a value that makes a symmetric D look asymmetric.  Per stream, an asymmetric call
    (the flag written) then a symmetric one, which must take the symmetric form."""
    p = W.rho_problem(seed=23, nshell=120, nao=380, ngrid=64 * 9 + 7, nset=2, density=0.4)
    bad = p.dms.copy()
    lst = np.flatnonzero(p.tab[:, 0])
    bad[1, int(lst[2]), int(lst[-1])] += 10.0
    old = cfgmm.tune_get(cfgmm.TUNE_RHO_SYMMETRIC)
    try:
        cfgmm.tune_set(cfgmm.TUNE_RHO_SYMMETRIC, 1)
        gen = _rho_dev(p, p.dms, p.phi, p.phi)
        cfgmm.tune_set(cfgmm.TUNE_RHO_SYMMETRIC, 0)
        ref = _oracle(p, p.dms, p.phi, p.phi)
        for _ in range(12):  # _rho_dev makes a stream, its workspace, and destroys the stream
            _rho_dev(p, bad, p.phi, p.phi)
            sym = _rho_dev(p, p.dms, p.phi, p.phi)
            assert rel_fro(sym, ref) <= TOL
            assert not np.array_equal(sym, gen)  # the symmetric form ran
    finally:
        cfgmm.tune_set(cfgmm.TUNE_RHO_SYMMETRIC, old)


def test_rho_symmetric_to_rounding():
    """D symmetric only to rounding (entries of the upper triangle moved by one or two units in the
    last place, as a general product C C^T leaves them) still takes the symmetric form, held to the
    oracle on that D; a difference of 1e-13 relative (about 450 units) keeps the general form,
    bitwise.  The resident handle's test at creation is the same."""
    p = W.rho_problem(seed=29, nshell=200, nao=620, ngrid=64 * 12 + 9, nset=2, density=0.4)
    rng = np.random.default_rng(3)
    near, far = p.dms.copy(), p.dms.copy()
    iu = np.triu_indices(p.nao, 1)
    for s in range(p.nset):
        m = rng.random(len(iu[0])) < 0.5
        x = near[s][iu]
        near[s][iu] = np.where(m, np.nextafter(np.nextafter(x, np.inf), np.inf), x)
    far[1, iu[0][5], iu[1][5]] *= 1.0 + 1e-13
    assert not np.array_equal(near[0], near[0].T)
    old = cfgmm.tune_get(cfgmm.TUNE_RHO_SYMMETRIC)
    try:
        cfgmm.tune_set(cfgmm.TUNE_RHO_SYMMETRIC, 1)
        gen_near = _rho_dev(p, near, p.phi, p.phi)
        gen_far = _rho_dev(p, far, p.phi, p.phi)
        cfgmm.tune_set(cfgmm.TUNE_RHO_SYMMETRIC, 0)
        sym_near = _rho_dev(p, near, p.phi, p.phi)
        got_far = _rho_dev(p, far, p.phi, p.phi)
    finally:
        cfgmm.tune_set(cfgmm.TUNE_RHO_SYMMETRIC, old)
    assert not np.array_equal(sym_near, gen_near)  # the symmetric form ran
    assert rel_fro(sym_near, _oracle(p, near, p.phi, p.phi)) <= TOL
    assert np.array_equal(got_far, gen_far)  # the general form ran
    # resident handles: the symmetric layout (about half the bytes) for `near`, the general one for `far`
    s = cfgmm.Stream()
    d_phi = cfgmm.DeviceBuffer.from_array(p.phi)
    d_t = cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(p.tab))
    sizes = []
    for dm in (near, far):
        d_d = cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(dm))
        h = cfgmm.RhoResident(s, d_d, d_phi, d_phi, p.nao, p.ngrid, p.nao, p.ngrid, p.nset, d_t, p.nblk, p.blksize,
                              cfgmm.PREC_OZAKI_INT8, 5)
        d_r = cfgmm.DeviceBuffer(8 * p.nset * p.ngrid)
        d_r.zero()
        h.eval(s, d_r, p.ngrid)
        s.synchronize()
        out = d_r.to_array(np.float64, (p.nset * p.ngrid,))
        sizes.append(h.bytes)
        assert rel_fro(out, _oracle(p, dm, p.phi, p.phi)) <= 1e-10
        h.destroy()
        d_d.free()
        d_r.free()
    d_phi.free()
    d_t.free()
    s.destroy()
    assert sizes[0] < 0.75 * sizes[1]


@pytest.mark.parametrize("blksize,nset", [(64, 3), (96, 5)])
def test_rho_symmetric_form(blksize, nset):
    """Bitwise-symmetric D_s with lhs and rhs one buffer take the upper-triangle form: the oracle
    to 1e-12, the general form (CFGMM_TUNE_RHO_SYMMETRIC = 1) to 1e-13, bitwise repeatable.  The
    general form is kept whenever the symmetric one would be wrong: one large asymmetric entry
    inside the mask, one asymmetric set among symmetric ones, lhs and rhs in separate buffers."""
    p = W.rho_problem(seed=21, nshell=260, nao=800, ngrid=blksize * 21 + 13, nset=nset, blksize=blksize,
                      density=0.4)
    assert all(np.array_equal(d, d.T) for d in p.dms)
    ref = _oracle(p, p.dms, p.phi, p.phi)
    old = cfgmm.tune_get(cfgmm.TUNE_RHO_SYMMETRIC)
    try:
        cfgmm.tune_set(cfgmm.TUNE_RHO_SYMMETRIC, 0)
        sym = _rho_dev(p, p.dms, p.phi, p.phi)
        sym2 = _rho_dev(p, p.dms, p.phi, p.phi)
        cfgmm.tune_set(cfgmm.TUNE_RHO_SYMMETRIC, 1)
        gen = _rho_dev(p, p.dms, p.phi, p.phi)
        cfgmm.tune_set(cfgmm.TUNE_RHO_SYMMETRIC, 0)
        assert rel_fro(sym, ref) <= TOL
        assert rel_fro(sym, gen) <= 1e-13
        assert np.array_equal(sym, sym2)
        assert not np.array_equal(sym, gen)  # the symmetric form ran
        # one entry D[u][v] (u, v both active in block 0, in different chunks) made asymmetric
        lst = np.flatnonzero(p.tab[:, 0])
        u, v = int(lst[3]), int(lst[min(len(lst) - 1, 70)])
        bad = p.dms.copy()
        bad[nset - 1, u, v] += 1000.0
        got = _rho_dev(p, bad, p.phi, p.phi)
        assert rel_fro(got, _oracle(p, bad, p.phi, p.phi)) <= TOL
        # lhs and rhs equal in value but separate buffers: the general form (bitwise)
        assert np.array_equal(_rho_dev(p, p.dms, p.phi, p.phi, same_buffer=False), gen)
    finally:
        cfgmm.tune_set(cfgmm.TUNE_RHO_SYMMETRIC, old)


def test_rho_symmetric_form_strided_padding():
    """The symmetric form with D rows strided past nao (asymmetric garbage in the padding, which the
    test on the device must ignore) and phi and rho rows strided past ngrid: the oracle to 1e-12,
    the general form to 1e-13, the rho padding untouched."""
    p = W.rho_problem(seed=33, nshell=200, nao=650, ngrid=64 * 17 + 9, nset=2, density=0.35)
    sa, sg = p.nao + 5, p.ngrid + 11
    rng = np.random.default_rng(8)
    D = rng.standard_normal((p.nset, sa, sa))  # asymmetric padding
    D[:, :p.nao, :p.nao] = p.dms
    phi = np.zeros((p.nao, sg))
    phi[:, :p.ngrid] = p.phi
    phi[:, p.ngrid:] = 1e30
    r0 = np.full(p.nset * sg, 0.5)
    ref = r0.reshape(p.nset, sg)[:, :p.ngrid].copy()
    ref += _oracle_rho(p).reshape(p.nset, p.ngrid)
    s = cfgmm.Stream()
    d_d = cfgmm.DeviceBuffer.from_array(D)
    d_phi = cfgmm.DeviceBuffer.from_array(phi)
    d_t = cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(p.tab))
    outs = []
    old = cfgmm.tune_get(cfgmm.TUNE_RHO_SYMMETRIC)
    try:
        for off in (0, 1):
            cfgmm.tune_set(cfgmm.TUNE_RHO_SYMMETRIC, off)
            d_r = cfgmm.DeviceBuffer.from_array(r0)
            cfgmm.aodmao2rho_dev(s, d_r, d_d, d_phi, d_phi, sa, sg, p.nao, p.ngrid, p.nset, d_t, p.nblk, p.blksize)
            s.synchronize()
            outs.append(d_r.to_array(np.float64, (p.nset, sg)))
            d_r.free()
    finally:
        cfgmm.tune_set(cfgmm.TUNE_RHO_SYMMETRIC, old)
    for b in (d_d, d_phi, d_t):
        b.free()
    s.destroy()
    sym, gen = outs
    assert rel_fro(sym[:, :p.ngrid] - 0.5, ref - 0.5) <= TOL
    assert rel_fro(sym[:, :p.ngrid] - 0.5, gen[:, :p.ngrid] - 0.5) <= 1e-13
    assert not np.array_equal(sym, gen)  # the symmetric form ran
    assert np.all(sym[:, p.ngrid:] == 0.5)
