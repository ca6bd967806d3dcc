"""The symmetric forms (one buffer as lhs and rhs, D_s symmetric) and the general forms over
small and ragged shapes: nao below one 64-row tile, ngrid below one block, one block, empty and full masks, nset up to 5,
every blksize kind.  Per call, batched and resident evaluations against the oracle (1e-12 FP64,
1e-10 split precision); batched bitwise equal to per call."""
import numpy as np
import pytest

import cfgmm
import oracle
from helpers import rel_fro

pytestmark = pytest.mark.gpu

CASES = [  # (nao, ngrid, blksize, nset, density)
    (1, 17, 16, 1, 1.0), (5, 64, 64, 2, 0.6), (63, 65, 64, 1, 0.5), (64, 1, 16, 3, 0.7), (65, 200, 48, 5, 0.4),
    (130, 130, 128, 2, 0.9), (200, 640, 64, 4, 0.0), (257, 333, 16, 3, 1.0), (300, 1000, 96, 2, 0.3),
    (700, 256, 64, 1, 0.55),
]


def _case(nao, ngrid, blksize, nset, density, seed):
    rng = np.random.default_rng(seed)
    nblk = -(-ngrid // blksize)
    tab = (rng.random((nao, nblk)) < density).astype(np.uint8)
    phi = rng.standard_normal((nao, ngrid))
    C = rng.standard_normal((nset, nao, 7))
    dms = np.einsum("sik,sjk->sij", C, C)
    dms = (dms + dms.transpose(0, 2, 1)) / 2  # bitwise symmetric
    r0 = rng.standard_normal(nset * ngrid)
    ref = r0.copy()
    assert oracle.aodmao2rho(ref, dms.ravel(), phi.ravel(), phi.ravel(), nao, ngrid, nao, ngrid, nset, tab, nblk,
                             blksize) == 0
    return tab, phi, dms, r0, ref, nblk


@pytest.mark.parametrize("i", range(len(CASES)))
def test_symmetric_forms_small_shapes(i):
    nao, ngrid, blksize, nset, density = CASES[i]
    tab, phi, dms, r0, ref, nblk = _case(nao, ngrid, blksize, nset, density, seed=100 + i)
    s = cfgmm.Stream()
    d_d = cfgmm.DeviceBuffer.from_array(dms)
    d_phi = cfgmm.DeviceBuffer.from_array(phi)
    d_t = cfgmm.DeviceBuffer.from_array(tab)
    # per call
    d_r = cfgmm.DeviceBuffer.from_array(r0)
    cfgmm.aodmao2rho_dev(s, d_r, d_d, d_phi, d_phi, nao, ngrid, nao, ngrid, nset, d_t, nblk, blksize)
    s.synchronize()
    single = d_r.to_array(np.float64, r0.shape)
    # batched (the same problem twice, one buffer each)
    d_b = [cfgmm.DeviceBuffer.from_array(r0) for _ in range(2)]
    probs = [cfgmm.RhoProblem(b.ptr, d_d.ptr, d_phi.ptr, d_phi.ptr, nao, ngrid, nao, ngrid, nset, d_t.ptr, nblk,
                              blksize) for b in d_b]
    cfgmm.aodmao2rho_batched_dev(s, probs)
    s.synchronize()
    batched = [b.to_array(np.float64, r0.shape) for b in d_b]
    # resident, five slices
    d_z = cfgmm.DeviceBuffer.from_array(r0)
    h = cfgmm.RhoResident(s, d_d, d_phi, d_phi, nao, ngrid, nao, ngrid, nset, d_t, nblk, blksize,
                          cfgmm.PREC_OZAKI_INT8, 5)
    h.eval(s, d_z, ngrid)
    s.synchronize()
    res = d_z.to_array(np.float64, r0.shape)
    h.destroy()
    for b in [d_d, d_phi, d_t, d_r, d_z] + d_b:
        b.free()
    s.destroy()
    inc = ref - r0
    if not tab.any():  # empty mask: nothing added
        for out in [single, res] + batched:
            assert np.array_equal(out, r0)
        return
    assert rel_fro(single - r0, inc) <= 1e-12
    for b in batched:
        assert np.array_equal(b, single)
    assert rel_fro(res - r0, inc) <= 1e-10


@pytest.mark.parametrize("i", range(len(CASES)))
def test_vxc_symmetric_small_shapes(i):
    """The AO-grid-AO contraction with one buffer as lhs and rhs (upper tiles of S_b, the mirror
    into V's lower triangle) against numpy, V padded and accumulated into."""
    nao, ngrid, blksize, _, density = CASES[i]
    rng = np.random.default_rng(200 + i)
    nblk = -(-ngrid // blksize)
    tab = (rng.random((nao, nblk)) < density).astype(np.uint8)
    phi = rng.standard_normal((nao, ngrid))
    w = rng.uniform(0.1, 2.0, ngrid)
    ldv = nao + 2
    V0 = rng.standard_normal(nao * ldv)
    ref = np.zeros((nao, nao))
    for b in range(nblk):
        cols = slice(b * blksize, min((b + 1) * blksize, ngrid))
        on = tab[:, b].astype(bool)
        A = np.where(on[:, None], phi[:, cols] * w[cols], 0.0)
        B = np.where(on[:, None], phi[:, cols], 0.0)
        ref += A @ B.T
    s = cfgmm.Stream()
    d_v = cfgmm.DeviceBuffer.from_array(V0)
    d_phi = cfgmm.DeviceBuffer.from_array(phi)
    d_w = cfgmm.DeviceBuffer.from_array(w)
    d_t = cfgmm.DeviceBuffer.from_array(tab)
    cfgmm.aogridao2mat_dev(s, d_v, ldv, d_phi, d_phi, d_w, ngrid, nao, ngrid, d_t, nblk, blksize)
    s.synchronize()
    got = d_v.to_array(np.float64, (nao, ldv))
    for b in (d_v, d_phi, d_w, d_t):
        b.free()
    s.destroy()
    V0m = V0.reshape(nao, ldv)
    assert np.array_equal(got[:, nao:], V0m[:, nao:])
    inc = got[:, :nao] - V0m[:, :nao]
    if not tab.any():
        assert np.array_equal(got, V0m)
        return
    assert rel_fro(inc, ref) <= 1e-12


@pytest.mark.parametrize("i", range(len(CASES)))
def test_general_forms_small_shapes(i):
    """The same shapes in the general form: lhs and rhs different buffers, D not symmetric."""
    nao, ngrid, blksize, nset, density = CASES[i]
    rng = np.random.default_rng(300 + i)
    nblk = -(-ngrid // blksize)
    tab = (rng.random((nao, nblk)) < density).astype(np.uint8)
    lhs = rng.standard_normal((nao, ngrid))
    rhs = rng.standard_normal((nao, ngrid))
    dms = rng.standard_normal((nset, nao, nao))
    r0 = rng.standard_normal(nset * ngrid)
    ref = r0.copy()
    assert oracle.aodmao2rho(ref, dms.ravel(), lhs.ravel(), rhs.ravel(), nao, ngrid, nao, ngrid, nset, tab, nblk,
                             blksize) == 0
    s = cfgmm.Stream()
    d_d = cfgmm.DeviceBuffer.from_array(dms)
    d_l = cfgmm.DeviceBuffer.from_array(lhs)
    d_rr = cfgmm.DeviceBuffer.from_array(rhs)
    d_t = cfgmm.DeviceBuffer.from_array(tab)
    d_r = cfgmm.DeviceBuffer.from_array(r0)
    cfgmm.aodmao2rho_dev(s, d_r, d_d, d_l, d_rr, nao, ngrid, nao, ngrid, nset, d_t, nblk, blksize)
    d_b = cfgmm.DeviceBuffer.from_array(r0)
    cfgmm.aodmao2rho_batched_dev(s, [cfgmm.RhoProblem(d_b.ptr, d_d.ptr, d_l.ptr, d_rr.ptr, nao, ngrid, nao, ngrid, nset,
                                                      d_t.ptr, nblk, blksize)])
    d_z = cfgmm.DeviceBuffer.from_array(r0)
    h = cfgmm.RhoResident(s, d_d, d_l, d_rr, nao, ngrid, nao, ngrid, nset, d_t, nblk, blksize,
                          cfgmm.PREC_OZAKI_INT8, 6)
    h.eval(s, d_z, ngrid)
    s.synchronize()
    single = d_r.to_array(np.float64, r0.shape)
    batched = d_b.to_array(np.float64, r0.shape)
    res = d_z.to_array(np.float64, r0.shape)
    h.destroy()
    for b in (d_d, d_l, d_rr, d_t, d_r, d_b, d_z):
        b.free()
    s.destroy()
    if not tab.any():
        for out in (single, batched, res):
            assert np.array_equal(out, r0)
        return
    inc = ref - r0
    assert rel_fro(single - r0, inc) <= 1e-12
    assert np.array_equal(batched, single)
    assert rel_fro(res - r0, inc) <= 1e-10
