"""The AO-grid-AO contraction after rho (cfgmm_aogridao2mat_dev) against numpy float64.

V[u][v] += sum_b sum_{g in block b} [u in L_b][v in L_b] lhs[u][g] w[g] rhs[v][g]: the reference has
no such function, so numpy (per-block products of the masked operands) is the check, to 1e-12
relative Frobenius; the mask is authoritative (huge values under off bits are ignored) and
repeated calls are bitwise identical."""
import numpy as np
import pytest

import cfgmm
from cfgmm import workloads as W
from helpers import rel_fro

pytestmark = pytest.mark.gpu


def vxc_numpy(lhs, rhs, w, tab, nao, ngrid, blksize):
    V = np.zeros((nao, nao))
    nblk = -(-ngrid // blksize)
    for b in range(nblk):
        cols = slice(b * blksize, min((b + 1) * blksize, ngrid))
        on = tab[:nao, b].astype(bool)
        A = np.where(on[:, None], lhs[:nao, cols] * (1.0 if w is None else w[cols]), 0.0)
        B = np.where(on[:, None], rhs[:nao, cols], 0.0)
        V += A @ B.T
    return V


def run(lhs, rhs, w, tab, nao, ngrid, blksize, ldv=None, V0=None, same=False):
    """same: lhs and rhs in one device buffer (the symmetric form: upper tiles of S_b only)."""
    ldv = nao if ldv is None else ldv
    stride = lhs.shape[1]
    V0 = np.zeros(nao * ldv) if V0 is None else V0
    s = cfgmm.Stream()
    bufs = [cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(x)) for x in (V0, lhs, rhs, tab)]
    if same:
        bufs[2].free()
        bufs[2] = bufs[1]
    d_w = None if w is None else cfgmm.DeviceBuffer.from_array(w)
    cfgmm.aogridao2mat_dev(s, bufs[0], ldv, bufs[1], bufs[2], d_w, stride, nao, ngrid, bufs[3], tab.shape[1], blksize)
    s.synchronize()
    out = bufs[0].to_array(np.float64, (nao, ldv))
    for b in {id(b): b for b in bufs}.values():
        b.free()
    if d_w is not None:
        d_w.free()
    s.destroy()
    return out


@pytest.mark.parametrize("blksize,ngrid", [(64, 64 * 30), (64, 1001), (96, 1000), (16, 1000)])
def test_vxc_symmetric_form(blksize, ngrid):
    """One buffer as lhs and rhs: only the upper tiles of each S_b are computed and mirrored.
    numpy to 1e-12, the general form (separate buffers) to 1e-14, V0 accumulated into, the padding
    of V untouched, bitwise repeatable; blocks of 64 points, a partial one (41 points), blocks of
    96 and 16."""
    p = W.rho_problem(seed=blksize + ngrid, nshell=230, nao=700, ngrid=ngrid, nset=1, blksize=blksize, density=0.4)
    rng = np.random.default_rng(5)
    w = rng.uniform(0.1, 2.0, p.ngrid)
    tab = np.ascontiguousarray(p.tab)
    ldv = p.nao + 3
    V0 = rng.standard_normal(p.nao * ldv)
    sym = run(p.phi, p.phi, w, tab, p.nao, p.ngrid, p.blksize, ldv=ldv, V0=V0.copy(), same=True)
    gen = run(p.phi, p.phi, w, tab, p.nao, p.ngrid, p.blksize, ldv=ldv, V0=V0.copy())
    ref = vxc_numpy(p.phi, p.phi, w, tab, p.nao, p.ngrid, p.blksize)
    V0m = V0.reshape(p.nao, ldv)
    assert rel_fro(sym[:, :p.nao] - V0m[:, :p.nao], ref) <= 1e-12
    assert rel_fro(sym[:, :p.nao] - V0m[:, :p.nao], gen[:, :p.nao] - V0m[:, :p.nao]) <= 1e-14
    assert np.array_equal(sym[:, p.nao:], V0m[:, p.nao:])
    again = run(p.phi, p.phi, w, tab, p.nao, p.ngrid, p.blksize, ldv=ldv, V0=V0.copy(), same=True)
    assert np.array_equal(sym, again)


def test_vxc_generated_molecule():
    p = W.rho_problem(seed=40, nshell=170, nao=520, ngrid=64 * 30, nset=1, density=0.35)
    rng = np.random.default_rng(2)
    w = rng.uniform(0.1, 2.0, p.ngrid)
    tab = np.ascontiguousarray(p.tab)
    got = run(p.phi, p.phi, w, tab, p.nao, p.ngrid, p.blksize)
    ref = vxc_numpy(p.phi, p.phi, w, tab, p.nao, p.ngrid, p.blksize)
    assert rel_fro(got, ref) <= 1e-12
    assert rel_fro(got, got.T) <= 1e-14  # lhs = rhs: symmetric up to rounding
    # weights on the other operand: lhs*w . rhs^T == (rhs*w . lhs^T)^T up to rounding
    got_t = run(p.phi * w, p.phi, None, tab, p.nao, p.ngrid, p.blksize)
    assert rel_fro(got_t, ref) <= 1e-12
    again = run(p.phi, p.phi, w, tab, p.nao, p.ngrid, p.blksize)
    assert np.array_equal(got, again)


@pytest.mark.parametrize("blksize,ngrid", [(16, 1000), (48, 1000), (64, 1000), (64, 1001), (96, 1000), (128, 1000)])
def test_vxc_strided_ragged_mask_authority(blksize, ngrid):
    """lhs != rhs, rows strided past ngrid, V padded and accumulated into, a partial last block
    (41 points for ngrid 1001: not a multiple of 4), blocks longer than 64 points, a periodic
    mask with huge values under its off bits, no weights."""
    nao = 333
    stride = ngrid + 7
    rng = np.random.default_rng(blksize + ngrid)
    nblk = -(-ngrid // blksize)
    tab = W.periodic_mask(nao, nblk + 2, W.MASK_30)
    lhs = rng.standard_normal((nao, stride))
    rhs = rng.standard_normal((nao, stride))
    on = np.repeat(tab[:, :nblk].astype(bool), blksize, axis=1)[:, :ngrid]
    lhs[:, :ngrid][~on] = 1e30
    rhs[:, :ngrid][~on] = -1e30
    ldv = nao + 5
    V0 = rng.standard_normal(nao * ldv)
    got = run(lhs, rhs, None, tab, nao, ngrid, blksize, ldv=ldv, V0=V0.copy())
    ref = vxc_numpy(np.where(np.pad(on, ((0, 0), (0, stride - ngrid))), lhs, 0.0),
                    np.where(np.pad(on, ((0, 0), (0, stride - ngrid))), rhs, 0.0), None, tab, nao, ngrid, blksize)
    G = got[:, :nao] - V0.reshape(nao, ldv)[:, :nao]
    assert rel_fro(G, ref) <= 1e-12
    assert np.array_equal(got[:, nao:], V0.reshape(nao, ldv)[:, nao:])  # V padding untouched


def test_vxc_empty_mask_and_errors():
    nao, ngrid = 40, 128
    tab = np.zeros((nao, 2), np.uint8)
    phi = np.ones((nao, ngrid))
    V0 = np.full(nao * nao, 3.0)
    got = run(phi, phi, None, tab, nao, ngrid, 64, V0=V0.copy())
    assert np.all(got == 3.0)
    s = cfgmm.Stream()
    d = cfgmm.DeviceBuffer(8 * nao * nao)
    with pytest.raises(cfgmm.CfgmmError) as e:
        cfgmm.aogridao2mat_dev(s, d, nao - 1, d, d, None, ngrid, nao, ngrid, d, 2, 64)  # ldv < nao
    assert e.value.status == cfgmm.EINVAL_DIMS
    d.free()
    s.destroy()
