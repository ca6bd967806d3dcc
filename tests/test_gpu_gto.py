"""On-device AO tabulation (cfgmm_eval_gto_dev) against the numpy oracle (oracle/gto.py), and
directly against mathematics: Gaussian normalisation integrals by Gauss-Hermite quadrature on
the device output; then the tabulated phi and mask feed rho on the device."""
import math

import numpy as np
import pytest

import cfgmm
import oracle
from oracle import gto
from helpers import rel_fro

pytestmark = pytest.mark.gpu


def tabulate(shells, exps, coefs, coords, nao, ngrid, blksize, cutoff, stride=None, nao_rows=None, ldtab=None):
    stride = ngrid if stride is None else stride
    rows = nao if nao_rows is None else nao_rows
    nblk = -(-ngrid // blksize)
    ldtab = nblk if ldtab is None else ldtab
    s = cfgmm.Stream()
    d_phi = cfgmm.DeviceBuffer.from_array(np.full((rows, stride), 7.0))
    d_tab = cfgmm.DeviceBuffer.from_array(np.full((rows, ldtab), 0xAB, np.uint8))
    d_xyz = cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(coords, dtype=np.float64))
    arr = cfgmm.gto_shells(*zip(*shells))
    cfgmm.eval_gto_dev(s, d_phi, stride, d_tab, ldtab, arr, exps, coefs, d_xyz, ngrid, rows, blksize, cutoff)
    s.synchronize()
    phi = d_phi.to_array(np.float64, (rows, stride))
    tab = d_tab.to_array(np.uint8, (rows, ldtab))
    for b in (d_phi, d_tab, d_xyz):
        b.free()
    s.destroy()
    return phi, tab


@pytest.mark.parametrize("blksize", [16, 64, 128])
def test_gto_vs_oracle_strided_ragged(blksize):
    """All l = 0..4 with 1-3 primitives, ngrid not a multiple of blksize, phi rows strided past
    ngrid, three spare AO rows owned by no shell and a wider mask: untouched by the call."""
    rng = np.random.default_rng(blksize)
    shells, exps, coefs, nao = gto.random_basis(rng, 40)
    ngrid, cutoff = 1000, 1e-3
    coords = rng.uniform(-1.0, 7.0, (ngrid, 3))
    phi, tab = tabulate(shells, exps, coefs, coords, nao, ngrid, blksize, cutoff, stride=ngrid + 5, nao_rows=nao + 3,
                        ldtab=-(-ngrid // blksize) + 2)
    ref_raw = gto.eval_gto(shells, exps, coefs, coords, nao)
    ref, ref_tab = gto.block_mask(ref_raw, shells, ngrid, blksize, cutoff)
    nblk = -(-ngrid // blksize)
    # a block whose maximum sits within rounding of the cutoff may fall either way
    tie = np.zeros_like(ref_tab, dtype=bool)
    for _, l, _, _, ao in shells:
        nf = (l + 1) * (l + 2) // 2
        for b in range(nblk):
            m = np.max(np.abs(ref_raw[ao:ao + nf, b * blksize:(b + 1) * blksize]))
            tie[ao:ao + nf, b] = abs(m - cutoff) <= 1e-12 * cutoff
    agree = (tab[:nao, :nblk] == ref_tab) | tie
    assert agree.all()
    on = np.repeat(tab[:nao, :nblk].astype(bool), blksize, axis=1)[:, :ngrid]
    assert rel_fro(np.where(on, phi[:nao, :ngrid], 0.0), np.where(on, ref_raw, 0.0)) <= 1e-13
    assert np.all(phi[:nao, :ngrid][~on] == 0.0)  # zeroed under off bits
    assert np.all(phi[:nao, ngrid:] == 7.0) and np.all(phi[nao:] == 7.0)
    assert np.all(tab[:nao, nblk:] == 0xAB) and np.all(tab[nao:] == 0xAB)
    assert 0.2 < ref_tab.mean() < 0.98  # a mask with both kinds of block


def block_maxima(raw, shells, ngrid, blksize):
    nblk = -(-ngrid // blksize)
    out = []
    for _, l, _, _, ao in shells:
        nf = (l + 1) * (l + 2) // 2
        out.append([np.max(np.abs(raw[ao:ao + nf, b * blksize:(b + 1) * blksize])) for b in range(nblk)])
    return np.array(out)


@pytest.mark.parametrize("blksize", [16, 64, 96])
def test_gto_clustered_blocks_cutoff_sweep(blksize):
    """Blocks of points clustered as on a molecular grid, so that most (block, shell) pairs are far
    apart and can be skipped without evaluation; cutoffs at quantiles of the block maxima put many
    pairs near the threshold.  Every tab bit and every value equals the oracle's.  The basis holds
    shells of l = 0..4, negative coefficients, a shell of 70 primitives (more than one group's
    record holds), shells of 30 primitives (groups cut short by primitives) and a shell with a zero
    exponent (no decay: never skipped)."""
    rng = np.random.default_rng(100 + blksize)
    shells, exps, coefs, nao = gto.random_basis(rng, 48, box=12.0)
    exps, coefs = list(exps), list(coefs)
    coefs = [c * (-1.0 if i % 5 == 0 else 1.0) for i, c in enumerate(coefs)]
    extra = [(70, 4), (30, 1), (30, 2), (30, 0), (1, 1)]
    for n, l in extra:
        shells.append((rng.uniform(0.0, 12.0, 3), l, n, len(exps), nao))
        e = list(10.0 ** rng.uniform(-1.0, 1.0, n))
        if n == 1:
            e = [0.0]
        exps.extend(e)
        coefs.extend(rng.uniform(0.2, 1.5, n) * (1e-6 if n == 1 else 1.0))
        nao += (l + 1) * (l + 2) // 2
    exps, coefs = np.array(exps), np.array(coefs)
    ngrid = blksize * 23 - 5
    nblk = -(-ngrid // blksize)
    centres = rng.uniform(0.0, 12.0, (nblk, 3))
    coords = centres[np.arange(ngrid) // blksize] + rng.uniform(-0.6, 0.6, (ngrid, 3))
    raw = gto.eval_gto(shells, exps, coefs, coords, nao)
    maxima = block_maxima(raw, shells, ngrid, blksize)
    srt = np.sort(maxima[maxima > 0])
    for q in (0.5, 0.7, 0.9):
        i = int(q * len(srt))
        cutoff = float(np.sqrt(srt[i] * srt[i + 1]))  # between two block maxima: no ties
        assert np.min(np.abs(maxima - cutoff)) > 1e-9 * cutoff
        phi, tab = tabulate(shells, exps, coefs, coords, nao, ngrid, blksize, cutoff)
        ref, ref_tab = gto.block_mask(raw, shells, ngrid, blksize, cutoff)
        assert np.array_equal(tab, ref_tab), f"cutoff quantile {q}"
        assert rel_fro(phi, ref) <= 1e-13
        assert np.all(phi[~np.repeat(ref_tab.astype(bool), blksize, axis=1)[:, :ngrid]] == 0.0)
        assert 0.4 < 1.0 - ref_tab.mean() < 0.97  # mostly off: screening has work to skip


def test_gto_normalisation_integrals_on_device():
    """int phi_c^2 = prod (2l-1)!! / (4a)^l (pi/2a)^(3/2) for one primitive of each l <= 4, by
    Gauss-Hermite quadrature of the device values."""
    a, center = 1.3, (0.25, -0.4, 0.1)
    x, w = np.polynomial.hermite.hermgauss(40)
    s = 1.0 / math.sqrt(2.0 * a)
    pts = np.stack(np.meshgrid(x * s, x * s, x * s, indexing="ij"), -1).reshape(-1, 3) + np.array(center)
    wts = (w[:, None, None] * w[None, :, None] * w[None, None, :]).ravel() * s ** 3
    shells, ao = [], 0
    for l in range(5):
        shells.append((center, l, 1, 0, ao))
        ao += (l + 1) * (l + 2) // 2
    phi, _ = tabulate(shells, np.array([a]), np.array([1.0]), pts, ao, len(pts), 64, 0.0)
    d = pts - np.array(center)
    e = np.exp(-2 * a * np.einsum("gi,gi->g", d, d))
    for _, l, _, _, off in shells:
        for c, lxyz in enumerate(gto.cart_components(l)):
            got = (phi[off + c] ** 2 / e) @ wts
            want = np.prod([math.prod(range(2 * li - 1, 0, -2)) for li in lxyz]) / (4 * a) ** l * (math.pi / (2 * a)) ** 1.5
            assert abs(got - want) <= 1e-12 * want


def test_gto_then_rho_on_device():
    """Tabulated phi and mask feed cfgmm_aodmao2rho_dev without leaving the device; the result
    matches the oracle rho of the oracle's phi and mask."""
    rng = np.random.default_rng(3)
    shells, exps, coefs, nao = gto.random_basis(rng, 60, box=8.0, lmax=2)
    ngrid, blksize, nset = 64 * 20, 64, 2
    coords = rng.uniform(0.0, 8.0, (ngrid, 3))
    C = rng.standard_normal((nset, nao, 30))
    dms = np.einsum("sik,sjk->sij", C, C) / 30
    s = cfgmm.Stream()
    nblk = ngrid // blksize
    d_phi = cfgmm.DeviceBuffer(8 * nao * ngrid)
    d_tab = cfgmm.DeviceBuffer(nao * nblk)
    d_xyz = cfgmm.DeviceBuffer.from_array(coords)
    cfgmm.eval_gto_dev(s, d_phi, ngrid, d_tab, nblk, cfgmm.gto_shells(*zip(*shells)), exps, coefs, d_xyz, ngrid, nao,
                       blksize, 1e-5)
    d_d = cfgmm.DeviceBuffer.from_array(dms)
    d_r = cfgmm.DeviceBuffer(8 * nset * ngrid)
    d_r.zero()
    cfgmm.aodmao2rho_dev(s, d_r, d_d, d_phi, d_phi, nao, ngrid, nao, ngrid, nset, d_tab, nblk, blksize)
    s.synchronize()
    got = d_r.to_array(np.float64, (nset * ngrid,))
    tab_dev = d_tab.to_array(np.uint8, (nao, nblk))
    for b in (d_phi, d_tab, d_xyz, d_d, d_r):
        b.free()
    s.destroy()
    ref_phi, ref_tab = gto.block_mask(gto.eval_gto(shells, exps, coefs, coords, nao), shells, ngrid, blksize, 1e-5)
    assert np.array_equal(tab_dev, ref_tab)
    ref = np.zeros(nset * ngrid)
    assert oracle.aodmao2rho(ref, dms.ravel(), ref_phi.ravel(), ref_phi.ravel(), nao, ngrid, nao, ngrid, nset,
                             np.ascontiguousarray(ref_tab), nblk, blksize) == 0
    assert rel_fro(got, ref) <= 1e-12


def test_gto_errors():
    s = cfgmm.Stream()
    d = cfgmm.DeviceBuffer(8 * 64 * 64)
    arr = cfgmm.gto_shells([(0.0, 0.0, 0.0)], [5], [1], [0], [0])
    with pytest.raises(cfgmm.CfgmmError) as e:
        cfgmm.eval_gto_dev(s, d, 64, d, 1, arr, np.array([1.0]), np.array([1.0]), d, 64, 30, 64, 0.0)
    assert e.value.status == cfgmm.EINVAL
    arr = cfgmm.gto_shells([(0.0, 0.0, 0.0)], [2], [1], [0], [0])
    with pytest.raises(cfgmm.CfgmmError) as e:  # six AOs do not fit nao = 5
        cfgmm.eval_gto_dev(s, d, 64, d, 1, arr, np.array([1.0]), np.array([1.0]), d, 64, 5, 64, 0.0)
    assert e.value.status == cfgmm.EINVAL
    with pytest.raises(cfgmm.CfgmmError) as e:  # ldtab 1 < 2 blocks
        cfgmm.eval_gto_dev(s, d, 128, d, 1, arr, np.array([1.0]), np.array([1.0]), d, 128, 6, 64, 0.0)
    assert e.value.status == cfgmm.EINVAL_DIMS
    d.free()
    s.destroy()
