"""Pins the GTO tabulation oracle (oracle/gto.py) by mathematics: Gaussian normalisation integrals
by Gauss-Hermite quadrature, a known-answer component order, the trace identity xx+yy+zz = r^2 s,
translation invariance and a brute-force block mask."""
import math

import numpy as np

from oracle import gto


def test_component_order_known_answer():
    """d-shell at d = (1, 2, 3) with a unit radial part: xx, xy, xz, yy, yz, zz (PySCF order)."""
    phi = gto.eval_gto([((0.0, 0.0, 0.0), 2, 1, 0, 0)], np.array([0.0]), np.array([1.0]), [[1.0, 2.0, 3.0]], 6)
    assert np.array_equal(phi[:, 0], [1.0, 2.0, 3.0, 4.0, 6.0, 9.0])
    assert gto.cart_components(3) == [(3, 0, 0), (2, 1, 0), (2, 0, 1), (1, 2, 0), (1, 1, 1), (1, 0, 2),
                                      (0, 3, 0), (0, 2, 1), (0, 1, 2), (0, 0, 3)]


def _gauss_hermite_grid(alpha, n=40, center=(0.3, -0.2, 0.5)):
    """Product Gauss-Hermite nodes for integrals of f(r) exp(-2 alpha |r - R|^2) (weights absorb it)."""
    x, w = np.polynomial.hermite.hermgauss(n)
    s = 1.0 / math.sqrt(2.0 * alpha)
    pts = np.stack(np.meshgrid(x * s, x * s, x * s, indexing="ij"), -1).reshape(-1, 3) + np.array(center)
    wts = (w[:, None, None] * w[None, :, None] * w[None, None, :]).ravel() * s ** 3
    return pts, wts


def test_normalisation_integrals_closed_form():
    """int x^(2lx) y^(2ly) z^(2lz) exp(-2a r^2) = prod (2l-1)!! / (4a)^l * (pi / 2a)^(3/2)."""
    a = 0.8
    center = (0.3, -0.2, 0.5)
    pts, wts = _gauss_hermite_grid(a, center=center)
    for l in range(5):
        phi = gto.eval_gto([(center, l, 1, 0, 0)], np.array([a]), np.array([1.0]), pts, (l + 1) * (l + 2) // 2)
        d = pts - np.array(center)
        # the quadrature weight already holds exp(-2a r^2): divide it out of phi^2
        got = (phi ** 2 / np.exp(-2 * a * np.einsum("gi,gi->g", d, d))) @ wts
        for c, lxyz in enumerate(gto.cart_components(l)):
            dfact = np.prod([math.prod(range(2 * li - 1, 0, -2)) for li in lxyz])
            want = dfact / (4 * a) ** l * (math.pi / (2 * a)) ** 1.5
            assert abs(got[c] - want) <= 1e-12 * want


def test_trace_identity_and_translation():
    rng = np.random.default_rng(0)
    pts = rng.uniform(-2, 2, (300, 3))
    exps, coefs = np.array([0.5, 2.0]), np.array([0.7, 0.4])
    s = gto.eval_gto([((0.1, 0.2, 0.3), 0, 2, 0, 0)], exps, coefs, pts, 1)[0]
    d = gto.eval_gto([((0.1, 0.2, 0.3), 2, 2, 0, 0)], exps, coefs, pts, 6)
    r2 = np.sum((pts - np.array([0.1, 0.2, 0.3])) ** 2, axis=1)
    assert np.allclose(d[0] + d[3] + d[5], r2 * s, rtol=1e-13, atol=0)
    shift = np.array([5.0, -3.0, 1.5])
    d2 = gto.eval_gto([((5.1, -2.8, 1.8), 2, 2, 0, 0)], exps, coefs, pts + shift, 6)
    assert np.allclose(d2, d, rtol=1e-12, atol=1e-300)


def test_block_mask_brute_force():
    rng = np.random.default_rng(1)
    shells, exps, coefs, nao = gto.random_basis(rng, 12)
    ngrid, blksize, cutoff = 200, 16, 1e-3
    pts = rng.uniform(0, 6, (ngrid, 3))
    phi = gto.eval_gto(shells, exps, coefs, pts, nao)
    out, tab = gto.block_mask(phi, shells, ngrid, blksize, cutoff)
    for _, l, _, _, ao in shells:
        for c in range((l + 1) * (l + 2) // 2):
            for b in range(tab.shape[1]):
                vals = [abs(phi[ao + cc][g]) for cc in range((l + 1) * (l + 2) // 2)
                        for g in range(b * blksize, min((b + 1) * blksize, ngrid))]
                on = max(vals) > cutoff
                assert tab[ao + c, b] == on
                seg = out[ao + c, b * blksize:min((b + 1) * blksize, ngrid)]
                assert np.array_equal(seg, phi[ao + c, b * blksize:min((b + 1) * blksize, ngrid)] if on
                                      else np.zeros_like(seg))
