"""The allocation-free rho entry point (cfgmm_aodmao2rho_dev_noalloc) against cfgmm_aodmao2rho_dev.

Both run the same kernels on the same plan, so the results must be bitwise equal; the first
call is also held to the oracle at the FP64 bar (relative Frobenius <= 1e-12).  Two calls in a
row, the second on a smaller problem, show that what the shared scratch keeps from the last
call does not leak into the next; a problem larger than the scratch is refused before any
launch."""
import numpy as np
import pytest

import cfgmm
import oracle
from cfgmm import workloads as W
from helpers import rel_fro

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _both(s, p, lhs_is_rhs=True):
    """(rho from cfgmm_aodmao2rho_dev, rho from the allocation-free entry point), both starting at
    the same nonzero rho."""
    rng = np.random.default_rng(p.nao)
    rho0 = rng.standard_normal(p.nset * p.ngrid)
    d_d, d_phi = cfgmm.DeviceBuffer.from_array(p.dms), cfgmm.DeviceBuffer.from_array(p.phi)
    d_rhs = d_phi if lhs_is_rhs else cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(p.phi[::-1]))
    d_t = cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(p.tab))
    outs = []
    for fn in (cfgmm.lib().cfgmm_aodmao2rho_dev, cfgmm.lib().cfgmm_aodmao2rho_dev_noalloc):
        d_r = cfgmm.DeviceBuffer.from_array(rho0)
        cfgmm.check(fn(s.handle, d_r.ptr, d_d.ptr, d_phi.ptr, d_rhs.ptr, p.nao, p.ngrid, p.nao, p.ngrid, p.nset,
                       d_t.ptr, p.nblk, p.blksize))
        s.synchronize()
        outs.append(d_r.to_array(np.float64, (p.nset * p.ngrid,)))
        d_r.free()
    for b in {d_d, d_phi, d_rhs, d_t}:
        b.free()
    return rho0, outs[0], outs[1]


def test_noalloc_equals_dev_twice_shrinking():
    s = cfgmm.Stream()
    big = W.rho_problem(seed=0)  # nao 2228, ngrid 16384, nset 3, 30 %
    small = W.rho_problem(seed=5, nshell=150, nao=450, ngrid=64 * 20 + 17, nset=2, blksize=48, density=0.35)
    for i, p in enumerate((big, small)):
        rho0, ref_dev, got = _both(s, p)
        print(f"call {i}: nao={p.nao} ngrid={p.ngrid} nset={p.nset} bitwise={np.array_equal(got, ref_dev)}",
              flush=True)
        assert np.array_equal(got, ref_dev)
        want = rho0.copy()
        assert oracle.aodmao2rho(want, p.dms.ravel(), p.phi.ravel(), p.phi.ravel(), p.nao, p.ngrid, p.nao, p.ngrid,
                                 p.nset, np.ascontiguousarray(p.tab), p.nblk, p.blksize) == 0
        assert rel_fro(got - rho0, want - rho0) <= TOL
    s.destroy()


def test_noalloc_general_form_after_symmetric():
    """lhs != rhs (the general form) right after a symmetric-form call: equal to the per-stream path."""
    s = cfgmm.Stream()
    p = W.rho_problem(seed=8, nshell=200, nao=600, ngrid=64 * 30, nset=2, density=0.3)
    for lhs_is_rhs in (True, False):
        _, ref_dev, got = _both(s, p, lhs_is_rhs)
        assert np.array_equal(got, ref_dev)
    s.destroy()


def test_noalloc_rejects_problem_beyond_scratch():
    """nao 4000, ngrid 65536, nset 4 needs about 136 MB of partials: refused, nothing launched."""
    s = cfgmm.Stream()
    buf = cfgmm.DeviceBuffer(8)
    nao, ngrid, nset = 4000, 65536, 4
    st = cfgmm.lib().cfgmm_aodmao2rho_dev_noalloc(s.handle, buf.ptr, buf.ptr, buf.ptr, buf.ptr, nao, ngrid, nao,
                                                  ngrid, nset, buf.ptr, ngrid // 64, 64)
    assert st == cfgmm.EINVAL
    assert b"scratch" in cfgmm.lib().cfgmm_last_error()
    buf.free()
    s.destroy()
