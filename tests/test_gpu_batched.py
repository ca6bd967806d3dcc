"""Batched rho (one persistent launch over many problems) through the C ABI.

Each problem of a batch must come out bitwise equal to a separate per-problem
call (every task writes its own partial slot and the reduction order is fixed),
for every setting of the batched-launch tunables, and equal to the CPU oracle
within the FP64 tolerance (relative Frobenius <= 1e-12)."""
import numpy as np
import pytest

import cfgmm
import oracle
from cfgmm import workloads as W
from helpers import rel_fro

pytestmark = pytest.mark.gpu


def _problems(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        nshell = int(rng.integers(20, 90))
        nao = 3 * nshell + 2 * int(rng.integers(0, 5))
        blksize = int(rng.choice([48, 64]))
        ngrid = int(blksize * rng.integers(1, 12) - rng.integers(0, 7))
        nset = int(rng.integers(1, 5))
        dens = float(rng.uniform(0.1, 0.6))
        p = W.rho_problem(seed=seed * 1000 + i, nshell=nshell, nao=nao, ngrid=ngrid, nset=nset, blksize=blksize,
                          density=dens)
        out.append(p)
    return out


class Uploaded:
    def __init__(self, p, rho0):
        self.p = p
        self.r = cfgmm.DeviceBuffer.from_array(rho0)
        self.d = cfgmm.DeviceBuffer.from_array(p.dms)
        self.phi = cfgmm.DeviceBuffer.from_array(p.phi)
        self.t = cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(p.tab))

    def desc(self):
        p = self.p
        return cfgmm.RhoProblem(self.r.ptr, self.d.ptr, self.phi.ptr, self.phi.ptr, p.nao, p.ngrid, p.nao, p.ngrid,
                                p.nset, self.t.ptr, p.nblk, p.blksize)

    def result(self):
        return self.r.to_array(np.float64, (self.p.nset, self.p.ngrid))


@pytest.fixture(scope="module")
def batch():
    return _problems(24, seed=11)


def _run_single(probs, rho0s, precision=cfgmm.PREC_FP64, ns=0):
    s = cfgmm.Stream()
    outs = []
    for p, r0 in zip(probs, rho0s):
        u = Uploaded(p, r0)
        cfgmm.aodmao2rho_dev(s, u.r, u.d, u.phi, u.phi, p.nao, p.ngrid, p.nao, p.ngrid, p.nset, u.t, p.nblk,
                             p.blksize, precision, ns)
        s.synchronize()
        outs.append(u.result())
    s.destroy()
    return outs


def _run_batched(probs, rho0s, precision=cfgmm.PREC_FP64, ns=0):
    s = cfgmm.Stream()
    ups = [Uploaded(p, r0) for p, r0 in zip(probs, rho0s)]
    cfgmm.aodmao2rho_batched_dev(s, [u.desc() for u in ups], precision, ns)
    s.synchronize()
    outs = [u.result() for u in ups]
    s.destroy()
    return outs


@pytest.mark.parametrize("wgs,sets", [(1, 1), (2, 2), (4, 4), (3, 2)])
def test_batched_equals_single_fp64(batch, wgs, sets):
    rng = np.random.default_rng(wgs * 10 + sets)
    rho0s = [rng.standard_normal(p.nset * p.ngrid) for p in batch]
    single = _run_single(batch, rho0s)
    old = (cfgmm.tune_get(cfgmm.TUNE_BATCH_WGS_PER_CU), cfgmm.tune_get(cfgmm.TUNE_BATCH_MAX_SETS))
    try:
        cfgmm.tune_set(cfgmm.TUNE_BATCH_WGS_PER_CU, wgs)
        cfgmm.tune_set(cfgmm.TUNE_BATCH_MAX_SETS, sets)
        batched = _run_batched(batch, rho0s)
    finally:
        cfgmm.tune_set(cfgmm.TUNE_BATCH_WGS_PER_CU, old[0])
        cfgmm.tune_set(cfgmm.TUNE_BATCH_MAX_SETS, old[1])
    for a, b in zip(single, batched):
        if sets >= 4 or all(p.nset <= sets for p in batch):
            assert np.array_equal(a, b)
        else:  # a different grouping of density matrices changes nothing but the task split
            assert rel_fro(b, a) <= 1e-15


def test_batched_vs_oracle(batch):
    outs = _run_batched(batch, [np.zeros(p.nset * p.ngrid) for p in batch])
    for p, got in zip(batch, outs):
        ref = np.zeros(p.nset * p.ngrid)
        assert oracle.aodmao2rho(ref, p.dms.ravel(), p.phi.ravel(), p.phi.ravel(), p.nao, p.ngrid, p.nao, p.ngrid,
                                 p.nset, np.ascontiguousarray(p.tab), p.nblk, p.blksize) == 0
        assert rel_fro(got, ref.reshape(p.nset, p.ngrid)) <= 1e-12


def test_batched_split_precision_equals_single(batch):
    sub = batch[:8]
    rho0s = [np.zeros(p.nset * p.ngrid) for p in sub]
    single = _run_single(sub, rho0s, cfgmm.PREC_OZAKI_INT8, 6)
    batched = _run_batched(sub, rho0s, cfgmm.PREC_OZAKI_INT8, 6)
    for a, b in zip(single, batched):
        assert np.array_equal(a, b)


def test_batched_form_choice_matches_single_calls():
    """D symmetric bitwise, symmetric only to rounding (two units in the last place on half the
    upper triangle), and asymmetric by 1e-13 relative in one entry, in one FP64 batch: each result
    bitwise that of its separate call, so the batch chose each problem's form as the per-call path
    does; the rounding-symmetric problem took the symmetric form (it differs from the general form)."""
    import dataclasses
    # lists longer than one 64-row chunk (with one chunk per block both forms compute the same)
    base = [W.rho_problem(seed=31 + i, nshell=150 + 20 * i, nao=460 + 60 * i, ngrid=64 * (5 + i) + 7 * i,
                          nset=2, blksize=64, density=0.4) for i in range(3)]
    rng = np.random.default_rng(5)
    near = base[1].dms.copy()
    iu = np.triu_indices(base[1].nao, 1)
    for s in range(base[1].nset):
        m = rng.random(len(iu[0])) < 0.5
        x = near[s][iu]
        near[s][iu] = np.where(m, np.nextafter(np.nextafter(x, np.inf), np.inf), x)
    far = base[2].dms.copy()
    far[0, 0, 1] *= 1.0 + 1e-13
    probs = [base[0], dataclasses.replace(base[1], dms=near), dataclasses.replace(base[2], dms=far)]
    rho0s = [np.zeros((p.nset, p.ngrid)) for p in probs]
    single = _run_single(probs, rho0s)
    bat = _run_batched(probs, rho0s)
    for a, b in zip(single, bat):
        assert np.array_equal(a, b)
    old = cfgmm.tune_get(cfgmm.TUNE_RHO_SYMMETRIC)
    try:
        cfgmm.tune_set(cfgmm.TUNE_RHO_SYMMETRIC, 1)
        general = _run_single(probs[1:2], rho0s[1:2])[0]
    finally:
        cfgmm.tune_set(cfgmm.TUNE_RHO_SYMMETRIC, old)
    assert not np.array_equal(single[1], general)
    assert rel_fro(single[1], general) <= 1e-12


def test_batched_empty_and_bad_problem():
    s = cfgmm.Stream()
    cfgmm.aodmao2rho_batched_dev(s, [])
    bad = cfgmm.RhoProblem(0, 0, 0, 0, 10, 10, 10, 10, 1, 0, 1, 40)  # blksize not a multiple of 16
    with pytest.raises(cfgmm.CfgmmError) as e:
        cfgmm.aodmao2rho_batched_dev(s, [bad])
    assert e.value.status == cfgmm.EINVAL
    s.destroy()


def test_tunables_validate():
    with pytest.raises(cfgmm.CfgmmError):
        cfgmm.tune_set(cfgmm.TUNE_BATCH_WGS_PER_CU, 0)
    with pytest.raises(cfgmm.CfgmmError):
        cfgmm.tune_set(99, 1)


def test_batched_overlapping_outputs_add_in_call_order():
    """Problems whose rho ranges overlap (three problems into one buffer, a fourth offset by half a
    row into it, a fifth elsewhere) add into rho in call order: bitwise the result of separate
    calls one after another."""
    probs = _problems(5, seed=23)
    # the same grid for all: every problem writes its nset rows of ngrid into a shared buffer
    ng = min(p.ngrid for p in probs)
    probs = [W.rho_problem(seed=900 + i, nshell=40, nao=120, ngrid=ng, nset=p.nset, blksize=p.blksize,
                           density=0.4) for i, p in enumerate(probs)]
    rng = np.random.default_rng(4)
    nrow = max(p.nset for p in probs)
    shared0 = rng.standard_normal(nrow * ng + ng)
    other0 = rng.standard_normal(nrow * ng)
    # (buffer, offset in elements) per problem
    where = [("shared", 0), ("other", 0), ("shared", 0), ("shared", ng // 2), ("shared", 0)]
    s = cfgmm.Stream()
    ups = [Uploaded(p, np.zeros(p.nset * p.ngrid)) for p in probs]

    def run(batched):
        bufs = {"shared": cfgmm.DeviceBuffer.from_array(shared0), "other": cfgmm.DeviceBuffer.from_array(other0)}
        descs = []
        for u, (name, off) in zip(ups, where):
            d = u.desc()
            d.rhos = bufs[name].ptr + 8 * off
            descs.append(d)
        if batched:
            cfgmm.aodmao2rho_batched_dev(s, descs)
        else:
            for d in descs:
                cfgmm.aodmao2rho_batched_dev(s, [d])
        s.synchronize()
        out = {k: b.to_array(np.float64, (b.nbytes // 8,)) for k, b in bufs.items()}
        for b in bufs.values():
            b.free()
        return out

    sep, bat = run(False), run(True)
    assert np.array_equal(sep["shared"], bat["shared"])
    assert np.array_equal(sep["other"], bat["other"])
    assert not np.array_equal(sep["shared"], shared0)
    s.destroy()
