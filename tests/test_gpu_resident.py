"""Resident rho operands (cfgmm_rho_resident_*): split precision over per-block compacted tiles.

Held to the CPU oracle normwise (relative Frobenius over rho) at 1e-10 with 6
int8 slices, as the per-call split-precision path; evaluations of leading grid
blocks / density matrices must equal the per-call FP64 result on that range
within the same tolerance and leave everything else untouched; repeated and
concurrent evaluations are bitwise identical (fixed reduction order)."""
import numpy as np
import pytest

import cfgmm
import oracle
from cfgmm import workloads as W
from helpers import rel_fro

pytestmark = pytest.mark.gpu


def _oracle(p, rhos0=None, lhs=None, rhs=None, dms=None):
    lhs = p.phi if lhs is None else lhs
    rhs = p.phi if rhs is None else rhs
    dms = p.dms if dms is None else dms
    out = np.zeros(p.nset * p.ngrid) if rhos0 is None else rhos0.copy()
    assert oracle.aodmao2rho(out, np.ascontiguousarray(dms).ravel(), np.ascontiguousarray(lhs).ravel(),
                             np.ascontiguousarray(rhs).ravel(), p.nao, p.ngrid, p.nao, p.ngrid, p.nset,
                             np.ascontiguousarray(p.tab), p.nblk, p.blksize) == 0
    return out.reshape(p.nset, p.ngrid)


class Resident:
    """same: lhs and rhs in one device buffer (the symmetric form when D is symmetric)."""

    def __init__(self, p, ns=6, lhs=None, rhs=None, dms=None, same=False):
        self.s = cfgmm.Stream()
        self.bufs = [cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(x)) for x in
                     (p.dms if dms is None else dms, p.phi if lhs is None else lhs, p.phi if rhs is None else rhs,
                      p.tab)]
        if same:
            self.bufs[2].free()
            self.bufs[2] = self.bufs[1]
        d, l, r, t = self.bufs
        self.h = cfgmm.RhoResident(self.s, d, l, r, p.nao, p.ngrid, p.nao, p.ngrid, p.nset, t, p.nblk, p.blksize,
                                   cfgmm.PREC_OZAKI_INT8, ns)
        self.p = p

    def eval(self, rhos0=None, stride=None, ngrid=None, nset=None, stream=None):
        p = self.p
        stride = p.ngrid if stride is None else stride
        r0 = np.zeros(p.nset * stride) if rhos0 is None else rhos0
        d_r = cfgmm.DeviceBuffer.from_array(r0)
        s = self.s if stream is None else stream
        self.h.eval(s, d_r, stride, ngrid, nset)
        s.synchronize()
        out = d_r.to_array(np.float64, r0.shape)
        d_r.free()
        return out

    def close(self):
        self.h.destroy()
        for b in {id(b): b for b in self.bufs}.values():
            b.free()
        self.s.destroy()


@pytest.mark.parametrize("ns,same", [(6, False), (5, True)])
def test_resident_full_workload(ns, same):
    """The full workload with six slices (general form) and with five (symmetric form)."""
    p = W.rho_problem(seed=0)
    ref = _oracle(p)
    r = Resident(p, ns=ns, same=same)
    got = r.eval().reshape(p.nset, p.ngrid)
    again = r.eval().reshape(p.nset, p.ngrid)
    nbytes = r.h.bytes
    r.close()
    err = rel_fro(got, ref)
    print(f"resident ns={ns} rel_fro {err:.3e}, {nbytes / 2**30:.2f} GiB")
    assert err <= 1e-10
    assert np.array_equal(got, again)


@pytest.mark.parametrize("ns,tol", [(4, 1e-7), (5, 1e-10)])
def test_resident_fewer_slices(ns, tol):
    """Balanced digits (about 8 bits per int8 slice): five slices within 1e-10, four within 1e-7."""
    p = W.rho_problem(seed=3, nshell=200, nao=640, ngrid=2048, nset=2, density=0.3)
    r = Resident(p, ns=ns)
    got = r.eval().reshape(p.nset, p.ngrid)
    r.close()
    assert rel_fro(got, _oracle(p)) <= tol


def test_resident_wide_range_dm_and_accumulate():
    """D magnitudes 10^U(-8, 2); rho accumulates into a nonzero start value."""
    p = W.rho_problem(seed=7, nshell=200, nao=640, ngrid=2048, nset=2, density=0.3, dm_logspan=(-8, 2))
    rng = np.random.default_rng(1)
    r0 = rng.standard_normal(p.nset * p.ngrid)
    ref = _oracle(p, rhos0=r0)
    r = Resident(p)
    got = r.eval(rhos0=r0).reshape(p.nset, p.ngrid)
    r.close()
    assert rel_fro(got - r0.reshape(p.nset, p.ngrid), ref - r0.reshape(p.nset, p.ngrid)) <= 1e-10


def test_resident_leading_range_and_stride():
    """Leading 12 of 20 blocks and 2 of 3 sets, rho rows strided by 1400 > ngrid: the rest is untouched."""
    p = W.rho_problem(seed=5, nshell=160, nao=520, ngrid=1280, nset=3, density=0.4)
    stride = 1400
    r0 = np.full(p.nset * stride, 7.0)
    r = Resident(p)
    got = r.eval(rhos0=r0.copy(), stride=stride, ngrid=12 * 64, nset=2).reshape(p.nset, stride)
    r.close()
    ref = _oracle(p).reshape(p.nset, p.ngrid)
    assert rel_fro(got[:2, :768] - 7.0, ref[:2, :768]) <= 1e-10
    assert np.all(got[:2, 768:] == 7.0) and np.all(got[2] == 7.0)


def test_resident_blksize48_partial_block_lhs_ne_rhs():
    """blksize 48 (three 16-column tiles), ngrid not a multiple of it, lhs != rhs, mask authority."""
    p = W.rho_problem(seed=9, nshell=120, nao=400, ngrid=1000, nset=2, blksize=48, density=0.35)
    rng = np.random.default_rng(4)
    lhs = p.phi * rng.uniform(0.5, 1.5, size=p.phi.shape)
    on = np.repeat(p.tab.astype(bool), p.blksize, axis=1)[:, : p.ngrid]
    lhs[~on] = 1e30  # under off bits: must be ignored
    rhs = p.phi.copy()
    rhs[~on] = -1e30
    ref = _oracle(p, lhs=np.where(on, lhs, 0.0), rhs=np.where(on, rhs, 0.0))
    r = Resident(p, lhs=lhs, rhs=rhs)
    got = r.eval().reshape(p.nset, p.ngrid)
    r.close()
    assert rel_fro(got, ref) <= 1e-10


def test_resident_concurrent_streams_bitwise():
    p = W.rho_problem(seed=6, nshell=200, nao=640, ngrid=4096, nset=2, density=0.3)
    r = Resident(p)
    s2 = cfgmm.Stream()
    a = r.eval()
    b = r.eval(stream=s2)
    r.close()
    s2.destroy()
    assert np.array_equal(a, b)


def test_resident_golden_fixture():
    import os
    f = np.load(os.path.join(os.path.dirname(__file__), "golden", "rho_sincos.npz"), allow_pickle=False)
    sa, sg, nao, ngrid, nset, ldtab, blksize = (int(x) for x in f["dims"])
    s = cfgmm.Stream()
    bufs = {k: cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(f[k])) for k in ("rhos0", "dms", "lhs", "rhs", "tab")}
    h = cfgmm.RhoResident(s, bufs["dms"], bufs["lhs"], bufs["rhs"], sa, sg, nao, ngrid, nset, bufs["tab"], ldtab,
                          blksize)
    h.eval(s, bufs["rhos0"], sg)
    s.synchronize()
    got = bufs["rhos0"].to_array(np.float64, f["rhos0"].shape)
    h.destroy()
    for b in bufs.values():
        b.free()
    s.destroy()
    assert rel_fro(got, f["rhos"]) <= 1e-10


def test_resident_errors():
    p = W.rho_problem(seed=1, nshell=40, nao=128, ngrid=256, nset=1, density=0.5)
    s = cfgmm.Stream()
    d, l, t = (cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(x)) for x in (p.dms, p.phi, p.tab))
    with pytest.raises(cfgmm.CfgmmError) as e:
        cfgmm.RhoResident(s, d, l, l, p.nao, p.ngrid, p.nao, p.ngrid, p.nset, t, p.nblk, p.blksize,
                          cfgmm.PREC_FP64, 0)
    assert e.value.status == cfgmm.EINVAL
    h = cfgmm.RhoResident(s, d, l, l, p.nao, p.ngrid, p.nao, p.ngrid, p.nset, t, p.nblk, p.blksize)
    out = cfgmm.DeviceBuffer(8 * p.ngrid)
    with pytest.raises(cfgmm.CfgmmError) as e:
        h.eval(s, out, p.ngrid, 100, 1)  # not a multiple of blksize
    assert e.value.status == cfgmm.EINVAL_DIMS
    with pytest.raises(cfgmm.CfgmmError) as e:
        h.eval(s, out, p.ngrid, p.ngrid, 2)  # more sets than resident
    assert e.value.status == cfgmm.EINVAL_DIMS
    h.destroy()
    for b in (d, l, t, out):
        b.free()
    s.destroy()


@pytest.mark.parametrize("depth,cols", [(2, 0), (1, 32), (1, 16)])
def test_resident_launch_tunables_bitwise(depth, cols):
    """Prefetch depth and columns per workgroup change the schedule only: results are bitwise equal."""
    p = W.rho_problem(seed=8, nshell=200, nao=640, ngrid=2048, nset=2, density=0.35)
    base = Resident(p)
    ref = base.eval()
    base.close()
    old = (cfgmm.tune_get(cfgmm.TUNE_RESIDENT_DEPTH), cfgmm.tune_get(cfgmm.TUNE_RESIDENT_COLS))
    try:
        cfgmm.tune_set(cfgmm.TUNE_RESIDENT_DEPTH, depth)
        cfgmm.tune_set(cfgmm.TUNE_RESIDENT_COLS, cols)
        r = Resident(p)  # the column width is fixed at creation
        got = r.eval()
        r.close()
    finally:
        cfgmm.tune_set(cfgmm.TUNE_RESIDENT_DEPTH, old[0])
        cfgmm.tune_set(cfgmm.TUNE_RESIDENT_COLS, old[1])
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("same", [True, False])
def test_resident_persistent_many_items(same):
    """2048 blocks of 16 points and five density matrices: 40 (symmetric form) or 60 (general form)
    items per workgroup of the persistent evaluation, more than its 32 item records staged in LDS at
    a time.  Held to the oracle, and bitwise equal to one workgroup per item, at both prefetch depths.
    A leading range of 3/4 of the grid and 4 of the 5 density matrices
    (60 % of the items: the persistent kernel, skipping the others) against the per-item kernel
    bitwise, the rest of rho untouched."""
    p = W.rho_problem(seed=12, nshell=200, nao=600, ngrid=32768, nset=5, blksize=16, density=0.3)
    ref = _oracle(p)
    r = Resident(p, ns=5, same=same)
    ng, ns = 3 * p.ngrid // 4, 4
    old = (cfgmm.tune_get(cfgmm.TUNE_RESIDENT_PERSISTENT), cfgmm.tune_get(cfgmm.TUNE_RESIDENT_DEPTH))
    outs, parts = {}, {}
    try:
        for pers, depth in ((0, 1), (0, 2), (1, 1)):
            cfgmm.tune_set(cfgmm.TUNE_RESIDENT_PERSISTENT, pers)
            cfgmm.tune_set(cfgmm.TUNE_RESIDENT_DEPTH, depth)
            outs[(pers, depth)] = r.eval().reshape(p.nset, p.ngrid)
            parts[(pers, depth)] = r.eval(ngrid=ng, nset=ns).reshape(p.nset, p.ngrid)
    finally:
        cfgmm.tune_set(cfgmm.TUNE_RESIDENT_PERSISTENT, old[0])
        cfgmm.tune_set(cfgmm.TUNE_RESIDENT_DEPTH, old[1])
        r.close()
    got = outs[(0, 1)]
    err = rel_fro(got, ref)
    print(f"persistent resident, same={same}: rel_fro {err:.3e}")
    assert err <= 1e-10
    assert np.array_equal(outs[(0, 2)], got) and np.array_equal(outs[(1, 1)], got)
    part = parts[(0, 1)]
    assert rel_fro(part[:ns, :ng], ref[:ns, :ng]) <= 1e-10
    assert np.all(part[:, ng:] == 0.0) and np.all(part[ns:] == 0.0)
    assert np.array_equal(parts[(0, 2)], part) and np.array_equal(parts[(1, 1)], part)


def test_resident_eval_batched_bitwise():
    """One batched launch over several handles (different sizes, leading ranges, strides, slices)
    equals separate evaluations bitwise; empty entries are no-ops."""
    probs = [W.rho_problem(seed=30 + i, nshell=90 + 40 * i, nao=270 + 120 * i, ngrid=64 * (12 + 7 * i), nset=1 + i % 3,
                           density=0.25 + 0.1 * i) for i in range(3)]
    rs = [Resident(p, ns=6 if i != 1 else 5) for i, p in enumerate(probs)]
    jobs = [(0, probs[0].ngrid, probs[0].nset, probs[0].ngrid), (1, 64 * 5, 1, 64 * 5 + 3), (2, probs[2].ngrid, 2, probs[2].ngrid),
            (0, 0, 1, 64), (2, 64 * 3, probs[2].nset, 64 * 3)]
    refs, bufs, entries = [], [], []
    for k, (i, ng, ns, stride) in enumerate(jobs):
        r = rs[i]
        if ng:
            refs.append(r.eval(rhos0=np.zeros(r.p.nset * stride) + k, stride=stride, ngrid=ng, nset=ns))
        else:
            refs.append(np.zeros(r.p.nset * stride) + k)
        buf = cfgmm.DeviceBuffer.from_array(np.zeros(r.p.nset * stride) + k)
        bufs.append(buf)
        entries.append((r.h, buf, stride, ng, ns))
    s = cfgmm.Stream()
    cfgmm.rho_resident_eval_batched(s, entries)
    s.synchronize()
    outs = [b.to_array(np.float64, ref.shape) for b, ref in zip(bufs, refs)]
    for b in bufs:
        b.free()
    for r in rs:
        r.close()
    s.destroy()
    for o, ref in zip(outs, refs):
        assert np.array_equal(o, ref)


def test_resident_eval_batched_persistent_plans():
    """Five-slice handles in both forms in one batched launch (the persistent batch: records of
    every evaluation dealt at the call, the plan kept for the same batch): bitwise equal to
    separate evaluations and to the batch of one workgroup per item, on the first call and on
    repeats (the kept plan), over full and leading ranges with padded strides, and for more
    distinct batches than plans are kept."""
    probs = [W.rho_problem(seed=50 + i, nshell=100 + 50 * i, nao=300 + 150 * i, ngrid=64 * (10 + 9 * i), nset=1 + i % 3,
                           density=0.2 + 0.1 * i) for i in range(4)]
    rs = [Resident(p, ns=5, same=(i % 2 == 0)) for i, p in enumerate(probs)]
    jobs = [(0, probs[0].ngrid, probs[0].nset, probs[0].ngrid), (1, 64 * 7, 1, 64 * 7 + 5),
            (2, probs[2].ngrid, probs[2].nset, probs[2].ngrid + 3), (3, 64 * 20, 1, 64 * 21)]
    refs = [rs[i].eval(rhos0=np.zeros(rs[i].p.nset * st) + k, stride=st, ngrid=ng, nset=ns)
            for k, (i, ng, ns, st) in enumerate(jobs)]
    s = cfgmm.Stream()

    def run(sel):
        bufs = [cfgmm.DeviceBuffer.from_array(np.zeros(rs[jobs[k][0]].p.nset * jobs[k][3]) + k) for k in sel]
        cfgmm.rho_resident_eval_batched(s, [(rs[jobs[k][0]].h, b, jobs[k][3], jobs[k][1], jobs[k][2])
                                            for k, b in zip(sel, bufs)])
        s.synchronize()
        outs = [b.to_array(np.float64, refs[k].shape) for k, b in zip(sel, bufs)]
        for b in bufs:
            b.free()
        return outs

    old = cfgmm.tune_get(cfgmm.TUNE_RESIDENT_PERSISTENT)
    try:
        everything = list(range(len(jobs)))
        for outs in (run(everything), run(everything)):  # built, then kept
            for k, o in zip(everything, outs):
                assert np.array_equal(o, refs[k])
        for sel in ([0, 2], [1, 3], [3], [2, 1, 0], [0, 1, 3], everything):  # past the kept plans
            for k, o in zip(sel, run(sel)):
                assert np.array_equal(o, refs[k])
        cfgmm.tune_set(cfgmm.TUNE_RESIDENT_PERSISTENT, 1)
        for k, o in zip(everything, run(everything)):
            assert np.array_equal(o, refs[k])
    finally:
        cfgmm.tune_set(cfgmm.TUNE_RESIDENT_PERSISTENT, old)
        for r in rs:
            r.close()
        s.destroy()
    for k, (i, ng, ns, st) in enumerate(jobs):
        p = probs[i]
        ref = _oracle(p).reshape(p.nset, p.ngrid)
        got = refs[k].reshape(p.nset, st)
        assert rel_fro(got[:ns, :ng] - k, ref[:ns, :ng]) <= 1e-10


def test_resident_eval_batched_plans_two_threads():
    """Two host threads, each on its own stream, run batched evaluations over more distinct batches
    than plans are kept (plans built, kept and evicted while the other thread launches): every
    result bitwise equal to the separate evaluation."""
    import threading
    probs = [W.rho_problem(seed=70 + i, nshell=80 + 30 * i, nao=240 + 90 * i, ngrid=64 * (6 + 5 * i), nset=1 + i % 2,
                           density=0.3) for i in range(4)]
    rs = [Resident(p, ns=5, same=(i % 2 == 1)) for i, p in enumerate(probs)]
    refs = [r.eval() for r in rs]
    sels = [[0, 1], [2, 3], [1, 2], [0, 3], [3], [0, 1, 2, 3], [2], [1, 3, 0]]
    errors = []

    def worker(order):
        s = cfgmm.Stream()
        try:
            for _ in range(3):
                for sel in order:
                    bufs = [cfgmm.DeviceBuffer(8 * rs[k].p.nset * rs[k].p.ngrid) for k in sel]
                    for b in bufs:
                        b.zero()
                    cfgmm.rho_resident_eval_batched(s, [(rs[k].h, b, rs[k].p.ngrid, rs[k].p.ngrid, rs[k].p.nset)
                                                        for k, b in zip(sel, bufs)])
                    s.synchronize()
                    for k, b in zip(sel, bufs):
                        if not np.array_equal(b.to_array(np.float64, refs[k].shape), refs[k]):
                            errors.append((sel, k))
                        b.free()
        except Exception as e:  # reported below, on the main thread
            errors.append(repr(e))
        finally:
            s.destroy()

    threads = [threading.Thread(target=worker, args=(order,)) for order in (sels, sels[::-1])]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for r in rs:
        r.close()
    assert not errors, errors


def test_resident_symmetric_form():
    """Symmetric D with lhs and rhs one buffer: the upper-triangle tiles only (about half the bytes),
    the oracle to 1e-10 over a partial last block, a leading range, both prefetch depths (bitwise
    equal) and a batched launch beside a general-form handle (bitwise equal to separate calls).
    D magnitudes 10^U(-8, 2) too.  One asymmetric entry keeps the general form."""
    p = W.rho_problem(seed=41, nshell=240, nao=760, ngrid=64 * 30 + 20, nset=3, density=0.35)
    ref = _oracle(p)
    rs, rg = Resident(p, same=True), Resident(p)
    assert rs.h.bytes < 0.75 * rg.h.bytes  # the symmetric layout was taken
    got = rs.eval().reshape(p.nset, p.ngrid)
    assert rel_fro(got, ref) <= 1e-10
    assert rel_fro(rg.eval().reshape(p.nset, p.ngrid), ref) <= 1e-10
    part = rs.eval(ngrid=64 * 7, nset=2).reshape(p.nset, p.ngrid)
    assert rel_fro(part[:2, :448], ref[:2, :448]) <= 1e-10 and np.all(part[:, 448:] == 0.0) and np.all(part[2] == 0.0)
    old = cfgmm.tune_get(cfgmm.TUNE_RESIDENT_DEPTH)
    try:
        cfgmm.tune_set(cfgmm.TUNE_RESIDENT_DEPTH, 2)
        assert np.array_equal(rs.eval().reshape(p.nset, p.ngrid), got)
    finally:
        cfgmm.tune_set(cfgmm.TUNE_RESIDENT_DEPTH, old)
    want = [rs.eval(), rg.eval()]
    bufs = [cfgmm.DeviceBuffer(8 * p.nset * p.ngrid) for _ in range(2)]
    for b in bufs:
        b.zero()
    s = cfgmm.Stream()
    cfgmm.rho_resident_eval_batched(s, [(rs.h, bufs[0], p.ngrid, p.ngrid, p.nset), (rg.h, bufs[1], p.ngrid, p.ngrid,
                                                                                    p.nset)])
    s.synchronize()
    for b, w in zip(bufs, want):
        assert np.array_equal(b.to_array(np.float64, w.shape), w)
        b.free()
    s.destroy()
    general_bytes = rg.h.bytes
    rs.close()
    rg.close()
    # one asymmetric entry inside the mask: the general form, still right
    lst = np.flatnonzero(p.tab[:, 0])
    bad = p.dms.copy()
    bad[1, lst[2], lst[min(len(lst) - 1, 90)]] += 5.0
    rb = Resident(p, same=True, dms=bad)
    assert rb.h.bytes == general_bytes
    assert rel_fro(rb.eval().reshape(p.nset, p.ngrid), _oracle(p, dms=bad)) <= 1e-10
    rb.close()
    # wide-range D
    pw = W.rho_problem(seed=42, nshell=200, nao=640, ngrid=2048, nset=2, density=0.3, dm_logspan=(-8, 2))
    for ns in (6, 5):
        rw = Resident(pw, ns=ns, same=True)
        assert rel_fro(rw.eval().reshape(pw.nset, pw.ngrid), _oracle(pw)) <= 1e-10
        rw.close()
