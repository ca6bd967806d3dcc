"""The runtime surface of the C ABI on the GPU: device queries, allocations, synchronous and
stream-ordered copies, memset, events, and the library's kernel tracing (hipEvent pairs around each
launch, read back per kernel name)."""
import numpy as np
import pytest

import cfgmm
from cfgmm import workloads as W

pytestmark = pytest.mark.gpu


def test_device_queries_and_memory():
    assert cfgmm.device_count() >= 1
    free, total = cfgmm.mem_get_info()
    assert 0 < free <= total
    rng = np.random.default_rng(0)
    x = rng.standard_normal(100_003)
    b = cfgmm.DeviceBuffer.from_array(x)  # synchronous upload
    assert np.array_equal(b.to_array(np.float64, x.shape), x)
    b.zero()
    assert not np.any(b.to_array(np.float64, x.shape))
    # stream-ordered copies: upload, device-to-device, download, all on one stream
    s = cfgmm.Stream()
    c = cfgmm.DeviceBuffer(x.nbytes)
    b.upload(x, s)
    cfgmm.check(cfgmm.lib().cfgmm_memcpy_dtod_async(c.ptr, b.ptr, x.nbytes, s.handle))
    out = np.empty_like(x)
    c.download(out, s)
    s.synchronize()
    assert np.array_equal(out, x)
    for buf in (b, c):
        buf.free()
    s.destroy()


def test_events_order_and_elapsed():
    """Two events around a DGEMM on one stream: both complete after the stream synchronises, the
    elapsed time is positive, and an event recorded on an idle stream completes."""
    g = W.gemm_problem(seed=3, m=512, n=512, k=512)
    s = cfgmm.Stream()
    d_a, d_b = cfgmm.DeviceBuffer.from_array(g.a), cfgmm.DeviceBuffer.from_array(g.b)
    d_c = cfgmm.DeviceBuffer(8 * g.m * g.ldc)
    d_c.zero()
    e0, e1 = cfgmm.Event(), cfgmm.Event()
    e0.record(s)
    for _ in range(5):
        cfgmm.dgemm_dev(s, d_c, d_a, d_b, g.m, g.n, g.k, g.lda, g.ldb, g.ldc, g.transa, g.transb)
    e1.record(s)
    e1.synchronize()
    assert e0.done() and e1.done()
    assert e0.elapsed_ms(e1) > 0.0
    s.synchronize()
    idle = cfgmm.Stream()
    e2 = cfgmm.Event()
    e2.record(idle)
    idle.synchronize()
    assert e2.done()
    for e in (e0, e1, e2):
        e.destroy()
    for buf in (d_a, d_b, d_c):
        buf.free()
    s.destroy()
    idle.destroy()


def test_kernel_tracing_counts_launches():
    """With tracing on, each launch of a traced kernel adds one count and a positive duration under
    its name; off, nothing is added; a reset clears the table."""
    p = W.rho_problem(seed=5, nshell=60, nao=180, ngrid=640, nset=2, density=0.4)
    s = cfgmm.Stream()
    d_d, d_phi = cfgmm.DeviceBuffer.from_array(p.dms), cfgmm.DeviceBuffer.from_array(p.phi)
    d_tab = cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(p.tab))
    d_r = cfgmm.DeviceBuffer(8 * p.nset * p.ngrid)

    def rho():
        cfgmm.aodmao2rho_dev(s, d_r, d_d, d_phi, d_phi, p.nao, p.ngrid, p.nao, p.ngrid, p.nset, d_tab, p.nblk,
                             p.blksize)

    try:
        cfgmm.trace_reset()
        cfgmm.trace_enable(True)
        for _ in range(3):
            rho()
        s.synchronize()
        ms, n = cfgmm.trace_read("rho_chunk_kernel")
        assert n == 3 and ms > 0.0
        ms_l, n_l = cfgmm.trace_read("build_ao_lists")
        assert n_l == 3 and ms_l > 0.0
        assert cfgmm.trace_read("no_such_kernel") == (0.0, 0)
        cfgmm.trace_enable(False)
        rho()
        s.synchronize()
        assert cfgmm.trace_read("rho_chunk_kernel")[1] == 3
        cfgmm.trace_reset()
        assert cfgmm.trace_read("rho_chunk_kernel") == (0.0, 0)
    finally:
        cfgmm.trace_enable(False)
        cfgmm.trace_reset()
        for buf in (d_d, d_phi, d_tab, d_r):
            buf.free()
        s.destroy()
