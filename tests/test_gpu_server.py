"""The serving loop (cfgmm_server_*): every result is the same alone and in the mix.

Each request of a Poisson mix (rho over resident molecules, DGEMMs with
device-generated operands) is first served alone, then all are replayed with
overlapping arrivals; the device checksums (a hash of the result bits) must be
identical.  Sampled results are checked against the CPU oracle (rho: FP64
1e-12, 6-slice split precision 1e-10, relative Frobenius) and numpy (GEMM,
1e-12), and the host-side hash of each returned result must equal its device
checksum.  A tight HBM budget forces evictions without changing any result."""
import ctypes

import numpy as np
import pytest

import cfgmm
import oracle
from cfgmm import workloads as W
from helpers import rel_fro

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pool():
    return [W.Molecule(0, W.rho_problem(seed=21, nshell=110, nao=330, ngrid=64 * 40, nset=3, density=0.4)),
            W.Molecule(1, W.rho_problem(seed=22, nshell=170, nao=520, ngrid=64 * 24, nset=4, density=0.25)),
            W.Molecule(2, W.rho_problem(seed=23, nshell=70, nao=210, ngrid=64 * 60, nset=2, density=0.55))]


def make_requests(pool, n, rate, seed):
    return W.serving_requests(pool, n, rate, seed=seed, gemm_range=(48, 400), min_blocks=1)


def to_ctypes(reqs, outs=None):
    out = []
    for i, r in enumerate(reqs):
        o = None if outs is None or outs[i] is None else outs[i].ctypes.data
        out.append(cfgmm.Request(r.kind, r.mol, r.ngrid, r.nset, r.m, r.n, r.k, r.seed, r.arrival_s, o))
    return out


def make_server(pool, precision, budget=64 << 30, slots=4, tile_cap=0):
    sv = cfgmm.Server(budget, slots=slots, rho_precision=precision, rho_nslices=6, max_batch=8,
                      split_tile_cap_bytes=tile_cap)
    for i, m in enumerate(pool):
        p = m.prob
        sv.add_molecule(i, p.dms, p.phi, np.ascontiguousarray(p.tab), p.blksize)
    return sv


def solo_checksums(sv, reqs):
    sums = []
    for r in to_ctypes(reqs):
        r.arrival_s = 0.0
        res = sv.run([r])
        assert res[0].status == 0
        sums.append(res[0].checksum)
    return sums


def rho_oracle(p, ngrid, nset):
    """Compact nset x ngrid result; the oracle's rho rows are strided by the molecule's grid, like phi."""
    out = np.zeros((nset - 1) * p.ngrid + ngrid)
    assert oracle.aodmao2rho(out, p.dms.ravel(), p.phi.ravel(), p.phi.ravel(), p.nao, p.ngrid, p.nao, ngrid, nset,
                             np.ascontiguousarray(p.tab), p.nblk, p.blksize) == 0
    return np.stack([out[s * p.ngrid: s * p.ngrid + ngrid] for s in range(nset)]).ravel()


@pytest.mark.parametrize("precision,tol", [(cfgmm.PREC_FP64, 1e-12), (cfgmm.PREC_OZAKI_INT8, 1e-10)])
def test_server_alone_equals_mix(pool, precision, tol):
    reqs = make_requests(pool, 48, rate=4000.0, seed=5)
    sv = make_server(pool, precision)
    sv.warm([0, 1, 2])
    alone = solo_checksums(sv, reqs)
    outs = [np.zeros(r.nset * r.ngrid if r.kind == 0 else r.m * r.n) if i % 7 == 0 else None
            for i, r in enumerate(reqs)]
    res = sv.run(to_ctypes(reqs, outs))
    stats = sv.stats()
    sv.destroy()
    mixed = [res[i].checksum for i in range(len(reqs))]
    assert all(res[i].status == 0 for i in range(len(reqs)))
    assert mixed == alone
    assert max(res[i].batch_size for i in range(len(reqs))) >= 1
    lat = np.array([res[i].done_s - reqs[i].arrival_s for i in range(len(reqs))])
    assert np.all(lat > 0)
    assert stats["rho_requests"] + stats["gemm_requests"] == 2 * len(reqs)
    for i, o in enumerate(outs):
        if o is None:
            continue
        r = reqs[i]
        assert W.result_checksum(o) == res[i].checksum
        if r.kind == 0:
            ref = rho_oracle(pool[r.mol].prob, r.ngrid, r.nset)
            assert rel_fro(o, ref) <= tol
        else:
            a, b = W.serving_gemm_operands(r.m, r.n, r.k, r.seed)
            assert rel_fro(o.reshape(r.m, r.n), a @ b) <= 1e-12


def test_server_tight_budget_evicts(pool):
    """A budget that holds about one molecule: requests still complete, with evictions, same results."""
    reqs = [r for r in make_requests(pool, 40, rate=3000.0, seed=9) if r.kind == 0]
    big = max(m.prob.phi.nbytes + m.prob.dms.nbytes for m in pool)
    sv = make_server(pool, cfgmm.PREC_FP64, budget=int(1.6 * big) + (4 << 20), slots=2)
    alone = solo_checksums(sv, reqs)
    res = sv.run(to_ctypes(reqs))
    stats = sv.stats()
    sv.destroy()
    assert [res[i].checksum for i in range(len(reqs))] == alone
    assert stats["evictions"] > 0 and stats["misses"] > 0


def test_server_rejects_bad_requests(pool):
    sv = make_server(pool, cfgmm.PREC_FP64)
    bad = cfgmm.Request(0, 7, 64, 1, 0, 0, 0, 0, 0.0, None)  # unknown molecule
    with pytest.raises(cfgmm.CfgmmError) as e:
        sv.run([bad])
    assert e.value.status == cfgmm.EINVAL
    bad = cfgmm.Request(0, 0, 100, 1, 0, 0, 0, 0, 0.0, None)  # not a multiple of blksize
    with pytest.raises(cfgmm.CfgmmError) as e:
        sv.run([bad])
    assert e.value.status == cfgmm.EINVAL_DIMS
    sv.destroy()


def test_server_failed_upload_leaves_molecule_clean():
    """A split-precision server whose tile cap admits molecule 0's tiles but whose budget holds only
    its D, phi and mask:
    the request on it fails with EOOM after the tiles are built, the molecule is freed and left
    non-resident with nothing charged, and the same server then serves the small molecule 1
    with the result it gives under a large budget, and fails molecule 0 again the same way."""
    pool = [W.Molecule(0, W.rho_problem(seed=31, nshell=100, nao=300, ngrid=64 * 48, nset=2, density=0.6)),
            W.Molecule(1, W.rho_problem(seed=32, nshell=20, nao=60, ngrid=64 * 4, nset=2, density=0.5))]

    def held_after_warm(prec, mid):
        sv = make_server(pool, prec)
        sv.warm([mid])
        held = sv.stats()["held_bytes"]
        sv.destroy()
        return held
    base0 = held_after_warm(cfgmm.PREC_FP64, 0)
    tiles0 = held_after_warm(cfgmm.PREC_OZAKI_INT8, 0) - base0
    full1 = held_after_warm(cfgmm.PREC_OZAKI_INT8, 1)
    budget = base0 + tiles0 // 2
    assert tiles0 > (8 << 20) and full1 + (4 << 20) < budget  # the sizes this test relies on
    r0 = cfgmm.Request(0, 0, pool[0].prob.ngrid, 2, 0, 0, 0, 0, 0.0, None)
    r1 = cfgmm.Request(0, 1, pool[1].prob.ngrid, 2, 0, 0, 0, 0, 0.0, None)
    ref = make_server(pool, cfgmm.PREC_OZAKI_INT8)
    want1 = ref.run([r1])[0].checksum
    ref.destroy()
    sv = make_server(pool, cfgmm.PREC_OZAKI_INT8, budget=budget, tile_cap=1 << 40)  # tiles built, then too big
    for _ in range(2):
        with pytest.raises(cfgmm.CfgmmError) as e:
            sv.run([r0])
        assert e.value.status == cfgmm.EOOM
        assert sv.stats()["held_bytes"] == 0
        res = sv.run([r1])
        assert res[0].status == 0 and res[0].checksum == want1
    sv.destroy()


def test_server_tile_cap_serves_large_molecules_in_fp64():
    """Split precision with a tile cap between the two molecules' tile sizes: requests on the large
    molecule run through the FP64 kernels (the same checksums as an FP64 server), those on the
    small one over resident tiles (the same checksums as a split server without a cap)."""
    pool = [W.Molecule(0, W.rho_problem(seed=41, nshell=100, nao=300, ngrid=64 * 8, nset=3, density=0.3)),
            W.Molecule(1, W.rho_problem(seed=42, nshell=200, nao=600, ngrid=64 * 40, nset=3, density=0.5))]
    reqs = [r for r in make_requests(pool, 40, rate=3000.0, seed=13) if r.kind == 0]
    held = []
    for i in range(len(pool)):
        sv = make_server(pool, cfgmm.PREC_OZAKI_INT8)
        base = make_server(pool, cfgmm.PREC_FP64)
        sv.warm([i])
        base.warm([i])
        held.append(sv.stats()["held_bytes"] - base.stats()["held_bytes"])
        sv.destroy()
        base.destroy()
    cap = 2 * held[0]  # between the two molecules' tile sizes (the server checks a bound of the tiles)
    assert cap < held[1], held  # the sizes this test relies on
    want = {}
    for prec in (cfgmm.PREC_FP64, cfgmm.PREC_OZAKI_INT8):
        sv = make_server(pool, prec)
        want[prec] = [sv.run([r])[0].checksum for r in to_ctypes(reqs)]
        sv.destroy()
    sv = make_server(pool, cfgmm.PREC_OZAKI_INT8, tile_cap=cap)
    res = sv.run(to_ctypes(reqs))
    stats = sv.stats()
    sv.destroy()
    assert all(res[i].status == 0 for i in range(len(reqs)))
    for i, r in enumerate(reqs):
        assert res[i].checksum == want[cfgmm.PREC_FP64 if r.mol == 1 else cfgmm.PREC_OZAKI_INT8][i], (i, r.mol)
    assert stats["rho_fp64_requests"] == sum(r.mol == 1 for r in reqs) > 0
    assert 0 < sum(r.mol == 0 for r in reqs)
