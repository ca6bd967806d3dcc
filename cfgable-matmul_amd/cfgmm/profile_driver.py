"""Runs one hot-path operation a few times on resident inputs, for profilers.

    python -m cfgmm.profile_driver <mode> [reps]

Modes:
  rho            masked FP64 rho, nao 2228, ngrid 16384, nset 3, 30 % mask (rho_chunk_kernel)
  rho_split<S>   the same rho in split precision with S int8 slices (rho_ozaki_kernel)
  rho_resident<S> the same rho over resident operands, S int8 slices (rho_resident_kernel; the
                 one-time preparation runs before the repetitions)
  dgemm_large    FP64 DGEMM 3527x9583x6581 NN (dgemm_kernel)
  dgemm_small    FP64 DGEMM 1031x977x1543 NT
  dgemm_small_tuned  the same with the autotuned small plan (the cached choice of class dgemm_small,
                 else 64x64 tiles x 6 K ranges)
  masked_gemm    the reference's masked GEMM test: 2228x16384x2228 NN, sin/cos, 30 % periodic mask,
                 BLKSIZE 48 (masked_gemm_kernel)
  gemm_split<S>  the large GEMM in split precision (ozaki_gemm_kernel)
  vxc            V += phi w phi^T over the same mask (vxc_block_kernel, vxc_reduce_kernel)
  gto            AO tabulation of workloads.gto_workload (700 shells, 1890 AOs, 16384 points)
  copy           a 256 MiB device-to-device copy (byte-counter calibration)
  batched        512 uneven rho problems over 32 generated molecules in one FP64 batched launch
                 (workloads.batched_problems; rho_batched kernels)
  batched_resident<S> the same 512 problems as leading ranges of the 32 molecules' resident
                 handles (S int8 slices; the handles are created before the repetitions) in one
                 cfgmm_rho_resident_eval_batched call
  indep          512 independent rho problems (workloads.independent_problems(n=512, seed=4000)) in
                 one FP64 batched launch
  indep_resident<S> the same problems, each sliced into its own resident handle (before the
                 repetitions), in one cfgmm_rho_resident_eval_batched call

Each pass of a hardware-counter profiler is a separate process running this
module; the inputs are generated, uploaded once and then the operation runs
`reps` times.
"""
from __future__ import annotations

import sys

import numpy as np

import cfgmm
from cfgmm import workloads as W


def main(argv: list[str]) -> int:
    mode = argv[1] if len(argv) > 1 else "rho"
    reps = int(argv[2]) if len(argv) > 2 else 3
    s = cfgmm.Stream()
    if mode == "rho" or mode.startswith("rho_split"):
        prec, ns = (cfgmm.PREC_FP64, 0) if mode == "rho" else (cfgmm.PREC_OZAKI_INT8, int(mode[len("rho_split"):]))
        p = W.rho_problem(seed=0)
        d_r = cfgmm.DeviceBuffer(8 * p.nset * p.ngrid)
        d_r.zero()
        d_d = cfgmm.DeviceBuffer.from_array(p.dms)
        d_phi = cfgmm.DeviceBuffer.from_array(p.phi)
        d_tab = cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(p.tab))
        for _ in range(reps):
            cfgmm.aodmao2rho_dev(s, d_r, d_d, d_phi, d_phi, p.nao, p.ngrid, p.nao, p.ngrid, p.nset, d_tab, p.nblk,
                                 p.blksize, prec, ns)
    elif mode.startswith("rho_resident"):
        p = W.rho_problem(seed=0)
        d_r = cfgmm.DeviceBuffer(8 * p.nset * p.ngrid)
        d_r.zero()
        d_d = cfgmm.DeviceBuffer.from_array(p.dms)
        d_phi = cfgmm.DeviceBuffer.from_array(p.phi)
        d_tab = cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(p.tab))
        h = cfgmm.RhoResident(s, d_d, d_phi, d_phi, p.nao, p.ngrid, p.nao, p.ngrid, p.nset, d_tab, p.nblk,
                              p.blksize, cfgmm.PREC_OZAKI_INT8, int(mode[len("rho_resident"):]))
        for _ in range(reps):
            h.eval(s, d_r, p.ngrid)
    elif mode in ("dgemm_large", "dgemm_small", "dgemm_small_tuned") or mode.startswith("gemm_split"):
        g = W.dgemm_small_nt() if mode.startswith("dgemm_small") else W.dgemm_large_nn()
        if mode == "dgemm_small_tuned":
            from cfgmm import autotune
            if not autotune.apply("dgemm_small"):  # no cached choice: the usual one
                cfgmm.tune_set(cfgmm.TUNE_DGEMM_SMALL_TILE, 64)
                cfgmm.tune_set(cfgmm.TUNE_DGEMM_SMALL_SPLITS, 6)
        prec, ns = (cfgmm.PREC_FP64, 0) if mode.startswith("dgemm") else (cfgmm.PREC_OZAKI_INT8,
                                                                           int(mode[len("gemm_split"):]))
        d_a = cfgmm.DeviceBuffer.from_array(g.a)
        d_b = cfgmm.DeviceBuffer.from_array(g.b)
        d_c = cfgmm.DeviceBuffer(8 * g.m * g.ldc)
        d_c.zero()
        for _ in range(reps):
            cfgmm.dgemm_dev(s, d_c, d_a, d_b, g.m, g.n, g.k, g.lda, g.ldb, g.ldc, g.transa, g.transb, prec, ns)
    elif mode == "masked_gemm":
        m = k = 2228
        n, bs = 16384, 48
        nblk = -(-n // bs)
        a, b = W.sin_cos_operands(m * k, k * n)
        tab = W.periodic_mask(k, nblk, W.MASK_30)
        W.zero_under_mask(b, tab, bs, n)
        d_a, d_b = cfgmm.DeviceBuffer.from_array(a), cfgmm.DeviceBuffer.from_array(b)
        d_t = cfgmm.DeviceBuffer.from_array(tab)
        d_c = cfgmm.DeviceBuffer(8 * m * n)
        d_c.zero()
        for _ in range(reps):
            cfgmm.dgemm_non0tab_dev(s, d_c, d_a, d_b, m, n, k, k, n, n, False, False, d_t, nblk, bs)
    elif mode == "indep" or mode.startswith("indep_resident"):
        probs = W.independent_problems(n=512, seed=4000)
        dev = [(cfgmm.DeviceBuffer.from_array(q.dms), cfgmm.DeviceBuffer.from_array(q.phi),
                cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(q.tab))) for q in probs]
        offs = np.concatenate([[0], np.cumsum([q.nset * q.ngrid for q in probs])])
        out = cfgmm.DeviceBuffer(8 * int(offs[-1]))
        out.zero()
        if mode == "indep":
            arr = (cfgmm.RhoProblem * len(probs))(*[
                cfgmm.RhoProblem(out.ptr + 8 * int(offs[i]), d.ptr, phi.ptr, phi.ptr, q.nao, q.ngrid, q.nao, q.ngrid,
                                 q.nset, tab.ptr, q.nblk, q.blksize) for i, (q, (d, phi, tab)) in enumerate(zip(probs, dev))])
            for _ in range(reps):
                cfgmm.aodmao2rho_batched_dev(s, arr, cfgmm.PREC_FP64, 0)
        else:
            nslices = int(mode[len("indep_resident"):])
            handles = [cfgmm.RhoResident(s, d, phi, phi, q.nao, q.ngrid, q.nao, q.ngrid, q.nset, tab, q.nblk,
                                         q.blksize, cfgmm.PREC_OZAKI_INT8, nslices) for q, (d, phi, tab) in zip(probs, dev)]
            evals = [(h, out.ptr + 8 * int(offs[i]), q.ngrid, q.ngrid, q.nset)
                     for i, (h, q) in enumerate(zip(handles, probs))]
            for _ in range(reps):
                cfgmm.rho_resident_eval_batched(s, evals)
    elif mode == "vxc":
        p = W.rho_problem(seed=0)
        w = np.random.default_rng(6).uniform(0.1, 2.0, p.ngrid)
        d_v = cfgmm.DeviceBuffer(8 * p.nao * p.nao)
        d_v.zero()
        d_phi = cfgmm.DeviceBuffer.from_array(p.phi)
        d_w = cfgmm.DeviceBuffer.from_array(w)
        d_tab = cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(p.tab))
        for _ in range(reps):
            cfgmm.aogridao2mat_dev(s, d_v, p.nao, d_phi, d_phi, d_w, p.ngrid, p.nao, p.ngrid, d_tab, p.nblk, p.blksize)
    elif mode == "gto":
        gw = W.gto_workload()
        shells = cfgmm.gto_shells(gw["centers"], gw["ls"], gw["nprim"], gw["poff"], gw["aoff"])
        d_phi = cfgmm.DeviceBuffer(8 * gw["nao"] * gw["ngrid"])
        d_tab = cfgmm.DeviceBuffer(gw["nao"] * gw["nblk"])
        d_xyz = cfgmm.DeviceBuffer.from_array(gw["coords"])
        for _ in range(reps):
            cfgmm.eval_gto_dev(s, d_phi, gw["ngrid"], d_tab, gw["nblk"], shells, gw["exps"], gw["coefs"], d_xyz,
                               gw["ngrid"], gw["nao"], gw["blksize"], gw["cutoff"])
    elif mode == "batched" or mode.startswith("batched_resident"):
        pool = W.molecule_pool(n=32, seed=100)
        probs = W.batched_problems(pool, n=512, seed=7)
        dev = [(cfgmm.DeviceBuffer.from_array(m.prob.dms), cfgmm.DeviceBuffer.from_array(m.prob.phi),
                cfgmm.DeviceBuffer.from_array(np.ascontiguousarray(m.prob.tab))) for m in pool]
        # rho rows strided by the molecule's grid, as its phi rows are
        sizes = [(q.nset - 1) * pool[q.mol].prob.ngrid + q.ngrid for q in probs]
        offs = np.concatenate([[0], np.cumsum(sizes)])
        out = cfgmm.DeviceBuffer(8 * int(offs[-1]))
        out.zero()
        if mode == "batched":
            descs = []
            for i, q in enumerate(probs):
                pm = pool[q.mol].prob
                d, phi, tab = dev[q.mol]
                descs.append(cfgmm.RhoProblem(out.ptr + 8 * int(offs[i]), d.ptr, phi.ptr, phi.ptr, pm.nao, pm.ngrid,
                                              pm.nao, q.ngrid, q.nset, tab.ptr, pm.nblk, pm.blksize))
            arr = (cfgmm.RhoProblem * len(descs))(*descs)
            for _ in range(reps):
                cfgmm.aodmao2rho_batched_dev(s, arr, cfgmm.PREC_FP64, 0)
        else:
            nslices = int(mode[len("batched_resident"):])
            handles = []
            for m, (d, phi, tab) in zip(pool, dev):
                pm = m.prob
                handles.append(cfgmm.RhoResident(s, d, phi, phi, pm.nao, pm.ngrid, pm.nao, pm.ngrid, pm.nset, tab,
                                                 pm.nblk, pm.blksize, cfgmm.PREC_OZAKI_INT8, nslices))
            evals = [(handles[q.mol], out.ptr + 8 * int(offs[i]), pool[q.mol].prob.ngrid, q.ngrid, q.nset)
                     for i, q in enumerate(probs)]
            for _ in range(reps):
                cfgmm.rho_resident_eval_batched(s, evals)
    elif mode == "copy":
        nbytes = 256 << 20
        a = cfgmm.DeviceBuffer(nbytes)
        b = cfgmm.DeviceBuffer(nbytes)
        a.zero()
        for _ in range(reps):
            cfgmm.check(cfgmm.lib().cfgmm_memcpy_dtod_async(b.ptr, a.ptr, nbytes, s.handle))
    else:
        print(f"unknown mode {mode!r}", file=sys.stderr)
        return 2
    s.synchronize()
    print(f"profile_driver: {mode} x{reps} done", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
