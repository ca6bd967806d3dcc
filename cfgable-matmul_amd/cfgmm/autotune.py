"""Autotuner for the library's launch tunables (cfgmm_tune_set / cfgmm_tune_get).

It times each candidate setting on the caller's own resident workload with
hipEvents on the launch stream, keeps the fastest, and stores the choice in a
JSON cache keyed by device architecture and workload class
(``tuning/<arch>.json`` next to this package unless a path is given).

Tunables:
  batched rho  -- persistent workgroups per CU (1..4) and density matrices per
                  task (1, 2, 4);
  DGEMM        -- tile rows per raster group (1, 2, 4, 8, 16); for shapes with
                  fewer 128x128 tiles than workgroup slots, the built-in plan
                  (the cost model's pick, or the in-workgroup K split up to
                  k = 1024), the in-workgroup K split at any k, or the tile
                  edge (64, 128) and the number of K ranges per tile (1..12);
  FP64 rho     -- density matrices per workgroup (1..4).
"""
from __future__ import annotations

import itertools
import json
import os

import cfgmm

TUNING_DIR = os.path.join(cfgmm.PKG_ROOT, "tuning")
ARCH = "gfx950"


def _time(fn, stream, reps=3, warmup=1) -> float:
    for _ in range(warmup):
        fn()
    stream.synchronize()
    e0, e1 = cfgmm.Event(), cfgmm.Event()
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / reps
    e0.destroy()
    e1.destroy()
    return ms


def tune_batched_rho(run_batch, stream, wgs=(1, 2, 3, 4), sets=(1, 2, 4), reps=3):
    """run_batch(): enqueues one batched rho call on `stream`.  Returns (best settings, all timings)."""
    old = {p: cfgmm.tune_get(p) for p in (cfgmm.TUNE_BATCH_WGS_PER_CU, cfgmm.TUNE_BATCH_MAX_SETS)}
    results = []
    try:
        for w, s in itertools.product(wgs, sets):
            cfgmm.tune_set(cfgmm.TUNE_BATCH_WGS_PER_CU, w)
            cfgmm.tune_set(cfgmm.TUNE_BATCH_MAX_SETS, s)
            results.append(({"wgs_per_cu": w, "max_sets": s}, _time(run_batch, stream, reps)))
    finally:
        for p, v in old.items():
            cfgmm.tune_set(p, v)
    best = min(results, key=lambda r: r[1])
    return best[0], results


def tune_dgemm_group(run_gemm, stream, groups=(1, 2, 4, 8, 16), reps=3):
    old = cfgmm.tune_get(cfgmm.TUNE_DGEMM_GROUP_M)
    results = []
    try:
        for g in groups:
            cfgmm.tune_set(cfgmm.TUNE_DGEMM_GROUP_M, g)
            results.append(({"group_m": g}, _time(run_gemm, stream, reps)))
    finally:
        cfgmm.tune_set(cfgmm.TUNE_DGEMM_GROUP_M, old)
    best = min(results, key=lambda r: r[1])
    return best[0], results


def tune_dgemm_small(run_gemm, stream, tiles=(64, 128), splits=(1, 2, 3, 4, 5, 6, 7, 8, 10, 12), reps=20):
    """Plan of a DGEMM too small to fill the chip with 128x128 tiles: the in-workgroup K split (where
    it applies), or the tile edge and K ranges of the split over workgroups."""
    params = (cfgmm.TUNE_DGEMM_KGROUPS, cfgmm.TUNE_DGEMM_SMALL_TILE, cfgmm.TUNE_DGEMM_SMALL_SPLITS)
    old = {p: cfgmm.tune_get(p) for p in params}
    cands = [{"kgroups": 0, "small_tile": 0, "small_splits": 0}, {"kgroups": 2, "small_tile": 0, "small_splits": 0}] + \
        [{"kgroups": 1, "small_tile": t, "small_splits": sp} for t, sp in itertools.product(tiles, splits)]
    results = []
    try:
        for cand in cands:
            for p, key in zip(params, ("kgroups", "small_tile", "small_splits")):
                cfgmm.tune_set(p, cand[key])
            results.append((cand, _time(run_gemm, stream, reps, warmup=3)))
    finally:
        for p, v in old.items():
            cfgmm.tune_set(p, v)
    best = min(results, key=lambda r: r[1])
    return best[0], results


def tune_rho_sets(run_rho, stream, sets=(1, 2, 3, 4), reps=10):
    old = cfgmm.tune_get(cfgmm.TUNE_RHO_MAX_SETS)
    results = []
    try:
        for v in sets:
            cfgmm.tune_set(cfgmm.TUNE_RHO_MAX_SETS, v)
            results.append(({"rho_max_sets": v}, _time(run_rho, stream, reps)))
    finally:
        cfgmm.tune_set(cfgmm.TUNE_RHO_MAX_SETS, old)
    best = min(results, key=lambda r: r[1])
    return best[0], results


def cache_path(path=None) -> str:
    return path or os.path.join(TUNING_DIR, f"{ARCH}.json")


def save(entries: dict, path=None) -> str:
    """Merge {workload class: settings} into the cache file."""
    path = cache_path(path)
    data = load_all(path)
    data.update(entries)
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(data, f, indent=2, sort_keys=True)
    return path


def load_all(path=None) -> dict:
    path = cache_path(path)
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def apply(workload_class: str, path=None) -> dict:
    """Apply the cached settings of a workload class (if any) to the library; returns them."""
    cfg = load_all(path).get(workload_class, {})
    if "wgs_per_cu" in cfg:
        cfgmm.tune_set(cfgmm.TUNE_BATCH_WGS_PER_CU, int(cfg["wgs_per_cu"]))
    if "max_sets" in cfg:
        cfgmm.tune_set(cfgmm.TUNE_BATCH_MAX_SETS, int(cfg["max_sets"]))
    if "group_m" in cfg:
        cfgmm.tune_set(cfgmm.TUNE_DGEMM_GROUP_M, int(cfg["group_m"]))
    if "small_tile" in cfg:
        cfgmm.tune_set(cfgmm.TUNE_DGEMM_SMALL_TILE, int(cfg["small_tile"]))
    if "small_splits" in cfg:
        cfgmm.tune_set(cfgmm.TUNE_DGEMM_SMALL_SPLITS, int(cfg["small_splits"]))
    if "kgroups" in cfg:
        cfgmm.tune_set(cfgmm.TUNE_DGEMM_KGROUPS, int(cfg["kgroups"]))
    if "rho_max_sets" in cfg:
        cfgmm.tune_set(cfgmm.TUNE_RHO_MAX_SETS, int(cfg["rho_max_sets"]))
    return cfg
