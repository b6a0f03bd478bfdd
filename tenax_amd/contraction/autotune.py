"""Tiling autotuner for the grouped block GEMMs.

The reference has no counterpart (it dispatches one einsum per block pair,
contractor.py:444-476, dmrg.py:1097-1109).  A grouped-GEMM launch can run
with several tilings (tnx_plan_set_tiling: tile shape, split K, XCD-aware
task order); which one is fastest depends on the block structure.  For a plan
with enough work, each candidate is timed on the plan's own buffers with HIP
events and the fastest is kept; the choice is cached per (plan kind, stage,
work bucket), so structurally similar plans -- every bond of a DMRG sweep at
a given chi -- reuse it without timing again.

Every tiling is deterministic, but different tilings round differently (in
the last bits); ``ENABLED = False`` (or TENAX_AMD_AUTOTUNE=0) keeps the
default 64x64 whole-K tiling everywhere for bitwise-reproducible runs.
"""

from __future__ import annotations

import math
import os
import threading

import torch

from tenax_amd import _lib

# (shape, kchunk, xcd): shape 0 = 64x64, 1 = 128x64, 2 = 64x128; kchunk -1 = split K automatically,
# -2 = the balanced two-wave split (whole tiles and heads of the long ones at
# once, short tail pieces after).  Near-ties go to the earliest candidate: the
# XCD-ordered whole-K tiling first (it moves ~30% fewer bytes from HBM than
# the plain order at equal kernel time, and needs no split-K reduction
# launch), the default tiling second, then the balanced split (fewer
# partial tiles than the plain split).
CANDIDATES = [(0, 0, 1), (0, 0, 0), (0, -2, 1), (0, -1, 1), (0, -1, 0), (1, 0, 1), (2, 0, 1), (1, -1, 1),
              (2, -1, 1)]
MIN_FLOPS = 2e9  # below this a launch is too short for the tiling to matter
REPS = 2
TIE = 0.05  # switch away from an earlier candidate only for a clear gain
ENABLED = os.environ.get("TENAX_AMD_AUTOTUNE", "1") != "0"

_best: dict = {}
_lock = threading.Lock()
STATS = {"timed": 0, "reused": 0}


def _bucket(flops: float) -> int:
    return int(round(2.0 * math.log2(max(flops, 1.0))))


def _time(run, reps: int) -> float:
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    run()
    s.record()
    for _ in range(reps):
        run()
    e.record()
    e.synchronize()
    return s.elapsed_time(e) / reps


def cached(stage: int, flops: float, kind: str):
    """The tiling already chosen for this (kind, stage, work bucket), or None."""
    with _lock:
        return _best.get((kind, stage, _bucket(flops)))


def tune(plan, stage: int, run, flops: float, kind: str):
    """Choose the tiling of one GEMM stage of ``plan``; ``run()`` launches that
    stage alone on the current stream (its inputs already in place)."""
    if not ENABLED or flops < MIN_FLOPS:
        return None
    sig = (kind, stage, _bucket(flops))
    with _lock:
        cfg = _best.get(sig)
    if cfg is None:
        timed = []
        for c in CANDIDATES:
            try:
                plan.set_tiling(stage, *c)
            except _lib.TnxError as e:
                # a split whose partial tiles would exceed the workspace cap
                # is not a candidate for this plan
                if e.code != _lib.TNX_ERR_CAPACITY:
                    raise
                continue
            timed.append((_time(run, REPS), c))
        # near-ties go to the earliest candidate (fixed order), so the choice
        # does not flip between runs on timing noise
        best = min(t for t, _ in timed)
        cfg = next(c for t, c in timed if t <= (1.0 + TIE) * best)
        with _lock:
            _best[sig] = cfg
        STATS["timed"] += 1
    else:
        STATS["reused"] += 1
    plan.set_tiling(stage, *cfg)
    return cfg


def clear():
    with _lock:
        _best.clear()
    STATS.update(timed=0, reused=0)
