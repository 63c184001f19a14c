"""Autotuner of the batched form's launch configuration.

Searched per shape class, on the live batch, timing each candidate with HIP
events on the launch stream (median of ``reps`` after ``warmup`` launches):

* GEMM tile bm x bn in {64, 128}^2, 128 x 256 or 256 x 128 (8 waves) per (Nr, Nt): timed on
  the G2 pass (w1 -> w0,
  no phase), which reads and writes scratch buffers only;
* nonlinear threads per row, Nt/8 or Nt/4, per (Nt, STE): timed on the
  nonlinear pass, which works in place on the scratch buffer w0;
* grid of each pass kind over the whole batch: one block per work item or a
  persistent grid of k x 256 blocks;
* launch mode per pass: those grouped launches, or one launch over the merged work
  list of every instance (fixed shapes), whichever takes less time.

Only the scratch buffers and the scale words are touched; the spectral state is
not. Before timing, w1 and w0 are loaded with copies of the state (so the GEMM
and nonlinear kernels see representative data), and the state's scale word is
recomputed at the end. Results are a plain dict that can be saved as JSON and
applied again with ``apply``.
"""
from __future__ import annotations

import ctypes
import json
import statistics

import torch

from . import _lib
from ._lib import call, stream_ptr

TILES = [(256, 128), (128, 256), (128, 128), (128, 64), (64, 128), (64, 64)]
GRIDS = [0, 256, 512, 1024, 2048]


def _time(fn, warmup, reps):
    st = torch.cuda.current_stream()
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(st)
        fn()
        b.record(st)
        b.synchronize()
        ts.append(a.elapsed_time(b))
    return statistics.median(ts)


def _sub_plan(batch, idx):
    descs = (_lib.BatchInstance * len(idx))()
    for k, i in enumerate(idx):
        descs[k] = batch._descs[i]
    plan = ctypes.c_void_p()
    call("fp_batch_plan_create", descs, len(idx), ctypes.byref(plan))
    return plan


def _prepare(batch):
    for m in batch.members:
        mx = m._mx
        mx.w1.copy_(mx.spec)
        mx.w0.copy_(mx.spec)
        mx.slots[1].copy_(mx.slots[0])


def tune(batch, warmup=2, reps=5, grids=True, verbose=False):
    """Pick tiles / threads / grids for `batch` (a BatchedIntegrator) and apply them."""
    lib = _lib.load()
    _prepare(batch)
    ms = batch.members
    res = {"gemm": {}, "nl": {}, "grid": {}, "timings_ms": {}}
    st = stream_ptr()

    shapes = {}
    for i, m in enumerate(ms):
        shapes.setdefault((m.Nr, m.Nt), []).append(i)
    for (Nr, Nt), idx in sorted(shapes.items()):
        plan = _sub_plan(batch, idx)
        try:
            best = None
            for bm, bn in TILES:
                if Nr % bm or (2 * Nt) % bn:
                    continue
                for k in range(len(idx)):
                    call("fp_batch_set_config", plan, k, bm, bn, 0)
                t = _time(lambda: call("fp_batch_pass", plan, 2, st), warmup, reps)
                res["timings_ms"][f"gemm {Nr}x{Nt} {bm}x{bn}"] = t
                if best is None or t < best[0]:
                    best = (t, bm, bn)
            res["gemm"][f"{Nr}x{Nt}"] = [best[1], best[2]]
        finally:
            lib.fp_batch_plan_destroy(plan)

    nl_classes = {}
    for i, m in enumerate(ms):
        nl_classes.setdefault((m.Nt, int(m.cfg.enable_ste)), []).append(i)
    for (Nt, ste), idx in sorted(nl_classes.items()):
        plan = _sub_plan(batch, idx)
        try:
            best = None
            for tpb in ((Nt // 16,) if Nt == 4096 else ()) + (Nt // 8, Nt // 4):
                for k in range(len(idx)):
                    call("fp_batch_set_config", plan, k, 0, 0, tpb)
                t = _time(lambda: call("fp_batch_pass", plan, 3, st), warmup, reps)
                res["timings_ms"][f"nl {Nt} ste{ste} tpb{tpb}"] = t
                if best is None or t < best[0]:
                    best = (t, tpb)
            res["nl"][f"{Nt}:{ste}"] = best[1]
        finally:
            lib.fp_batch_plan_destroy(plan)

    apply(batch, res)
    if grids:
        for kind, pas in (("plasma", 0), ("gemm", 2), ("nl", 3)):
            best = None
            for g in GRIDS:
                batch.set_grid(kind, g)
                t = _time(lambda: call("fp_batch_pass", batch._plan, pas, st), warmup, reps)
                res["timings_ms"][f"grid {kind} {g}"] = t
                if best is None or t < best[0] * 0.98:    # prefer the simpler grid on ties
                    best = (t, g)
            res["grid"][kind] = best[1]
            batch.set_grid(kind, best[1])

    # Launch mode per pass: the tuned grouped launches, or one launch over the merged work
    # list of every instance (fixed shapes; needs Nr % 256 == 0). Timed are the passes that
    # write only scratch buffers, scale words and the carrier densities: the scan, G1, G2, the
    # nonlinear step and G3; G4 (which writes the state) follows G2, the other product without
    # a phase. One launch wins a tie within 1 %.
    res["merged_passes"] = 0
    if all(m.Nr % 256 == 0 for m in ms):
        timed = (0, 1, 2, 3, 4)
        t_g = {p: _time(lambda: call("fp_batch_pass", batch._plan, p, st), warmup, reps)
               for p in timed}
        batch.set_merged_passes(0x3F)
        t_m = {p: _time(lambda: call("fp_batch_pass", batch._plan, p, st), warmup, reps)
               for p in timed}
        mask = 0
        for p in timed:
            res["timings_ms"][f"pass {p} grouped"] = t_g[p]
            res["timings_ms"][f"pass {p} merged"] = t_m[p]
            if t_m[p] <= 1.01 * t_g[p]:
                mask |= 1 << p
        if mask & (1 << 2):
            mask |= 1 << 5
        res["merged_passes"] = mask
        batch.set_merged_passes(mask)
    batch.refresh_scales()
    torch.cuda.synchronize()
    batch.config = res
    if verbose:
        print(json.dumps({k: v for k, v in res.items() if k != "timings_ms"}), flush=True)
    return res


def apply(batch, res):
    """Apply a tuning result (as returned by ``tune`` or loaded from JSON)."""
    for i, m in enumerate(batch.members):
        bm, bn = res["gemm"].get(f"{m.Nr}x{m.Nt}", (0, 0))
        tpb = res["nl"].get(f"{m.Nt}:{int(m.cfg.enable_ste)}", 0)
        batch.set_config(i, bm, bn, tpb)
    for kind, g in res.get("grid", {}).items():
        batch.set_grid(kind, g)
    batch.set_merged_passes(int(res.get("merged_passes", 0)))
    batch.config = res
