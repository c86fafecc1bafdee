"""Autotuner of the residual's implementation options.

Every option of ibis_residual_set_option listed here changes only how the
work is laid out (face-value layout, precomputed offsets, packing, face
kernel and its block size, LDS staging of offsets or of the cell stencil,
prefetching, cells per tile, which is also the block size of the tile
kernels, whether the cell kernel runs the boundary actions of its stencil's
ghosts), never the arithmetic: all choices give the same bits.  The tuner
times a caller-supplied workload under each candidate and keeps the
fastest, by coordinate descent.  A candidate the library refuses for the
mesh at hand (for example a tile whose face list does not fit in LDS) is
skipped.
"""
from __future__ import annotations

from typing import Callable, Dict, Mapping, Tuple

# each option with its candidate values; the first is the starting point
OPTION_SPACE = (("face_state_layout", (2, 1, 0)), ("precomputed_offsets", (1, 0)),
                ("pack_cells", (0, 1)), ("face_kernel", (1, 0, 64, 128, 256)),
                ("stage_offsets", (1, 0)), ("cell_prefetch", (0, 1)), ("stage_cells", (1, 0)),
                ("tile", (256, 128, 512)), ("fold_boundary", (1, 0)))
SHORT = {"face_state_layout": "layout", "precomputed_offsets": "offsets", "pack_cells": "pack",
         "face_kernel": "fk", "stage_offsets": "stage", "cell_prefetch": "pf",
         "stage_cells": "sc", "tile": "tile", "fold_boundary": "fold"}


def variant_key(opt: Mapping[str, int]) -> str:
    return "_".join(f"{SHORT[k]}{opt[k]}" for k, _ in OPTION_SPACE)


def canonical(opt: Mapping[str, int]) -> Dict[str, int]:
    """The same kernels under one name: cell_prefetch applies only to the
    cell kernel that reads its stencil from memory, so with stage_cells 1
    both of its values launch the same code."""
    out = dict(opt)
    if out.get("stage_cells", 0) == 1 and "cell_prefetch" in out:
        out["cell_prefetch"] = 0
    return out


def coordinate_descent(set_option: Callable[[str, int], None], measure: Callable[[], float],
                       fixed: Mapping[str, int] = None,
                       log: Callable[[str, float], None] = None,
                       sweeps: int = 2,
                       margin: float = 0.005) -> Tuple[Dict[str, int], float, Dict[str, float]]:
    """Sweeps over OPTION_SPACE, each trying every value of one option with
    the others at their best so far.  ``measure`` times the workload under
    the options set (lower is better); ``fixed`` pins options.  A candidate
    replaces the best only when it is faster by more than ``margin``
    (relative), so timing noise does not move the choice away from the
    earlier values of an option.  Leaves the best options set and returns
    them, their time and every measured variant (keyed by variant_key)."""
    fixed = dict(fixed or {})
    cur = {k: fixed.get(k, vals[0]) for k, vals in OPTION_SPACE}
    seen: Dict[str, float] = {}

    def score(opt):
        # equivalent options share one measurement, so noise between two
        # runs of the same kernels cannot pick either
        key = variant_key(canonical(opt))
        if key not in seen:
            try:
                for k, v in opt.items():
                    set_option(k, v)
            except Exception:  # refused for this mesh: never chosen
                seen[key] = float("inf")
                return seen[key]
            seen[key] = measure()
            if log:
                log(key, seen[key])
        return seen[key]

    best = score(cur)
    for _ in range(sweeps):
        for k, vals in OPTION_SPACE:
            if k in fixed:
                continue
            for v in vals:
                cand = dict(cur, **{k: v})
                t = score(cand)
                if t < best * (1.0 - margin):
                    best, cur = t, cand
    for k, v in cur.items():
        set_option(k, v)
    return cur, best, seen


def tune_residual(fv, prim, gb, dudt, reps: int = 10, batches: int = 3,
                  fixed: Mapping[str, int] = None, log: Callable[[str, float], None] = None):
    """Tune a FiniteVolume on one state: each candidate runs 2 untimed
    residuals, then ``batches`` batches of ``reps`` timed ones (HIP events
    on the current stream); its time is the fastest batch's µs per
    residual."""
    import torch
    stream = torch.cuda.current_stream()

    def measure():
        for _ in range(2):
            fv.compute_dudt(prim, gb, dudt)
        out = []
        for _ in range(batches):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            for _ in range(reps):
                fv.compute_dudt(prim, gb, dudt)
            b.record(stream)
            torch.cuda.synchronize()
            out.append(a.elapsed_time(b) * 1e3 / reps)
        return min(out)

    return coordinate_descent(fv.set_option, measure, fixed, log)
