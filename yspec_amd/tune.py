"""Autotuner of the batched path.

The batched forward (ysx_forward_events mode 1) has these knobs:

* the launch shape (ysx_set_batch_tuning): one block per work item, or
  persistent blocks that take `chunk` items per atomic queue access with
  `blocks_per_cu` blocks resident per compute unit;
* lanes per pair where the kernels offer a choice (ysx_set_lanes): 8 lanes
  sharing the stage coefficients of one l = 0 radial solve, or one lane per
  solve.

The tuner times the candidates on the events already loaded in the context
(one warm-up forward, then the best of `reps`): the launch shapes first,
then the lanes per radial solve under the fastest shape.  It leaves the
fastest setting applied.  The other kernels' lanes per pair follow from
their lockstep layout (one lane per (l, w) for the minor and toroidal
marches, one per receiver-side row for the b-vector rows), and their LDS
layout (the stage slots and the coefficient chunk of the marches) from the
four blocks per CU that their registers allow.
"""
from __future__ import annotations

CANDIDATES = [(0, 1, 4)] + [(1, c, b) for b in (4, 8, 16) for c in (1, 2, 4)]
RAD_LANES = (8, 1)


def _time(g, reps):
    g.forward_events(1)
    return min(g.forward_events(1) for _ in range(reps))


def tune(g, reps: int = 2, candidates=None, rad_lanes=RAD_LANES, log=None):
    """Time the batched forward of context ``g`` (an api.YSpec with events
    set) for each (persist, chunk, blocks_per_cu), then for each number of
    lanes per radial solve under the fastest shape; apply and return the
    fastest as a dict with every candidate's time in ms."""
    best, table = None, []
    for persist, chunk, bpc in candidates or CANDIDATES:
        g.set_batch_tuning(persist, chunk, bpc)
        ms = _time(g, reps)
        table.append({"persist": persist, "chunk": chunk, "blocks_per_cu": bpc, "ms": ms})
        if log:
            log(f"tune persist={persist} chunk={chunk} blocks_per_cu={bpc}: {ms:.1f} ms")
        if best is None or ms < best["ms"]:
            best = table[-1]
    g.set_batch_tuning(best["persist"], best["chunk"], best["blocks_per_cu"])
    best = dict(best, rad_lanes=0)
    lanes = []
    for n in rad_lanes:
        g.set_lanes(n)
        ms = _time(g, reps)
        lanes.append({"rad_lanes": n, "ms": ms})
        if log:
            log(f"tune rad_lanes={n}: {ms:.1f} ms")
        if ms < best["ms"]:
            best = dict(best, rad_lanes=n, ms=ms)
    g.set_lanes(best["rad_lanes"])
    return {"best": best, "candidates": table, "rad_lanes": lanes}
