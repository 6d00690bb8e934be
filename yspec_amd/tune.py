"""Autotuner of the batched path.

The batched forward (ysx_forward_events mode 1) has three knobs
(ysx_set_batch_tuning): one block per work item, or persistent blocks that
take `chunk` items per atomic queue access with `blocks_per_cu` blocks
resident per compute unit.  The tuner times each candidate on the events
already loaded in the context (one warm-up forward, then the best of
`reps`) and leaves the fastest one set.  Lanes per pair are fixed by the
kernels (one lane per (l, w) for the minor and toroidal marches, four --
one per source column -- for the b-vector); the LDS layout per block
follows from that.
"""
from __future__ import annotations

CANDIDATES = [(0, 1, 4)] + [(1, c, b) for b in (4, 8, 16) for c in (1, 2, 4)]


def tune(g, reps: int = 2, candidates=None, log=None):
    """Time the batched forward of context ``g`` (an api.YSpec with events
    set) for each (persist, chunk, blocks_per_cu); apply and return the
    fastest as a dict with every candidate's time in ms."""
    best, table = None, []
    for persist, chunk, bpc in candidates or CANDIDATES:
        g.set_batch_tuning(persist, chunk, bpc)
        g.forward_events(1)
        ms = min(g.forward_events(1) for _ in range(reps))
        table.append({"persist": persist, "chunk": chunk, "blocks_per_cu": bpc, "ms": ms})
        if log:
            log(f"tune persist={persist} chunk={chunk} blocks_per_cu={bpc}: {ms:.1f} ms")
        if best is None or ms < best["ms"]:
            best = table[-1]
    g.set_batch_tuning(best["persist"], best["chunk"], best["blocks_per_cu"])
    return {"best": best, "candidates": table}
