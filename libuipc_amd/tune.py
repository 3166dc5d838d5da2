"""Tuners: the MAS level count and the SpMV launch shape.

The number of active MAS levels (MASPreconditionerEngine::set_active_level_num,
finite_element/mas_preconditioner_engine.h:104-110) moves the iteration count
and the cost per iteration in opposite directions, and which wins depends on
the system: only whole solves to the tolerance compare them.  The tuner times
one solve per candidate on the caller's system (meant for the first Newton
iteration after a topology change; the choice then holds until the next one),
keeps the fastest, and leaves it set on the preconditioner.

The SpMV launch shape (threads per workgroup, rows per window, blocks per lane
and stage) changes the window layout and the kernel's occupancy but not the
product; the shape tuner times products on the caller's matrix and keeps the
fastest."""
from __future__ import annotations

import time

from .device import LinearSystem


def tune_mas_levels(ls: LinearSystem, handle: int, b, x, levels=None, repeats: int = 1, **solve_kw) -> dict:
    """Solves A x = b once per candidate level count (`levels`, default 1 ..
    the hierarchy's count) after one untimed solve each (graph capture), and
    sets the fastest.  Returns {"best": L, "table": {L: {"ms", "iters"}}};
    a candidate whose solve did not converge is never chosen."""
    n_levels = ls.mas_export(handle, matrices=False)["L"]
    cand = sorted({int(v) for v in (levels or range(1, n_levels + 1)) if 1 <= int(v) <= n_levels})
    if not cand:
        raise ValueError("tune_mas_levels: no candidate level count in 1..%d" % n_levels)
    table = {}
    for lv in cand:
        ls.mas_set_active_levels(handle, lv)
        ls.solve(b, x, **solve_kw)
        ls.ctx.synchronize()
        best_ms, st = float("inf"), None
        for _ in range(max(1, repeats)):
            t0 = time.perf_counter()
            st = ls.solve(b, x, **solve_kw)
            ls.ctx.synchronize()
            best_ms = min(best_ms, (time.perf_counter() - t0) * 1e3)
        table[lv] = {"ms": round(best_ms, 3), "iters": int(st["iters_exact"]), "converged": int(st["converged"])}
    ok = [lv for lv in cand if table[lv]["converged"]] or cand
    best = min(ok, key=lambda lv: table[lv]["ms"])
    ls.mas_set_active_levels(handle, best)
    return {"best": best, "table": table}


def tune_mas_levels_batched(ls: LinearSystem, handle: int, seg_rows, b, x, levels=None, **solve_kw) -> dict:
    """Per-system MAS level counts for a batched solve: the batch is solved
    once per candidate count (every system at that count) and each system
    keeps the count that took it the fewest iterations (ties: fewer levels,
    the cheaper iteration); the per-system counts are left set
    (mi_precond_mas_set_system_levels).  Returns {"levels": [per system],
    "table": {L: [iterations per system]}, "uniform_best": the single count
    with the fewest iterations of the slowest system}."""
    n_levels = ls.mas_export(handle, matrices=False)["L"]
    cand = sorted({int(v) for v in (levels or range(1, n_levels + 1)) if 1 <= int(v) <= n_levels})
    if not cand:
        raise ValueError("tune_mas_levels_batched: no candidate level count in 1..%d" % n_levels)
    table = {}
    for lv in cand:
        ls.mas_set_active_levels(handle, lv)
        sts = ls.solve_batched(seg_rows, b, x, **solve_kw)
        table[lv] = [int(st["iters_exact"]) if st["converged"] else 1 << 30 for st in sts]
    nsys = len(table[cand[0]])
    per = [min(cand, key=lambda lv: (table[lv][k], lv)) for k in range(nsys)]
    uniform = min(cand, key=lambda lv: (max(table[lv]), lv))
    ls.mas_set_system_levels(handle, seg_rows, per)
    return {"levels": per, "table": table, "uniform_best": uniform}


def tune_spmv_shape(ls: LinearSystem, x, reps: int = 20, shapes=None, margin: float = 0.02) -> dict:
    """Times `reps` products per SpMV layout -- the full BSR
    (MI_SPMV_SHAPE_FULL) and every window launch shape, each layout built for
    it, one untimed product first; the PCG's own product (mi_spmv_ordered on
    a window layout, y = A x on the full BSR) -- and keeps the fastest, as
    the reference fits its SpMV grid to each matrix (Spmv::fit_block_count,
    linear_system/spmv.cu:909-933).  The system's layout stays unless another
    is faster by more than `margin` (timing noise would otherwise flip
    between near-equal layouts).  Returns {"best": k, "table": {k: {"us",
    "threads", "rows", "per", "vec"}}}."""
    import numpy as np

    from .device import spmv_shapes
    from .hip_events import EventTimer

    from .abi import MI_SPMV_SHAPE_FULL

    info = spmv_shapes(ls.lib)["shapes"]
    lay0 = ls.spmv_layout()
    dflt = lay0["shape"] if lay0["layout"] == "window" else MI_SPMV_SHAPE_FULL  # the system's own layout
    from .abi import MI_FP64

    if shapes is None:
        # an FP64 system runs a four-blocks-per-load shape as its two-block
        # twin (16-byte loads hold two FP64 values): not a layout of its own
        cand = [MI_SPMV_SHAPE_FULL] + [k for k in range(len(info))
                                       if not (ls.precision == MI_FP64 and info[k]["vec"] == 4)]
    else:
        cand = [int(k) for k in shapes]
    ctx = ls.ctx
    stream = ctx.stream()
    y = ctx.alloc(ls.n, np.float64)
    ev = EventTimer(stream)
    table = {}
    for k in cand:
        ls.set_spmv_shape(k, stream)
        if k >= 0:
            def product():
                ctx.lib.mi_spmv_ordered(ls.h, x.ptr, y.ptr, None, stream)
            desc = info[k]
        else:
            def product():
                ctx.lib.mi_spmv(ls.h, 1.0, x.ptr, 0.0, y.ptr, stream)
            desc = {"threads": 256, "rows": 0, "per": 0, "vec": 0, "layout": "full"}
        product()
        ev.start()
        for _ in range(max(1, reps)):
            product()
        table[k] = {"us": round(1e3 * ev.stop() / max(1, reps), 2), **desc}
    best = min(table, key=lambda k: table[k]["us"])
    if dflt in table and table[best]["us"] > (1.0 - margin) * table[dflt]["us"]:
        best = dflt
    ls.set_spmv_shape(best, stream)
    ctx.synchronize()
    return {"best": best, "table": table}
