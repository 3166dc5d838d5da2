"""Level-count tuner for the MAS preconditioner.

The number of active MAS levels (MASPreconditionerEngine::set_active_level_num,
finite_element/mas_preconditioner_engine.h:104-110) moves the iteration count
and the cost per iteration in opposite directions, and which wins depends on
the system: only whole solves to the tolerance compare them.  The tuner times
one solve per candidate on the caller's system (meant for the first Newton
iteration after a topology change; the choice then holds until the next one),
keeps the fastest, and leaves it set on the preconditioner."""
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
