"""Batched PCG (config 4): independent FEM + contact systems of uneven sizes
on consecutive block rows, solved in the same captured iterations with
per-system alpha, beta and convergence (mi_pcg_solve_batched).  With the 3x3
block Jacobi each system must match its own FusedPCG solve (the oracle's,
and the GPU's single-system solve); with one MAS over the union (partitions
never cross systems) every system must reach its tolerance."""
import time

import numpy as np
import pytest

from libuipc_amd import abi, synth
from libuipc_amd.device import LinearSystem, rel_err
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def assembled(ctx, bt):
    ls = LinearSystem(ctx, bt.block_rows)
    r, c, v = ctx.to_device(bt.rows), ctx.to_device(bt.cols), ctx.to_device(bt.vals)
    ls.assemble(r, c, v)
    ls._keep = (r, c, v)
    return ls


def test_batch_diag3_matches_single_solves(ctx):
    bt = synth.batch_systems(10, seed=2, dof_lo=3e3, dof_hi=3e4)
    ls = assembled(ctx, bt)
    ls.add_preconditioner(abi.MI_PRECOND_DIAG3, 0, 3 * bt.block_rows)
    ls.assemble_preconditioners()
    x = ctx.alloc(3 * bt.block_rows, np.float64)
    sts = ls.solve_batched(bt.seg_rows, ctx.to_device(bt.rhs), x, criterion=abi.MI_CRIT_REL_RESIDUAL, tol=1e-6,
                           true_residual=1)
    xs = x.numpy()
    for k, s in enumerate(bt.systems):
        st = sts[k]
        seg = slice(3 * bt.seg_rows[k], 3 * bt.seg_rows[k + 1])
        assert st["converged"] == 1 and st["rel_res_true"] < 2e-6, (k, st)
        r, c, v, _ = O.convert_sym(s.block_rows, s.rows, s.cols, s.vals)
        D = O.diag3(r, c, v, 0, s.block_rows)
        xo, ro = O.pcg(s.block_rows, r, c, v, s.rhs, [dict(kind=1, blk_off=0, blk_cnt=s.block_rows, data=D)],
                       criterion=1, tol=1e-6)
        assert abs(st["iters_exact"] - ro["iters_exact"]) <= max(1, ro["iters_exact"] // 50), (k, st, ro)
        assert st["iters"] == ro["iters"] or abs(st["iters_exact"] - ro["iters_exact"]) > 0
        assert rel_err(xs[seg], xo) < 1e-6
    ls.close()


def test_batch_psd_guard_per_system(ctx):
    """one system of a batch has a singular diagonal block (z non-finite on a
    finite r): the guard solves that system as plain CG, as FusedPCG run on it
    alone would (linear_fused_pcg.cu:1843-1875); the other systems keep the
    3x3 block Jacobi and their solutions do not change.  A max_iter of 1
    stops every system before its first iteration, as the single solve does."""
    from tests.test_gpu_linsys import PSD_CASES, with_decoupled_vertex

    bt = synth.batch_systems(4, seed=5, dof_lo=3e3, dof_hi=1e4)
    systems = list(bt.systems)
    systems[2] = with_decoupled_vertex(systems[2], *PSD_CASES["singular"])
    seg = np.concatenate([[0], np.cumsum([t.block_rows for t in systems])]).astype(np.int32)
    rows = np.concatenate([t.rows + seg[k] for k, t in enumerate(systems)]).astype(np.int32)
    cols = np.concatenate([t.cols + seg[k] for k, t in enumerate(systems)]).astype(np.int32)
    vals = np.concatenate([np.asarray(t.vals).reshape(-1, 9) for t in systems])
    rhs = np.concatenate([t.rhs for t in systems])
    B = int(seg[-1])
    ls = LinearSystem(ctx, B)
    rd, cd, vd = ctx.to_device(rows), ctx.to_device(cols), ctx.to_device(vals)
    ls.assemble(rd, cd, vd)
    ls.add_preconditioner(abi.MI_PRECOND_DIAG3, 0, 3 * B)
    ls.assemble_preconditioners()
    x = ctx.alloc(3 * B, np.float64)
    kw = dict(criterion=abi.MI_CRIT_REL_RESIDUAL, tol=1e-6)
    sts = ls.solve_batched(seg, ctx.to_device(rhs), x, **kw)
    xs = x.numpy()
    for k, t in enumerate(systems):
        r, c, v, _ = O.convert_sym(t.block_rows, t.rows, t.cols, t.vals)
        D = O.diag3(r, c, v, 0, t.block_rows)
        xo, ro = O.pcg(t.block_rows, r, c, v, t.rhs, [dict(kind=1, blk_cnt=t.block_rows, data=D)], **kw)
        assert sts[k]["bypassed"] == ro["bypassed"] == (1 if k == 2 else 0), (k, sts[k], ro)
        # plain CG on the bypassed system runs ~1,500 iterations, over which
        # rounding differences of the SpMV order shift the count a little
        assert sts[k]["converged"] == 1, (k, sts[k], ro)
        assert abs(sts[k]["iters_exact"] - ro["iters_exact"]) <= max(1, ro["iters_exact"] // 50), (k, sts[k], ro)
        assert rel_err(xs[3 * seg[k]:3 * seg[k + 1]], xo) < (1e-5 if k == 2 else 1e-6)
    # without the fallback the non-finite r0.z0 is an error
    with pytest.raises(abi.MiError) as e:
        ls.solve_batched(seg, ctx.to_device(rhs), x, psd_fallback=0, **kw)
    assert e.value.code == abi.MI_ERR_NONFINITE
    # max_iter = 1 for every system: no iteration runs, x stays 0
    tiny = 1.5 / (3 * int(np.diff(seg).max()))
    sts = ls.solve_batched(seg, ctx.to_device(rhs), x, max_iter_ratio=tiny, **kw)
    for k, t in enumerate(systems):
        assert sts[k]["converged"] == 0 and sts[k]["iters"] == int(tiny * float(3 * (seg[k + 1] - seg[k]))), (k, sts[k])
    assert not np.any(x.numpy())
    ls.close()


def test_batch_mas_union(ctx):
    bt = synth.batch_systems(12, seed=3, dof_lo=3e3, dof_hi=6e4)
    ls = assembled(ctx, bt)
    ls.add_preconditioner(abi.MI_PRECOND_MAS, 0, 3 * bt.block_rows, part_ids=bt.part_ids, tets=bt.tets,
                          build=abi.MI_MAS_BUILD_DEVICE)
    ls.assemble_preconditioners()
    x = ctx.alloc(3 * bt.block_rows, np.float64)
    sts = ls.solve_batched(bt.seg_rows, ctx.to_device(bt.rhs), x, criterion=abi.MI_CRIT_REL_RESIDUAL, tol=1e-4,
                           true_residual=1)
    for k, s in enumerate(bt.systems):
        assert sts[k]["converged"] == 1 and sts[k]["rel_res_true"] < 2e-4, (k, sts[k])
        # the same system alone with its own MAS
        one = LinearSystem(ctx, s.block_rows)
        r, c, v = ctx.to_device(s.rows), ctx.to_device(s.cols), ctx.to_device(s.vals)
        one.assemble(r, c, v)
        one.add_preconditioner(abi.MI_PRECOND_MAS, 0, s.n, part_ids=s.part_ids, tets=s.tets)
        one.assemble_preconditioners()
        xo = ctx.alloc(s.n, np.float64)
        so = one.solve(ctx.to_device(s.rhs), xo, criterion=abi.MI_CRIT_REL_RESIDUAL, tol=1e-4)
        assert sts[k]["iters_exact"] <= 1.25 * so["iters_exact"] + 5, (k, sts[k], so)
        one.close()
    ls.close()


def test_batch_mas_per_system_levels(ctx):
    """A MAS level count per system (mi_precond_mas_set_system_levels): every
    system of the batch takes the iterations and the solution it takes when
    the whole batch runs at its count (the union's clusters and coarse levels
    never mix systems).  The batched tuner keeps, per system, the count with
    the fewest iterations and leaves it set; mi_precond_mas_set_active_levels
    returns to one count."""
    from libuipc_amd.tune import tune_mas_levels_batched

    bt = synth.batch_systems(6, seed=7, dof_lo=3e3, dof_hi=3e4)
    ls = assembled(ctx, bt)
    h = ls.add_preconditioner(abi.MI_PRECOND_MAS, 0, 3 * bt.block_rows, part_ids=bt.part_ids, tets=bt.tets,
                              build=abi.MI_MAS_BUILD_DEVICE)
    ls.assemble_preconditioners()
    L = ls.mas_export(h, matrices=False)["L"]
    cand = sorted({1, 2, L} & set(range(1, L + 1)))
    assert len(cand) >= 2
    b = ctx.to_device(bt.rhs)
    kw = dict(criterion=abi.MI_CRIT_REL_RESIDUAL, tol=1e-6)
    ref = {}
    for lv in cand:
        ls.mas_set_active_levels(h, lv)
        x = ctx.alloc(3 * bt.block_rows, np.float64)
        ref[lv] = (ls.solve_batched(bt.seg_rows, b, x, **kw), x.numpy())
    levels = [cand[k % len(cand)] for k in range(len(bt.systems))]
    ls.mas_set_system_levels(h, bt.seg_rows, levels)
    x = ctx.alloc(3 * bt.block_rows, np.float64)
    sts = ls.solve_batched(bt.seg_rows, b, x, **kw)
    xs = x.numpy()
    for k, lv in enumerate(levels):
        seg = slice(3 * bt.seg_rows[k], 3 * bt.seg_rows[k + 1])
        it = ref[lv][0][k]["iters_exact"]
        assert sts[k]["converged"] == 1 and abs(sts[k]["iters_exact"] - it) <= max(2, it // 50), (k, lv, sts[k], it)
        assert rel_err(xs[seg], ref[lv][1][seg]) < 1e-5, (k, lv)
    t = tune_mas_levels_batched(ls, h, bt.seg_rows, b, x, levels=cand, **kw)
    for k, lv in enumerate(t["levels"]):
        assert t["table"][lv][k] == min(t["table"][c][k] for c in cand), (k, t)
    sts = ls.solve_batched(bt.seg_rows, b, x, **kw)
    assert all(st["converged"] == 1 for st in sts)
    ls.mas_set_active_levels(h, L)
    sts = ls.solve_batched(bt.seg_rows, b, x, **kw)
    for k, st in enumerate(sts):
        it = ref[L][0][k]["iters_exact"]
        assert abs(st["iters_exact"] - it) <= max(2, it // 50), (k, st, it)
    with pytest.raises(abi.MiError):
        ls.mas_set_system_levels(h, [0, bt.block_rows // 2], [1])  # rows outside the systems
    ls.close()


@pytest.mark.parametrize("precision", [abi.MI_FP64, abi.MI_FP32], ids=["fp64", "fp32"])
def test_batch_mas_skips_stopped_systems(ctx, precision):
    """The SpMV chunks and the MAS fine clusters of stopped systems are
    skipped.  Solving the same batch (same assembled MAS) with one system's
    rhs zeroed stops that system at once, so its clusters are skipped for the
    whole solve: every other system must reach the same solution (to the
    solve's tolerance: the SpMV adds its in-window transposed products with
    LDS atomics, so the rounding differs from run to run, as with the
    reference's FP64 atomics), in the same iterations (within 2%), and the
    zeroed system must return x = 0."""
    bt = synth.batch_systems(8, seed=5, dof_lo=3e3, dof_hi=3e4)
    ls = LinearSystem(ctx, bt.block_rows, precision)
    r, c, v = ctx.to_device(bt.rows), ctx.to_device(bt.cols), ctx.to_device(bt.vals)
    ls.assemble(r, c, v)
    ls.add_preconditioner(abi.MI_PRECOND_MAS, 0, 3 * bt.block_rows, part_ids=bt.part_ids, tets=bt.tets,
                          build=abi.MI_MAS_BUILD_DEVICE, mas_precision=precision)
    ls.assemble_preconditioners()
    kw = dict(criterion=abi.MI_CRIT_REL_RESIDUAL, tol=1e-6, true_residual=1)
    x1, x2 = ctx.alloc(3 * bt.block_rows, np.float64), ctx.alloc(3 * bt.block_rows, np.float64)
    s1 = ls.solve_batched(bt.seg_rows, ctx.to_device(bt.rhs), x1, **kw)
    j = 3
    segj = slice(3 * bt.seg_rows[j], 3 * bt.seg_rows[j + 1])
    rhs2 = bt.rhs.copy()
    rhs2[segj] = 0.0
    s2 = ls.solve_batched(bt.seg_rows, ctx.to_device(rhs2), x2, **kw)
    a, b = x1.numpy(), x2.numpy()
    tol = 1e-5 if precision == abi.MI_FP64 else 1e-4
    for k in range(len(bt.systems)):
        seg = slice(3 * bt.seg_rows[k], 3 * bt.seg_rows[k + 1])
        if k == j:
            assert s2[k]["converged"] == 1 and s2[k]["iters_exact"] == 0 and not b[seg].any()
            continue
        assert s1[k]["converged"] == 1 and s2[k]["converged"] == 1, (k, s1[k], s2[k])
        assert abs(s1[k]["iters_exact"] - s2[k]["iters_exact"]) <= max(2, s1[k]["iters_exact"] // 50), (k, s1[k], s2[k])
        assert rel_err(b[seg], a[seg]) < tol, k
    ls.close()


def test_batch_config4_recipe(ctx):
    """config 4's recipe at a tenth of its size range: 64 systems, DoF
    log-uniform in [3e3, 6e4] (the full range, up to 6e5 DoFs per system,
    needs ~100 GB of host-generated triplets), MAS over the union"""
    t0 = time.perf_counter()
    bt = synth.batch_systems(64, seed=4, dof_lo=3e3, dof_hi=6e4)
    tgen = time.perf_counter() - t0
    ls = assembled(ctx, bt)
    ls.add_preconditioner(abi.MI_PRECOND_MAS, 0, 3 * bt.block_rows, part_ids=bt.part_ids, tets=bt.tets,
                          build=abi.MI_MAS_BUILD_DEVICE)
    ls.assemble_preconditioners()
    x = ctx.alloc(3 * bt.block_rows, np.float64)
    b = ctx.to_device(bt.rhs)
    ls.solve_batched(bt.seg_rows, b, x, criterion=abi.MI_CRIT_REL_RESIDUAL, tol=1e-4)  # warm-up (captures the graph)
    ctx.synchronize()
    t0 = time.perf_counter()
    sts = ls.solve_batched(bt.seg_rows, b, x, criterion=abi.MI_CRIT_REL_RESIDUAL, tol=1e-4)
    ctx.synchronize()
    ms = (time.perf_counter() - t0) * 1e3
    it = np.array([s["iters_exact"] for s in sts])
    assert all(s["converged"] == 1 for s in sts)
    print(f"\n[batch] {len(bt.systems)} systems, {3 * bt.block_rows} DoFs, {len(bt.rows)} triplets (generated in {tgen:.0f} s): "
          f"{ms:.1f} ms for the batch, iterations min {it.min()} median {int(np.median(it))} max {it.max()}",
          flush=True)
    ls.close()
