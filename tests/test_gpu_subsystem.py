"""The FEM-subsystem steps on either side of the solve, through the C ABI:
the fixed-vertex treatment of gradient and Hessian triplets
(fem_linear_subsystem.cu:23-53) and dx = -x (fem_linear_subsystem.cu:119-126)."""
import numpy as np
import pytest

from libuipc_amd import abi, synth
from libuipc_amd.abi import check
from libuipc_amd.device import LinearSystem

pytestmark = pytest.mark.gpu


def test_fix_vertices_triplets_and_gradient(ctx):
    s = synth.fem_system(6, 60, seed=4)
    rng = np.random.default_rng(1)
    V = s.block_rows
    fixed = (rng.random(V) < 0.2).astype(np.int32)
    rows, cols, vals = ctx.to_device(s.rows), ctx.to_device(s.cols), ctx.to_device(s.vals)
    g = ctx.to_device(s.rhs)
    fx = ctx.to_device(fixed)
    check(ctx.lib.mi_subsystem_fix_vertices(ctx.h, s.n_triplets, rows.ptr, cols.ptr, vals.ptr, 0, fx.ptr, V, g.ptr,
                                            None))
    ctx.synchronize()
    v = vals.numpy().reshape(-1, 9)
    touched = (fixed[s.rows] != 0) | (fixed[s.cols] != 0)
    diag = touched & (s.rows == s.cols)
    eye = np.eye(3).reshape(-1, order="F")
    assert np.array_equal(v[~touched], s.vals.reshape(-1, 9)[~touched])   # untouched triplets keep their bits
    assert np.array_equal(v[diag], np.broadcast_to(eye, (diag.sum(), 9)))
    assert not v[touched & ~diag].any()
    gg = g.numpy().reshape(-1, 3)
    assert not gg[fixed != 0].any()
    assert np.array_equal(gg[fixed == 0], s.rhs.reshape(-1, 3)[fixed == 0])
    # the fixed rows decouple: PCG leaves their unknowns exactly 0
    ls = LinearSystem(ctx, V)
    ls.assemble(rows, cols, vals)
    ls.add_preconditioner(abi.MI_PRECOND_DIAG3, 0, s.n)
    ls.assemble_preconditioners()
    x = ctx.alloc(s.n, np.float64)
    st = ls.solve(g, x, criterion=abi.MI_CRIT_REL_RESIDUAL, tol=1e-8, true_residual=1)
    xs = x.numpy().reshape(-1, 3)
    assert st["converged"] == 1 and st["rel_res_true"] < 2e-8
    assert not xs[fixed != 0].any()
    ls.close()


def test_fix_vertices_offset_and_unwritten_slots(ctx):
    # a subsystem at block offset 5 of a larger system; slot 2 is unwritten (row -1)
    rows = np.array([5, 5, -1, 6, 7, 9], np.int32)
    cols = np.array([5, 6, -1, 6, 9, 9], np.int32)
    vals = np.arange(6 * 9, dtype=np.float64) + 1.0
    fixed = np.array([0, 1, 0, 0, 1], np.int32)        # local vertices 1 (global 6) and 4 (global 9)
    r, c, v, fx = ctx.to_device(rows), ctx.to_device(cols), ctx.to_device(vals), ctx.to_device(fixed)
    check(ctx.lib.mi_subsystem_fix_vertices(ctx.h, 6, r.ptr, c.ptr, v.ptr, 5, fx.ptr, 5, None, None))
    out = v.numpy().reshape(6, 9)
    want = vals.reshape(6, 9).copy()
    want[1] = 0.0                                       # (5, 6): off-diagonal, 6 fixed
    want[3] = np.eye(3).reshape(-1, order="F")          # (6, 6): fixed diagonal
    want[4] = 0.0                                       # (7, 9)
    want[5] = np.eye(3).reshape(-1, order="F")          # (9, 9)
    assert np.array_equal(out, want)


def test_distribute_solution(ctx):
    x = np.random.default_rng(2).standard_normal(300)
    d = ctx.to_device(x)
    out = ctx.alloc(60, np.float64)
    check(ctx.lib.mi_subsystem_distribute(ctx.h, d.ptr, 30, 20, out.ptr, None))
    assert np.array_equal(out.numpy(), -x[30:90])
