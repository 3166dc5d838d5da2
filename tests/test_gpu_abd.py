"""mi_produce_abd against the oracle's affine-body restatement: the triplet
layout bit for bit (rows, cols, zero lower blocks, identity for fixed
bodies), the values to 1e-9 of the largest entry (the device projects the
9x9 ortho Hessian with its own Jacobi eigen-decomposition, the oracle with
numpy's eigh), and a solve of the assembled system with the 12x12 block
Jacobi preconditioner (ABDDiagPreconditioner)."""
import numpy as np
import pytest

from libuipc_amd import abi, synth
from libuipc_amd.abi import check
from libuipc_amd.device import LinearSystem, rel_err
from oracle import abd as OA
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def produce(ctx, a, fixed=None, body_offset=0):
    n = len(a.q)
    rows, cols, vals = ctx.alloc(16 * n, np.int32), ctx.alloc(16 * n, np.int32), ctx.alloc((16 * n, 9), np.float64)
    keep = [ctx.to_device(x) for x in (a.q, a.dyadic, a.volume, a.kappa)]
    fx = ctx.to_device(fixed.astype(np.int32)) if fixed is not None else None
    check(ctx.lib.mi_produce_abd(ctx.h, n, keep[0].ptr, keep[1].ptr, keep[2].ptr, keep[3].ptr,
                                 fx.ptr if fx is not None else None, a.dt, body_offset, rows.ptr, cols.ptr, vals.ptr,
                                 None))
    ctx.synchronize()
    return rows, cols, vals


@pytest.mark.parametrize("noise", [1e-2, 0.3])
def test_abd_matches_oracle(ctx, noise):
    a = synth.abd_inputs(300, seed=5, noise=noise)
    fixed = np.zeros(300, bool)
    fixed[::17] = True
    rows, cols, vals = produce(ctx, a, fixed, body_offset=11)
    r, c, v = OA.triplets(a.q, a.dyadic, a.volume, a.kappa, a.dt, fixed, body_offset=11)
    assert np.array_equal(rows.numpy(), r) and np.array_equal(cols.numpy(), c)
    got = vals.numpy()
    assert not got[r > c].any()
    assert np.abs(got - v).max() <= 1e-9 * np.abs(v).max()
    # fixed bodies: exactly the identity's upper blocks
    I = np.nonzero(fixed)[0][1]
    blk = got[16 * I:16 * I + 16]
    assert np.array_equal(blk, v[16 * I:16 * I + 16])


def test_abd_system_solve(ctx):
    """10,000 bodies (config 3's count) assembled and solved with DIAG12"""
    a = synth.abd_inputs(10_000, seed=6)
    n = len(a.q)
    rows, cols, vals = produce(ctx, a)
    B = 4 * n
    ls = LinearSystem(ctx, B)
    ls.assemble(rows, cols, vals)
    ls.add_preconditioner(abi.MI_PRECOND_DIAG12, 0, 3 * B)
    ls.assemble_preconditioners()
    rhs = np.random.default_rng(1).uniform(-1, 1, 3 * B)
    x = ctx.alloc(3 * B, np.float64)
    st = ls.solve(ctx.to_device(rhs), x, criterion=abi.MI_CRIT_REL_RESIDUAL, tol=1e-10, true_residual=1)
    assert st["converged"] == 1 and st["rel_res_true"] < 1e-9
    # block diagonal system: the solve is the per-body 12x12 solve
    r, c, v = OA.triplets(a.q[:3], a.dyadic[:3], a.volume[:3], a.kappa[:3], a.dt)
    orr, oc, ov, _ = O.convert_sym(12, r, c, v)
    xs = x.numpy()[:36]
    for I in range(3):
        H = OA.body_hessian(a.q[I], a.dyadic[I], a.volume[I], a.kappa[I], a.dt)
        assert rel_err(xs[12 * I:12 * I + 12], np.linalg.solve(H, rhs[12 * I:12 * I + 12])) < 1e-8
    ls.close()
