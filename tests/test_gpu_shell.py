"""mi_produce_bw_shell and mi_produce_bending against the oracle's
restatements of the strain-limiting Baraff-Witkin shell and the discrete shell
hinge: triplet rows and columns bit for bit, values to 1e-10 of the largest
entry (same closed forms, different summation order), and the assembled
256 x 256 cloth (config 3's grid: stretch/shear + bending + kinetic) solved
with the 3x3 block Jacobi preconditioner."""
import numpy as np
import pytest

from libuipc_amd import abi, synth
from libuipc_amd.abi import check
from libuipc_amd.device import LinearSystem
from oracle import oracle as O
from oracle import shell as OS

pytestmark = pytest.mark.gpu


def produce(ctx, c, vert_offset=0):
    T = len(c.tris)
    rows, cols, vals = ctx.alloc(6 * T, np.int32), ctx.alloc(6 * T, np.int32), ctx.alloc((6 * T, 9), np.float64)
    keep = [ctx.to_device(a) for a in (c.tris, c.rest_positions, c.positions)]
    check(ctx.lib.mi_produce_bw_shell(ctx.h, T, keep[0].ptr, keep[1].ptr, keep[2].ptr, c.stretch, c.shear,
                                      c.strain_rate, c.dt, vert_offset, rows.ptr, cols.ptr, vals.ptr, None))
    ctx.synchronize()
    return rows, cols, vals


@pytest.mark.parametrize("seed,noise", [(1, 0.05), (2, 0.4)])
def test_bw_shell_matches_oracle(ctx, seed, noise):
    c = synth.cloth_inputs(24, seed=seed, noise=noise)
    # reverse some triangles so that the transposed half-block path is taken
    c.tris[::3] = c.tris[::3, ::-1]
    rows, cols, vals = produce(ctx, c, vert_offset=5)
    r, cc, v = OS.triplets(c.tris, c.rest_positions, c.positions, c.stretch, c.shear, c.strain_rate, c.dt,
                           vert_offset=5)
    assert np.array_equal(rows.numpy(), r) and np.array_equal(cols.numpy(), cc)
    assert np.abs(vals.numpy() - v).max() <= 1e-10 * np.abs(v).max()


def produce_bending(ctx, c, hinges, vert_offset=0):
    H = len(hinges)
    rows, cols, vals = ctx.alloc(10 * H, np.int32), ctx.alloc(10 * H, np.int32), ctx.alloc((10 * H, 9), np.float64)
    keep = [ctx.to_device(a) for a in (hinges, c.rest_positions, c.positions)]
    check(ctx.lib.mi_produce_bending(ctx.h, H, keep[0].ptr, keep[1].ptr, keep[2].ptr, c.bending, c.dt, vert_offset,
                                     rows.ptr, cols.ptr, vals.ptr, None))
    ctx.synchronize()
    return rows, cols, vals


@pytest.mark.parametrize("seed,noise", [(1, 0.05), (2, 0.4)])
def test_bending_matches_oracle(ctx, seed, noise):
    c = synth.cloth_inputs(20, seed=seed, noise=noise)
    hg = synth.hinges_of(c.tris)
    hg[::4] = hg[::4, ::-1]  # reversed stencils take the transposed path
    rows, cols, vals = produce_bending(ctx, c, hg, vert_offset=3)
    r, cc, v = OS.bending_triplets(hg, c.rest_positions, c.positions, c.bending, c.dt, vert_offset=3)
    assert np.array_equal(rows.numpy(), r) and np.array_equal(cols.numpy(), cc)
    assert np.abs(vals.numpy() - v).max() <= 1e-10 * np.abs(v).max()


def test_bw_cloth_256_solve(ctx):
    c = synth.cloth_inputs(256, seed=3)
    V, T = len(c.rest_positions), len(c.tris)
    rows, cols, vals = produce(ctx, c)
    # kinetic blocks m I (density 200 kg/m^3, thickness 2 mm, lumped)
    mass = np.full(V, 200.0 * 2e-3 * 1.0 / V)
    mr, mc, mv = ctx.alloc(V, np.int32), ctx.alloc(V, np.int32), ctx.alloc((V, 9), np.float64)
    dm = ctx.to_device(mass)
    check(ctx.lib.mi_produce_kinetic(ctx.h, V, dm.ptr, 0, mr.ptr, mc.ptr, mv.ptr, None))
    hg = synth.hinges_of(c.tris)
    br, bc, bv = produce_bending(ctx, c, hg)
    r = np.concatenate([mr.numpy(), rows.numpy(), br.numpy()])
    cc = np.concatenate([mc.numpy(), cols.numpy(), bc.numpy()])
    v = np.concatenate([mv.numpy(), vals.numpy(), bv.numpy()])
    ls = LinearSystem(ctx, V)
    rd, cd, vd = ctx.to_device(r), ctx.to_device(cc), ctx.to_device(v)
    ls.assemble(rd, cd, vd)
    ls.add_preconditioner(abi.MI_PRECOND_DIAG3, 0, 3 * V)
    ls.assemble_preconditioners()
    rhs = np.random.default_rng(4).uniform(-1, 1, 3 * V)
    x = ctx.alloc(3 * V, np.float64)
    st = ls.solve(ctx.to_device(rhs), x, criterion=abi.MI_CRIT_REL_RESIDUAL, tol=1e-6, true_residual=1)
    assert st["converged"] == 1 and st["rel_res_true"] < 2e-6
    # the same system on the oracle: residual of the GPU solution
    orr, oc, ov, _ = O.convert_sym(V, r, cc, v)
    Ax, _ = O.spmv_sym(V, orr, oc, ov, x.numpy())
    assert np.linalg.norm(rhs - Ax) <= 2e-6 * np.linalg.norm(rhs)
    print(f"\n[cloth] 256x256: {T} triangles, {len(hg)} hinges, {st['iters_exact']} PCG iterations to 1e-6",
          flush=True)
    ls.close()
