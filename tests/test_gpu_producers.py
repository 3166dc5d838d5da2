"""Device Hessian producers (include/mi_producers.h) against the host
generator that writes the same system: the kinetic and contact triplets bit
for bit (same arithmetic), the stable Neo-Hookean triplets to 1e-9 of the
largest entry (the host projects with a numerical 9x9 eigen-decomposition,
the device with the analytic eigensystem), and the assembled systems."""
import numpy as np
import pytest

from libuipc_amd import abi, synth
from libuipc_amd.abi import check
from libuipc_amd.device import LinearSystem

pytestmark = pytest.mark.gpu


def produce(ctx, p, B):
    nt, nc = len(p.tets), len(p.contact_kappa)
    T = B + 10 * nt + 10 * nc
    rows = ctx.alloc(T, np.int32)
    cols = ctx.alloc(T, np.int32)
    vals = ctx.alloc((T, 9), np.float64)
    keep = [ctx.to_device(a) for a in (p.mass, p.tets, p.rest_positions, p.positions)]
    dm, dt_, dX, dx = keep
    lib = ctx.lib
    i32, f64 = 4, 8
    check(lib.mi_produce_kinetic(ctx.h, B, dm.ptr, 0, rows.ptr, cols.ptr, vals.ptr, None))
    off = B
    check(lib.mi_produce_snh(ctx.h, nt, dt_.ptr, dX.ptr, dx.ptr, p.mu, p.lam, p.dt, 0,
                             rows.ptr.value + i32 * off, cols.ptr.value + i32 * off, vals.ptr.value + 9 * f64 * off,
                             None))
    off = B + 10 * nt
    if nc:
        ci, cg, ck = ctx.to_device(p.contact_idx), ctx.to_device(p.contact_g), ctx.to_device(p.contact_kappa)
        keep += [ci, cg, ck]
        check(lib.mi_produce_contact(ctx.h, nc, ci.ptr, cg.ptr, ck.ptr, p.dt, 0, rows.ptr.value + i32 * off,
                                     cols.ptr.value + i32 * off, vals.ptr.value + 9 * f64 * off, None))
    ctx.synchronize()
    return rows, cols, vals, keep


@pytest.mark.parametrize("grid,contacts", [(5, 0), (8, 300), (16, 2000)])
def test_producers_match_host_generator(ctx, grid, contacts):
    s = synth.fem_system(grid, contacts, seed=9)
    p = synth.producer_inputs(grid, contacts, seed=9)
    B = s.block_rows
    nt = len(p.tets)
    assert np.array_equal(p.tets, s.tets)
    rows, cols, vals, _ = produce(ctx, p, B)
    r, c, v = rows.numpy(), cols.numpy(), vals.numpy()
    assert np.array_equal(r, s.rows) and np.array_equal(c, s.cols)
    kin = slice(0, B)
    fem = slice(B, B + 10 * nt)
    con = slice(B + 10 * nt, None)
    assert np.array_equal(v[kin], s.vals[kin])
    assert np.array_equal(v[con], s.vals[con])
    scale = np.abs(s.vals[fem]).max()
    err = np.abs(v[fem] - s.vals[fem]).max() / scale
    assert err < 1e-9, err
    # every produced tet block is symmetric PSD as a 12x12 matrix
    K = np.zeros((nt, 12, 12))
    vv = v[fem].reshape(nt, 10, 3, 3).transpose(0, 1, 3, 2)    # [t, slot, row, col]
    rr, cc = r[fem].reshape(nt, 10), c[fem].reshape(nt, 10)
    slot = 0
    for a in range(4):
        for b in range(a, 4):
            tq = p.tets
            swap = ~(tq[:, a] < tq[:, b])
            blk = np.where(swap[:, None, None], vv[:, slot].transpose(0, 2, 1), vv[:, slot])
            K[:, 3 * a:3 * a + 3, 3 * b:3 * b + 3] = blk
            K[:, 3 * b:3 * b + 3, 3 * a:3 * a + 3] = blk.transpose(0, 2, 1)
            slot += 1
    w = np.linalg.eigvalsh(K[: min(nt, 2000)])
    assert w.min() >= -1e-9 * np.abs(w).max()


def test_produced_system_solves_like_the_host_one(ctx):
    s = synth.fem_system(8, 300, seed=4)
    p = synth.producer_inputs(8, 300, seed=4)
    rows, cols, vals, keep = produce(ctx, p, s.block_rows)
    a = LinearSystem(ctx, s.block_rows)
    a.assemble(rows, cols, vals)
    h = LinearSystem(ctx, s.block_rows)
    hr, hc, hv = ctx.to_device(s.rows), ctx.to_device(s.cols), ctx.to_device(s.vals)
    h.assemble(hr, hc, hv)
    ra, ca, va, pa = a.export_bcoo()
    rh, ch, vh, ph = h.export_bcoo()
    assert np.array_equal(ra, rh) and np.array_equal(ca, ch) and np.array_equal(pa, ph)
    assert np.abs(va - vh).max() <= 1e-9 * np.abs(vh).max()
    b = ctx.to_device(s.rhs)
    for ls in (a, h):
        ls.add_preconditioner(abi.MI_PRECOND_DIAG3, 0, s.n)
        ls.assemble_preconditioners()
    xa, xh = ctx.alloc(s.n, np.float64), ctx.alloc(s.n, np.float64)
    sa = a.solve(b, xa, criterion=abi.MI_CRIT_REL_RESIDUAL, tol=1e-8)
    sh = h.solve(b, xh, criterion=abi.MI_CRIT_REL_RESIDUAL, tol=1e-8)
    assert sa["converged"] and sh["converged"]
    d = np.linalg.norm(xa.numpy() - xh.numpy()) / np.linalg.norm(xh.numpy())
    assert d < 1e-6, d
    a.close()
    h.close()
