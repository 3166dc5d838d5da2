"""The library-owned triplet buffer and gradient (GlobalLinearSystem::Impl::
_update_subsystem_extent and the fill of _assemble_linear_system,
global_linear_system.cu:321-497) through the C ABI: per-subsystem DoF and
triplet offsets in registration order, capacity growth by the reserve ratio,
the -1 fill that drops unwritten slots, a DoF change that resizes the system,
and an assembly and solve of the owned buffer against the oracle."""
import ctypes as C

import numpy as np
import pytest

from libuipc_amd import abi, synth
from libuipc_amd.device import DeviceArray, LinearSystem, rel_err
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _put(ctx, ptr, offset, arr):
    """copy a host array into a device buffer at element offset"""
    a = np.ascontiguousarray(arr)
    if a.size:
        abi.check(ctx.lib.mi_memcpy_h2d(ctx.h, C.c_void_p(ptr + offset * a.itemsize), a.ctypes.data_as(C.c_void_p),
                                        a.nbytes))


def _get(ctx, ptr, n, dtype):
    out = np.empty(n, dtype)
    if n:
        abi.check(ctx.lib.mi_memcpy_d2h(ctx.h, out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), out.nbytes))
    return out


def test_extents_offsets_growth_and_fill(ctx):
    ls = LinearSystem(ctx, 0, reserve_ratio=1.1)
    ext = [dict(is_diag=1, dof_count=30, block_count=100),
           dict(is_diag=0, block_count=7, rl_block_count=5, l=0, r=2),
           dict(is_diag=1, dof_count=12, block_count=40)]
    views, T, cap = ls.update_extents(ext)
    assert T == 152 and cap == int(152 * 1.1)
    assert ls.B == 14
    assert [v["dof_offset"] for v in views] == [0, -1, 30]
    assert [v["triplet_offset"] for v in views] == [0, 100, 112]
    assert views[1]["rl_triplet_offset"] == 107
    assert (views[0]["row_offset"], views[0]["col_offset"]) == (0, 0)
    assert (views[1]["row_offset"], views[1]["col_offset"]) == (0, 10)
    assert (views[2]["row_offset"], views[2]["col_offset"]) == (10, 10)
    bufs = ls.begin_assembly()
    ctx.synchronize()
    assert np.all(_get(ctx, bufs["rows"], T, np.int32) == -1) and np.all(_get(ctx, bufs["cols"], T, np.int32) == -1)
    assert not np.any(_get(ctx, bufs["b"], 42, np.float64))
    # fewer triplets: same capacity and buffers; more than the capacity: x1.1 of the total
    ext[0]["block_count"] = 90
    _, T2, cap2 = ls.update_extents(ext)
    assert T2 == 142 and cap2 == cap and ls.begin_assembly()["rows"] == bufs["rows"]
    ext[2]["block_count"] = 400
    _, T3, cap3 = ls.update_extents(ext)
    assert T3 == 502 and cap3 == int(502 * 1.1)
    # bad extents
    with pytest.raises(abi.MiError):
        ls.update_extents([dict(is_diag=1, dof_count=4, block_count=1)])
    with pytest.raises(abi.MiError):
        ls.update_extents([dict(is_diag=0, block_count=1, l=0, r=0)])
    ls.close()


def test_owned_assembly_and_solve_match_the_oracle(ctx):
    """two FEM subsystems and a coupling between them, written into their
    slices with their block offsets; the second subsystem reserves more
    triplets than it writes (those slots stay -1 and are dropped); the
    assembled BCOO is bit-exact with the oracle's conversion of the same
    global triplet array, and MAS-free PCG on the owned b matches the oracle"""
    a = synth.fem_system(8, 1500, seed=11)
    c = synth.fem_system(6, 600, seed=12)
    rng = np.random.default_rng(5)
    n_cpl = 300
    cl = rng.integers(0, a.block_rows, n_cpl).astype(np.int32)
    cr = rng.integers(0, c.block_rows, n_cpl).astype(np.int32)
    g = rng.standard_normal((n_cpl, 3))
    # a PSD coupling kappa [g; -g][g; -g]^T per pair: its diagonal blocks are
    # written by the two diagonal subsystems, its off-diagonal ones by the
    # coupling (lr; the rl transposes are lower blocks, dropped)
    gg = 1e2 * np.einsum("ki,kj->kij", g, g).reshape(n_cpl, 9)
    cv = -gg
    spare = 77
    ext = [dict(is_diag=1, dof_count=a.n, block_count=a.n_triplets + n_cpl),
           dict(is_diag=1, dof_count=c.n, block_count=c.n_triplets + n_cpl + spare),
           dict(is_diag=0, block_count=n_cpl, rl_block_count=n_cpl, l=0, r=1)]
    ls = LinearSystem(ctx, 0)
    views, T, _ = ls.update_extents(ext)
    B = a.block_rows + c.block_rows
    assert ls.B == B and T == a.n_triplets + c.n_triplets + spare + 4 * n_cpl
    bufs = ls.begin_assembly()
    # what each subsystem writes into its slice: indices + its block offsets
    host_rows = np.full(T, -1, np.int32)
    host_cols = np.full(T, -1, np.int32)
    host_vals = np.zeros((T, 9))
    def write(off, r, cc, v):
        host_rows[off:off + r.size], host_cols[off:off + r.size] = r, cc
        host_vals[off:off + r.size] = v
        _put(ctx, bufs["rows"], off, r.astype(np.int32))
        _put(ctx, bufs["cols"], off, cc.astype(np.int32))
        _put(ctx, bufs["vals"], 9 * off, np.asarray(v, np.float64).reshape(-1))
    va, vc, vx = views
    write(va["triplet_offset"], a.rows + va["row_offset"], a.cols + va["col_offset"], a.vals)
    write(va["triplet_offset"] + a.n_triplets, cl + va["row_offset"], cl + va["col_offset"], gg)
    write(vc["triplet_offset"], c.rows + vc["row_offset"], c.cols + vc["col_offset"], c.vals)
    write(vc["triplet_offset"] + c.n_triplets, cr + vc["row_offset"], cr + vc["col_offset"], gg)
    write(vx["triplet_offset"], cl + vx["row_offset"], cr + vx["col_offset"], cv)
    cvt = cv.reshape(-1, 3, 3).transpose(0, 2, 1).reshape(-1, 9)
    write(vx["rl_triplet_offset"], cr + vx["col_offset"], cl + vx["row_offset"], cvt)
    rhs = np.concatenate([a.rhs, c.rhs])
    _put(ctx, bufs["b"], va["dof_offset"], a.rhs)
    _put(ctx, bufs["b"], vc["dof_offset"], c.rhs)
    ls.assemble_owned()
    r, cc, v, rp = O.convert_sym(B, host_rows, host_cols, host_vals)
    gr, gc, gv, grp = ls.export_bcoo()
    assert np.array_equal(gr, r) and np.array_equal(gc, cc) and np.array_equal(grp, rp)
    assert np.array_equal(gv.view(np.uint64), v.view(np.uint64))
    ls.add_preconditioner(abi.MI_PRECOND_DIAG3, 0, 3 * B)
    ls.assemble_preconditioners()
    x = ctx.alloc(3 * B, np.float64)
    bd = DeviceArray(ctx, 3 * B, np.float64, borrow=bufs["b"])
    st = ls.solve(bd, x, criterion=abi.MI_CRIT_REL_RESIDUAL, tol=1e-8, true_residual=1)
    D = O.diag3(r, cc, v, 0, B)
    xo, ro = O.pcg(B, r, cc, v, rhs, [dict(kind=1, blk_off=0, blk_cnt=B, data=D)], criterion=1, tol=1e-8)
    assert st["converged"] == 1 and st["rel_res_true"] < 2e-8
    assert abs(st["iters_exact"] - ro["iters_exact"]) <= max(1, ro["iters_exact"] // 50)
    assert rel_err(x.numpy(), xo) < 1e-6
    # a DoF change resizes the system: drop the coupling and the second mesh
    views, T, _ = ls.update_extents([dict(is_diag=1, dof_count=a.n, block_count=a.n_triplets)])
    assert ls.B == a.block_rows and T == a.n_triplets
    bufs = ls.begin_assembly()
    _put(ctx, bufs["rows"], 0, a.rows)
    _put(ctx, bufs["cols"], 0, a.cols)
    _put(ctx, bufs["vals"], 0, np.asarray(a.vals, np.float64).reshape(-1))
    ls.assemble_owned()
    r, cc, v, rp = O.convert_sym(a.block_rows, a.rows, a.cols, a.vals)
    gr, gc, gv, grp = ls.export_bcoo()
    assert np.array_equal(gr, r) and np.array_equal(gc, cc) and np.array_equal(grp, rp)
    assert np.array_equal(gv.view(np.uint64), v.view(np.uint64))
    ls.close()
