"""The headline system at its exact size -- the 64^3-vertex grid (1,250,235 SNH
tets) plus 2,000,000 vertex-triangle contact stencils, seed 1, 32,764,494
triplets -- through the C ABI against the oracle: the assembled upper BCOO
bit for bit, the SpMV, the MAS tables bit for bit and its apply, and MAS-PCG
to ||r|| / ||b|| <= 1e-4.  The oracle side takes about a minute on 16 threads."""
import os
import time

import numpy as np
import pytest

from libuipc_amd import abi, synth
from libuipc_amd.device import LinearSystem, rel_err
from oracle import oracle as O

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


@pytest.fixture(scope="module")
def c2():
    t0 = time.perf_counter()
    s = synth.fem_system(64, 2_000_000, seed=1)
    assert s.n_triplets == 32_764_494
    r, c, v, rp = O.convert_sym(s.block_rows, s.rows, s.cols, s.vals)
    print(f"\n[c2] generated + oracle convert in {time.perf_counter() - t0:.1f}s: U={r.size}", flush=True)
    return s, (r, c, v, rp)


@pytest.fixture(scope="module")
def c2_mas(c2):
    s, (r, c, v, _) = c2
    t0 = time.perf_counter()
    m = O.Mas(s.block_rows, s.part_ids, s.tets)
    m.assemble(r, c, v)
    print(f"\n[c2] oracle MAS build {time.perf_counter() - t0:.1f}s", flush=True)
    return m


@pytest.fixture(scope="module")
def c2_gpu(ctx, c2):
    s, _ = c2
    ls = LinearSystem(ctx, s.block_rows)
    rows, cols, vals = ctx.to_device(s.rows), ctx.to_device(s.cols), ctx.to_device(s.vals)
    ls.assemble(rows, cols, vals)
    # the reference's column sweep for the cluster inverses, as the oracle
    # (the default FP64-MFMA panel sweep: test_c2_mas_apply_mfma)
    h = ls.add_preconditioner(abi.MI_PRECOND_MAS, 0, s.n, part_ids=s.part_ids, tets=s.tets,
                              invert=abi.MI_MAS_INVERT_SWEEP)
    ls.assemble_preconditioners()
    ls._keep = (rows, cols, vals)
    yield ls, h
    ls.close()


def test_c2_convert_bitexact(c2, c2_gpu):
    s, (r, c, v, rp) = c2
    ls, _ = c2_gpu
    assert ls.nnz() == r.size == 5_381_741
    gr, gc, gv, grp = ls.export_bcoo()
    assert np.array_equal(gr, r) and np.array_equal(gc, c) and np.array_equal(grp, rp)
    assert np.array_equal(gv.view(np.uint64), v.view(np.uint64))


def test_c2_spmv(ctx, c2, c2_gpu):
    s, (r, c, v, _) = c2
    ls, _ = c2_gpu
    x = np.random.default_rng(21).standard_normal(s.n)
    yref, dref = O.spmv_sym(s.block_rows, r, c, v, x)
    yd = ctx.alloc(s.n, np.float64)
    d = ls.spmv_dot(ctx.to_device(x), yd)
    y = yd.numpy()
    assert np.abs(y - yref).max() <= 1e-13 * np.abs(yref).max()
    assert d == pytest.approx(dref, rel=1e-12)
    # 262,144 rows: 512-row windows by size (512 windows, two per CU)
    lay = ls.spmv_layout()
    assert lay["layout"] == "window" and lay["window_rows"] == 512


def test_c2_mas_tables_and_apply(ctx, c2, c2_mas, c2_gpu):
    s, _ = c2
    ls, h = c2_gpu
    m = c2_mas
    g = ls.mas_export(h, matrices=False)
    t = m.tables()
    assert (g["V"], g["N"], g["L"], g["total"], g["level1_begin"]) == (m.V, m.N, m.L, m.total, m.level1_begin)
    assert g["L"] == 6
    for k in ("part_to_real", "real_to_part", "going_next", "coarse_tables", "level_sizes"):
        assert np.array_equal(g[k], t[k]), k
    rv = np.random.default_rng(22).standard_normal(s.n)
    zd = ctx.alloc(s.n, np.float64)
    ls.apply_preconditioner(ctx.to_device(rv), zd)
    z = zd.numpy()
    err = rel_err(z, m.apply(rv))
    print(f"\n[c2] MAS apply FP64 |z - z_oracle| / |z| = {err:.2e}", flush=True)
    assert err < 1e-9
    assert rv @ z > 0


def test_c2_mas_apply_mfma(ctx, c2, c2_mas):
    """the default cluster inverse (4-row pivot panels on FP64 matrix cores,
    the sweep in its Cholesky form) on the 64^3 + 2M-contact system: the MAS
    apply against the oracle's (whose inverse is the column sweep), to the
    rounding the two inverses differ by"""
    s, _ = c2
    ls = LinearSystem(ctx, s.block_rows)
    rows, cols, vals = ctx.to_device(s.rows), ctx.to_device(s.cols), ctx.to_device(s.vals)
    ls.assemble(rows, cols, vals)
    del rows, cols, vals
    ls.add_preconditioner(abi.MI_PRECOND_MAS, 0, s.n, part_ids=s.part_ids, tets=s.tets)
    ls.assemble_preconditioners()
    rv = np.random.default_rng(24).standard_normal(s.n)
    zd = ctx.alloc(s.n, np.float64)
    ls.apply_preconditioner(ctx.to_device(rv), zd)
    z = zd.numpy()
    err = rel_err(z, c2_mas.apply(rv))
    print(f"\n[c2] MAS apply, MFMA inverse: |z - z_oracle| / |z| = {err:.2e}", flush=True)
    assert err < 1e-7 and rv @ z > 0
    ls.close()


def test_c2_mas_apply_fp32(ctx, c2, c2_mas):
    """the precision arm: FP32 matrix and MAS streams against the oracle's FP32 apply"""
    s, _ = c2
    ls = LinearSystem(ctx, s.block_rows, abi.MI_FP32)
    rows, cols, vals = ctx.to_device(s.rows), ctx.to_device(s.cols), ctx.to_device(s.vals)
    ls.assemble(rows, cols, vals)
    del rows, cols, vals
    ls.add_preconditioner(abi.MI_PRECOND_MAS, 0, s.n, part_ids=s.part_ids, tets=s.tets, mas_precision=abi.MI_FP32)
    ls.assemble_preconditioners()
    rv = np.random.default_rng(23).standard_normal(s.n)
    zd = ctx.alloc(s.n, np.float64)
    ls.apply_preconditioner(ctx.to_device(rv), zd)
    err = rel_err(zd.numpy(), c2_mas.apply(rv, fp32=True))
    print(f"\n[c2] MAS apply FP32 |z - z_oracle| / |z| = {err:.2e}", flush=True)
    assert err < 2e-4
    ls.close()


def test_c2_mas_pcg(ctx, c2, c2_mas, c2_gpu):
    """FP64 MAS-PCG to 1e-4: true residual, iterations within 10% of the
    oracle's PCG on the same system, and the solutions' difference"""
    s, (r, c, v, _) = c2
    ls, h = c2_gpu
    b = ctx.to_device(s.rhs)
    x = ctx.alloc(s.n, np.float64)
    st = ls.solve(b, x, criterion=abi.MI_CRIT_REL_RESIDUAL, tol=1e-4, true_residual=1)
    t0 = time.perf_counter()
    xo, ro = O.pcg(s.block_rows, r, c, v, s.rhs, [dict(kind=3, blk_cnt=s.block_rows, mas=c2_mas)], criterion=1,
                   tol=1e-4, threads=THREADS)
    err = rel_err(x.numpy(), xo)
    print(f"\n[c2] MAS-PCG iters gpu={st['iters_exact']} oracle={ro['iters_exact']} "
          f"(oracle {time.perf_counter() - t0:.1f}s on {THREADS} threads) rel_res_true={st['rel_res_true']:.3e} "
          f"|x - x_oracle| / |x| = {err:.2e}", flush=True)
    assert st["converged"] == 1 and ro["converged"] == 1
    assert st["rel_res_true"] <= 2e-4
    assert abs(st["iters_exact"] - ro["iters_exact"]) <= max(2, 0.1 * ro["iters_exact"])
    # both stop near 1e-4: the solutions agree to about the stopping tolerance's effect
    assert err < 1e-2
