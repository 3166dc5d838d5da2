"""GPU parity of the MAS preconditioner (hierarchy, cluster assembly, inverse,
apply, MAS-PCG) against the oracle, through the C ABI."""
import numpy as np
import pytest

from libuipc_amd import abi, synth
from libuipc_amd.abi import MiError
from libuipc_amd.device import LinearSystem, rel_err
from oracle import oracle as O
from tests.test_oracle_mas import system

pytestmark = pytest.mark.gpu


def gpu_mas(ctx, s, precision=abi.MI_FP64, active_levels=0, invert=abi.MI_MAS_INVERT_SWEEP, shape=None):
    """shape: the SpMV layout (None: by size, the full BSR below ~200k rows;
    0: 512-row windows in the MAS part order, the PCG then iterating in it)"""
    ls = LinearSystem(ctx, s.block_rows)
    if shape is not None:
        ls.set_spmv_shape(shape)
    rows, cols, vals = ctx.to_device(s.rows), ctx.to_device(s.cols), ctx.to_device(s.vals)
    ls.assemble(rows, cols, vals)
    ls._keep = (rows, cols, vals)
    h = ls.add_preconditioner(abi.MI_PRECOND_MAS, 0, s.n, part_ids=s.part_ids, tets=s.tets,
                              mas_precision=precision, active_levels=active_levels, invert=invert)
    ls.assemble_preconditioners()
    return ls, h


@pytest.mark.parametrize("grid,contacts,partition", [(6, 50, "geometric"), (7, 80, "random"), (6, 40, "unpart"),
                                                     (16, 2000, "geometric")])
def test_mas_tables_and_clusters(ctx, grid, contacts, partition):
    s, (r, c, v) = system(grid, contacts, 11, partition)
    ls, h = gpu_mas(ctx, s)
    g = ls.mas_export(h)
    m = O.Mas(s.block_rows, s.part_ids, s.tets)
    m.assemble(r, c, v)
    t = m.tables()
    # integer hierarchy: bit-exact
    assert (g["V"], g["N"], g["L"], g["total"], g["level1_begin"]) == (m.V, m.N, m.L, m.total, m.level1_begin)
    for k in ("part_to_real", "real_to_part", "going_next", "coarse_tables"):
        assert np.array_equal(g[k], t[k]), k
    assert np.array_equal(g["level_sizes"], t["level_sizes"])
    # cluster Hessians: same sums, different order of the coarse additions
    H = m.hessians()
    assert np.allclose(g["hess"], H, rtol=1e-11, atol=1e-12 * np.abs(H).max())
    # FP64 inverses (same sweep, FMA rounding)
    I = m.inverses()
    assert np.abs(g["inv"] - I).max() <= 1e-8 * np.abs(I).max()


@pytest.mark.parametrize("precision,tol", [(abi.MI_FP64, 1e-9), (abi.MI_FP32, 2e-4)])
@pytest.mark.parametrize("partition", ["geometric", "random", "unpart"])
def test_mas_apply(ctx, precision, tol, partition):
    s, (r, c, v) = system(8, 300, 12, partition)
    ls, h = gpu_mas(ctx, s, precision)
    m = O.Mas(s.block_rows, s.part_ids, s.tets)
    m.assemble(r, c, v)
    rv = np.random.default_rng(3).standard_normal(s.n)
    zd = ctx.alloc(s.n, np.float64)
    ls.apply_preconditioner(ctx.to_device(rv), zd)
    z = zd.numpy()
    want = m.apply(rv)
    assert rel_err(z, want) < tol
    assert rv @ z > 0


def dense_clusters(blocks):
    """(C, 136, 9) upper blocks (column-major 3x3) -> (C, 48, 48) symmetric"""
    C = len(blocks)
    out = np.zeros((C, 48, 48))
    k = 0
    for a in range(16):
        for b in range(a, 16):
            B = blocks[:, k].reshape(C, 3, 3).transpose(0, 2, 1)
            out[:, 3 * a:3 * a + 3, 3 * b:3 * b + 3] = B
            out[:, 3 * b:3 * b + 3, 3 * a:3 * a + 3] = B.transpose(0, 2, 1)
            k += 1
    return out


@pytest.mark.parametrize("precision,tol", [(abi.MI_FP64, 1e-10), (abi.MI_FP32, 1e-6)])
@pytest.mark.parametrize("partition", ["geometric", "random"])
def test_mas_inverse_kernels(ctx, partition, precision, tol):
    """The MFMA block Gauss-Jordan and the column sweep give the same inverse to
    rounding, H C^-1 = I on every cluster (a padding slot's diagonal counts as
    1), and both reach about the same MAS-PCG iteration count."""
    s, _ = system(12, 3000, 15, partition)
    out = {}
    for inv in (abi.MI_MAS_INVERT_MFMA, abi.MI_MAS_INVERT_SWEEP):
        ls, h = gpu_mas(ctx, s, precision, invert=inv)
        g = ls.mas_export(h)
        x = ctx.alloc(s.n, np.float64)
        st = ls.solve(ctx.to_device(s.rhs), x, criterion=abi.MI_CRIT_REL_RESIDUAL, tol=1e-4)
        out[inv] = (g, st)
    (gm, sm), (gs, ss) = out[abi.MI_MAS_INVERT_MFMA], out[abi.MI_MAS_INVERT_SWEEP]
    # the cluster sums use FP64 atomics: equal to rounding between two builds
    assert np.allclose(gm["hess"], gs["hess"], rtol=1e-12, atol=1e-13 * np.abs(gs["hess"]).max())
    assert np.abs(gm["inv"] - gs["inv"]).max() <= max(1e-8, tol) * np.abs(gs["inv"]).max()
    H, Ci = dense_clusters(gm["hess"]), dense_clusters(gm["inv"])
    d = np.arange(48)
    pad = H[:, d, d] == 0
    H[:, d, d] = np.where(pad, 1.0, H[:, d, d])
    E = np.einsum("kij,kjl->kil", H, Ci) - np.eye(48)
    scale = np.abs(H).max(axis=(1, 2)) * np.abs(Ci).max(axis=(1, 2))
    assert (np.abs(E).max(axis=(1, 2)) <= tol * scale).all()
    print(f"\nMAS-PCG to 1e-4: {sm['iters_exact']} iterations (MFMA inverse), {ss['iters_exact']} (sweep)", flush=True)
    assert abs(sm["iters_exact"] - ss["iters_exact"]) <= max(2, ss["iters_exact"] // 20)


def test_mas_one_level_block_jacobi(ctx):
    s, (r, c, v) = system(8, 300, 13, "random")
    ls, h = gpu_mas(ctx, s, active_levels=1)
    m = O.Mas(s.block_rows, s.part_ids, s.tets)
    m.assemble(r, c, v)
    rv = np.random.default_rng(4).standard_normal(s.n)
    zd = ctx.alloc(s.n, np.float64)
    ls.apply_preconditioner(ctx.to_device(rv), zd)
    assert rel_err(zd.numpy(), m.apply(rv, active_levels=1)) < 1e-9


@pytest.mark.parametrize("levels", [0, 1])
@pytest.mark.parametrize("precision", [abi.MI_FP64, abi.MI_FP32])
@pytest.mark.parametrize("shape", [None, 0], ids=["auto", "window"])
def test_mas_pcg(ctx, precision, levels, shape):
    """MAS-PCG against the oracle's, all levels and one level; FP32 runs the
    fine apply fused with the r.z partials; on the size's layout (the full
    BSR here) and on the window layout, whose PCG iterates in the part order"""
    s, (r, c, v) = system(16, 20000, 14)
    ls, h = gpu_mas(ctx, s, precision, active_levels=levels, shape=shape)
    x = ctx.alloc(s.n, np.float64)
    st = ls.solve(ctx.to_device(s.rhs), x, criterion=abi.MI_CRIT_REL_RESIDUAL, tol=1e-4, true_residual=1)
    m = O.Mas(s.block_rows, s.part_ids, s.tets)
    m.assemble(r, c, v)
    xo, ro = O.pcg(s.block_rows, r, c, v, s.rhs, [dict(kind=3, blk_cnt=s.block_rows, mas=m,
                                                        mas_fp32=precision == abi.MI_FP32, active_levels=levels)],
                   criterion=1, tol=1e-4)
    assert st["converged"] == 1 and ro["converged"] == 1
    assert st["rel_res"] <= 1e-4 and st["rel_res_true"] < 2e-4
    assert abs(st["iters_exact"] - ro["iters_exact"]) <= max(2, ro["iters_exact"] // 10)
    # vs a tight FP64 solve
    xt, rt = O.pcg(s.block_rows, r, c, v, s.rhs, [dict(kind=3, blk_cnt=s.block_rows, mas=m)], criterion=1, tol=1e-11)
    assert rel_err(x.numpy(), xt) < 1e-2


@pytest.mark.parametrize("grid,partition", [(48, "geometric"), (48, "random"), (12, "unpart")])
def test_mas_pcg_iterate_after_a_few_iterations(ctx, grid, partition):
    """MAS-PCG stopped after 12 iterations gives the oracle's iterate to FP64
    rounding: every step of the iteration -- the x, r update with the
    level-0 (and, on the 48^3 grid's geometric parts, level-1) restriction
    folded in, the coarse and fine solves, the scalars -- is the oracle's.
    48^3 has a base level above 1 (6,912 level-1 nodes), so the level-1
    restriction runs; random parts break the alignment it needs and fall back
    to the separate kernel; un-partitioned vertices take the 3x3 fallback."""
    s, (r, c, v) = system(grid, 20 * grid ** 2, 31, partition)
    ls, h = gpu_mas(ctx, s, shape=0)
    ratio = 12.5 / s.n  # max_iter = 12
    x = ctx.alloc(s.n, np.float64)
    st = ls.solve(ctx.to_device(s.rhs), x, criterion=abi.MI_CRIT_REL_RESIDUAL, tol=1e-14, max_iter_ratio=ratio)
    m = O.Mas(s.block_rows, s.part_ids, s.tets)
    m.assemble(r, c, v)
    xo, ro = O.pcg(s.block_rows, r, c, v, s.rhs, [dict(kind=3, blk_cnt=s.block_rows, mas=m)], criterion=1, tol=1e-14,
                   max_iter_ratio=ratio, threads=8)
    assert ro["converged"] == 0 and ro["iters"] == 12 and st["converged"] == 0 and st["iters"] == 12
    assert rel_err(x.numpy(), xo) < 1e-9
    assert st["rel_res"] == pytest.approx(ro["rel_res"], rel=1e-8)
    ls.close()


@pytest.mark.parametrize("partition", ["random", "unpart"])
def test_mas_pcg_fp32_one_level(ctx, partition):
    """FP32, one level (the fine kernel fused with the r.z partials), with
    disconnected parts and with un-partitioned vertices (the 3x3 fallback):
    the solve matches the oracle's and the returned x solves the system"""
    s, (r, c, v) = system(12, 3000, 16, partition)
    ls, h = gpu_mas(ctx, s, abi.MI_FP32, active_levels=1, shape=0)
    x = ctx.alloc(s.n, np.float64)
    st = ls.solve(ctx.to_device(s.rhs), x, criterion=abi.MI_CRIT_REL_RESIDUAL, tol=1e-4, true_residual=1)
    m = O.Mas(s.block_rows, s.part_ids, s.tets)
    m.assemble(r, c, v)
    xo, ro = O.pcg(s.block_rows, r, c, v, s.rhs, [dict(kind=3, blk_cnt=s.block_rows, mas=m, mas_fp32=True,
                                                        active_levels=1)], criterion=1, tol=1e-4)
    assert st["converged"] == 1 and st["rel_res_true"] < 2e-4
    assert abs(st["iters_exact"] - ro["iters_exact"]) <= max(2, ro["iters_exact"] // 10)


def test_mas_set_active_levels_and_tuner(ctx):
    """mi_precond_mas_set_active_levels changes the apply to the oracle's with
    that many levels (0 = all, as the reference clamps), also for a solve
    captured before the change; the tuner times every level count and leaves
    the fastest set"""
    from libuipc_amd.tune import tune_mas_levels

    s, (r, c, v) = system(10, 500, 17, "geometric")
    ls, h = gpu_mas(ctx, s)
    m = O.Mas(s.block_rows, s.part_ids, s.tets)
    m.assemble(r, c, v)
    L = ls.mas_export(h, matrices=False)["L"]
    rv = np.random.default_rng(6).standard_normal(s.n)
    zd = ctx.alloc(s.n, np.float64)
    b, x = ctx.to_device(s.rhs), ctx.alloc(s.n, np.float64)
    iters = {}
    for lv in (1, 2, 0, L + 3):
        ls.mas_set_active_levels(h, lv)
        ls.apply_preconditioner(ctx.to_device(rv), zd)
        want = m.apply(rv, active_levels=lv if 0 < lv <= L else 0)
        assert rel_err(zd.numpy(), want) < 1e-9, lv
        iters[lv] = ls.solve(b, x, criterion=abi.MI_CRIT_REL_RESIDUAL, tol=1e-4)["iters_exact"]
    # both are all levels; over ~900 iterations the rounding of the SpMV's
    # LDS atomics moves the count by a few percent (the applies above agree)
    assert abs(iters[0] - iters[L + 3]) <= max(2, iters[0] // 20)
    with pytest.raises(MiError):
        d = LinearSystem(ctx, s.block_rows)
        hd = d.add_preconditioner(abi.MI_PRECOND_DIAG3, 0, s.n)
        d.mas_set_active_levels(hd, 1)
    t = tune_mas_levels(ls, h, b, x, criterion=abi.MI_CRIT_REL_RESIDUAL, tol=1e-4)
    assert sorted(t["table"]) == list(range(1, L + 1))
    assert t["best"] == min(t["table"], key=lambda k: t["table"][k]["ms"])
    ls.apply_preconditioner(ctx.to_device(rv), zd)
    assert rel_err(zd.numpy(), m.apply(rv, active_levels=t["best"])) < 1e-9


def test_mas_full_size_apply(ctx):
    """64^3 grid + 200k contacts: apply vs the oracle, and r^T z > 0"""
    s = synth.fem_system(64, 200_000, seed=3)
    ls, h = gpu_mas(ctx, s)
    r, c, v, _ = O.convert_sym(s.block_rows, s.rows, s.cols, s.vals)
    m = O.Mas(s.block_rows, s.part_ids, s.tets)
    m.assemble(r, c, v)
    g = ls.mas_export(h, matrices=False)
    t = m.tables()
    assert g["L"] == 6 and np.array_equal(g["going_next"], t["going_next"])
    rv = np.random.default_rng(5).standard_normal(s.n)
    zd = ctx.alloc(s.n, np.float64)
    ls.apply_preconditioner(ctx.to_device(rv), zd)
    z = zd.numpy()
    assert rel_err(z, m.apply(rv)) < 1e-8
    assert rv @ z > 0


def test_mas_dump_files(ctx, tmp_path):
    """The debug dump (reference file names, Matrix Market block layout, JSON
    metadata) carries exactly the exported state."""
    import json

    s, _ = system(8, 300, 12, "geometric")
    ls, h = gpu_mas(ctx, s)
    g = ls.mas_export(h)
    ls.mas_dump(h, str(tmp_path), frame=3, newton_iter=7)
    meta = json.loads((tmp_path / "mas_cluster_meta.f3.n7.json").read_text())
    C = g["clusters"]
    assert meta["total_nodes"] == g["V"] and meta["total_map_nodes"] == g["N"]
    assert meta["total_clusters"] == g["total"] and meta["num_cluster_blocks"] == C
    assert meta["banksize"] == 16 and meta["sym_block_count"] == 136 and meta["level_num"] == g["L"]
    assert meta["part_to_real"] == g["part_to_real"].tolist()
    assert meta["real_to_part"] == g["real_to_part"].tolist()
    assert meta["going_next"] == g["going_next"][:g["total"]].tolist()
    assert [(e["count"], e["offset"]) for e in meta["levels"]] == [tuple(x) for x in g["level_sizes"].tolist()]
    for kind in ("hess", "inv"):
        lines = (tmp_path / f"mas_cluster_{kind}.f3.n7.mtx").read_text().splitlines()
        assert lines[0] == "%%MatrixMarket matrix coordinate real general"
        n, m, e = (int(t) for t in lines[2].split())
        assert n == m == 48 * C and e == C * 136 * 9
        ent = np.array([ln.split() for ln in lines[3:]], dtype=np.float64)
        assert len(ent) == e
        i = ent[:, 0].astype(np.int64) - 1
        j = ent[:, 1].astype(np.int64) - 1
        want = g[kind].reshape(C, 136, 9)
        # entry order: cluster, upper block (row-major over the bank), row, column
        cl = i // 48
        br, bc = (i % 48) // 3, (j % 48) // 3
        k = 16 * br - br * (br + 1) // 2 + bc
        assert np.all(bc >= br) and np.array_equal(cl, j // 48)
        assert np.array_equal(ent[:, 2], want[cl, k, (i % 3) + 3 * (j % 3)])


@pytest.mark.parametrize("precision", [abi.MI_FP64, abi.MI_FP32])
def test_mas_subsystem_at_an_offset(ctx, precision):
    """A MAS preconditioner on a FEM subsystem that starts at block row 100 of
    the global system, behind a block-Jacobi range, with coupling blocks
    between the two ranges (dropped by both preconditioners): its apply and a
    MAS-PCG solve match the oracle run on the subsystem alone."""
    s, (r, c, v) = system(8, 200, 21, "geometric")
    off = 100
    rng = np.random.default_rng(5)
    B = s.block_rows + off
    # front range: SPD diagonal blocks; coupling: weak blocks (front row, FEM col)
    fr = np.arange(off, dtype=np.int32)
    fv = np.zeros((off, 9))
    fv[:, [0, 4, 8]] = 1e3 + rng.random((off, 3))
    cr = rng.integers(0, off, 200).astype(np.int32)
    cc = (off + rng.integers(0, s.block_rows, 200)).astype(np.int32)
    cv = 1e-3 * rng.standard_normal((200, 9))
    rows = np.concatenate([fr, s.rows + off, cr]).astype(np.int32)
    cols = np.concatenate([fr, s.cols + off, cc]).astype(np.int32)
    vals = np.concatenate([fv, s.vals, cv])
    ls = LinearSystem(ctx, B)
    dr, dc, dv = ctx.to_device(rows), ctx.to_device(cols), ctx.to_device(vals)
    ls.assemble(dr, dc, dv)
    ls.add_preconditioner(abi.MI_PRECOND_DIAG3, 0, 3 * off)
    ls.add_preconditioner(abi.MI_PRECOND_MAS, 3 * off, s.n, part_ids=s.part_ids, tets=s.tets,
                          mas_precision=precision)
    ls.assemble_preconditioners()
    m = O.Mas(s.block_rows, s.part_ids, s.tets)
    m.assemble(r, c, v)
    rv = rng.standard_normal(3 * B)
    zd = ctx.alloc(3 * B, np.float64)
    ls.apply_preconditioner(ctx.to_device(rv), zd)
    z = zd.numpy()
    tol = 1e-9 if precision == abi.MI_FP64 else 2e-4
    assert rel_err(z[3 * off:], m.apply(rv[3 * off:])) < tol
    front = rv[:3 * off].reshape(-1, 3) / fv[:, [0, 4, 8]]
    assert np.allclose(z[:3 * off], front.reshape(-1), rtol=1e-12, atol=0)
    x = ctx.alloc(3 * B, np.float64)
    b = rng.uniform(-1, 1, 3 * B)
    st = ls.solve(ctx.to_device(b), x, criterion=abi.MI_CRIT_REL_RESIDUAL, tol=1e-6, true_residual=1)
    assert st["converged"] == 1 and st["rel_res_true"] < 2e-6
    ls.close()
