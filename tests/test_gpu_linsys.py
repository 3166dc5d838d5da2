"""GPU parity of the assembled matrix, the SpMV, the vector reductions, the
block-Jacobi preconditioners and the PCG driver against the CPU oracle, all
called through the C ABI (libmi_linsys.so)."""
import json
import math
import os

import numpy as np
import pytest

from libuipc_amd import abi, synth
from libuipc_amd.device import LinearSystem, rel_err
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))


def upload_system(ctx, s):
    return ctx.to_device(s.rows), ctx.to_device(s.cols), ctx.to_device(s.vals)


def assembled(ctx, s, precision=abi.MI_FP64):
    ls = LinearSystem(ctx, s.block_rows, precision)
    r, c, v = upload_system(ctx, s)
    ls.assemble(r, c, v)
    ls._trip = (r, c, v)
    return ls


# ------------------------------------------------------------------ KATs
def test_dot_norm_kats(ctx):
    for case in GOLD["dot_norm_f32"]:
        x = ctx.to_device(np.array(case["x"], np.float32))
        if "y" in case:
            y = ctx.to_device(np.array(case["y"], np.float32))
            assert ctx.dot(x, y) == pytest.approx(case["dot"], rel=1e-6)
        assert ctx.norm(x) == pytest.approx(case["norm_x"], rel=case.get("norm_eps", 1e-6))
    g = GOLD["norm_f64_scaled"]
    v = ctx.norm(ctx.to_device(np.array(g["x"])))
    assert math.isfinite(v) and v == pytest.approx(g["norm"], rel=g["eps"])
    e = ctx.to_device(np.zeros(0))
    e.size = 0
    assert ctx.dot(e, e) == 0.0 and ctx.norm(e) == 0.0
    with pytest.raises(ValueError):
        ctx.dot(ctx.to_device(np.ones(1)), ctx.to_device(np.ones(2)))
    assert math.isinf(ctx.norm(ctx.to_device(np.array([np.inf, 1.0]))))
    assert math.isnan(ctx.norm(ctx.to_device(np.array([np.nan, 1.0]))))


def test_dot_large(ctx):
    rng = np.random.default_rng(3)
    x = rng.standard_normal(1_000_003)
    y = rng.standard_normal(1_000_003)
    assert ctx.dot(ctx.to_device(x), ctx.to_device(y)) == pytest.approx(O.dot(x, y), rel=1e-12)
    assert ctx.norm(ctx.to_device(x)) == pytest.approx(O.norm(x), rel=1e-12)


# ------------------------------------------------------------------ convert
def check_bcoo_bitexact(ls, s):
    r, c, v, rp = ls.export_bcoo()
    orr, oc, ov, orp = O.convert_sym(s.block_rows, s.rows, s.cols, s.vals)
    assert r.size == orr.size
    assert np.array_equal(r, orr) and np.array_equal(c, oc) and np.array_equal(rp, orp)
    # duplicates are summed in the same (source) order: bit-identical values
    assert np.array_equal(v.view(np.uint64), ov.view(np.uint64))


def test_convert_c1_bitexact(ctx, c1):
    ls = assembled(ctx, c1)
    assert ls.nnz() == 26416
    check_bcoo_bitexact(ls, c1)


def test_convert_contacts_bitexact(ctx, c1_contact):
    check_bcoo_bitexact(assembled(ctx, c1_contact), c1_contact)


def test_convert_f32_triplets(ctx, c1_contact):
    """mi_sys_assemble_f32: FP32 triplet values (the precision arm's
    producers) give, bit for bit, the blocks of the oracle's conversion of the
    same values widened to double (sums in FP64, ascending source order), and
    the FP32 system's product is the oracle's on those blocks; the same
    triplets as doubles give the same system; an FP64 system refuses them"""
    s = c1_contact
    v32 = s.vals.astype(np.float32)
    r, c = ctx.to_device(s.rows), ctx.to_device(s.cols)
    ls = LinearSystem(ctx, s.block_rows, abi.MI_FP32)
    ls.assemble(r, c, ctx.to_device(v32))
    gr, gc, gv, gp = ls.export_bcoo()
    orr, oc, ov, orp = O.convert_sym(s.block_rows, s.rows, s.cols, v32.astype(np.float64))
    assert np.array_equal(gr, orr) and np.array_equal(gc, oc) and np.array_equal(gp, orp)
    assert np.array_equal(gv.view(np.uint64), ov.view(np.uint64))
    x = np.random.default_rng(12).standard_normal(s.n)
    yref, dref = O.spmv_sym(s.block_rows, orr, oc, ov, x)
    yd = ctx.alloc(s.n, np.float64)
    d = ls.spmv_dot(ctx.to_device(x), yd)
    assert np.abs(yd.numpy() - yref).max() <= 1e-6 * np.abs(yref).max()
    assert d == pytest.approx(dref, rel=1e-5)
    ls2 = LinearSystem(ctx, s.block_rows, abi.MI_FP32)
    ls2.assemble(r, c, ctx.to_device(v32.astype(np.float64)))
    assert np.array_equal(ls2.export_bcoo()[2].view(np.uint64), gv.view(np.uint64))
    ls64 = LinearSystem(ctx, s.block_rows, abi.MI_FP64)
    with pytest.raises(abi.MiError):
        ls64.assemble(r, c, ctx.to_device(v32))
    for h in (ls, ls2, ls64):
        h.close()


@pytest.mark.parametrize("B,T,seed", [(1, 5, 0), (7, 300, 1), (300, 20000, 2), (5000, 70000, 3)])
def test_convert_random_bitexact(ctx, B, T, seed):
    """lower-triangle and invalid (-1) entries are dropped; heavy duplication"""
    rng = np.random.default_rng(seed)
    rows = rng.integers(0, B, T).astype(np.int32)
    cols = rng.integers(0, B, T).astype(np.int32)
    rows[rng.random(T) < 0.05] = -1
    vals = rng.standard_normal((T, 9))

    class S:
        pass

    s = S()
    s.block_rows, s.rows, s.cols, s.vals = B, rows, cols, vals
    check_bcoo_bitexact(assembled(ctx, s), s)


def test_convert_empty_and_reassemble(ctx, c1):
    ls = LinearSystem(ctx, 10)
    z = ctx.alloc(1, np.int32)
    zv = ctx.alloc((1, 9), np.float64)
    z.size = 0
    ls.assemble(z, z, zv)
    assert ls.nnz() == 0
    r, c, v, rp = ls.export_bcoo()
    assert rp.tolist() == [0] * 11
    # a second assembly of a different system on the same handle
    ls2 = assembled(ctx, c1)
    check_bcoo_bitexact(ls2, c1)
    r2, c2, v2 = upload_system(ctx, c1)
    ls2.assemble(r2, c2, v2)
    check_bcoo_bitexact(ls2, c1)


def test_convert_full_size_fem_bitexact(ctx):
    """64^3 FEM grid (the FEM part of the 64^3 + 2M contact system): U is the closed form and
    the structure and values are bit-exact with the oracle"""
    s = synth.fem_system(64, 0, seed=1)
    ls = assembled(ctx, s)
    assert ls.nnz() == 1_798_336
    check_bcoo_bitexact(ls, s)


# ------------------------------------------------------------------ SpMV
@pytest.mark.parametrize("precision", [abi.MI_FP64, abi.MI_FP32])
def test_spmv_dot(ctx, c1_contact, precision):
    s = c1_contact
    ls = assembled(ctx, s, precision)
    orr, oc, ov, _ = O.convert_sym(s.block_rows, s.rows, s.cols, s.vals)
    x = np.random.default_rng(4).standard_normal(s.n)
    yref, dref = O.spmv_sym(s.block_rows, orr, oc, ov, x)
    xd = ctx.to_device(x)
    yd = ctx.alloc(s.n, np.float64)
    d = ls.spmv_dot(xd, yd)
    y = yd.numpy()
    tol = 1e-13 if precision == abi.MI_FP64 else 1e-6
    assert np.abs(y - yref).max() <= tol * np.abs(yref).max()
    assert d == pytest.approx(dref, rel=tol * 10)
    # y = a A x + b y
    w = np.random.default_rng(5).standard_normal(s.n)
    wd = ctx.to_device(w)
    if precision == abi.MI_FP64:
        ls.spmv(-2.0, xd, 0.5, wd)
        assert np.abs(wd.numpy() - (-2.0 * yref + 0.5 * w)).max() <= 1e-12 * np.abs(yref).max()


@pytest.mark.parametrize("precision", [abi.MI_FP64, abi.MI_FP32])
@pytest.mark.parametrize("unpart", [False, True])
def test_spmv_in_part_order(ctx, c1_contact, precision, unpart):
    """with a MAS preconditioner attached, the SpMV's windows follow its part
    order (rows permuted; a block inside one window stored once and applied
    both ways there, a block across windows stored with both rows): the same
    product as the oracle's, when the MAS is added after the assembly (the
    layout is rebuilt by the preconditioner assembly) and on a reassembly;
    also with every seventh vertex unpartitioned (ranked after the parts)"""
    s = c1_contact
    parts = s.part_ids.copy()
    if unpart:
        parts[::7] = -1
    ls = assembled(ctx, s, precision)
    ls.set_spmv_shape(0)  # 4,096 rows take the full BSR by size
    ls.add_preconditioner(abi.MI_PRECOND_MAS, 0, s.n, part_ids=parts, tets=s.tets)
    ls.assemble_preconditioners()
    orr, oc, ov, _ = O.convert_sym(s.block_rows, s.rows, s.cols, s.vals)
    rng = np.random.default_rng(21)
    tol = 1e-13 if precision == abi.MI_FP64 else 1e-6
    for reassemble in (False, True):
        if reassemble:
            ls.assemble(*ls._trip)
            ls.assemble_preconditioners()
        x = rng.standard_normal(s.n)
        yref, dref = O.spmv_sym(s.block_rows, orr, oc, ov, x)
        yd = ctx.alloc(s.n, np.float64)
        d = ls.spmv_dot(ctx.to_device(x), yd)
        assert np.abs(yd.numpy() - yref).max() <= tol * np.abs(yref).max(), reassemble
        assert d == pytest.approx(dref, rel=tol * 10)
    ls.close()


@pytest.mark.parametrize("precision", [abi.MI_FP64, abi.MI_FP32])
def test_spmv_every_launch_shape(ctx, c1_contact, precision):
    """every SpMV launch shape (its window size, workgroup and stage) gives
    the oracle's product, with the windows in a MAS part order; the shape
    tuner times them all and leaves the fastest set"""
    from libuipc_amd.device import spmv_shapes
    from libuipc_amd.tune import tune_spmv_shape

    s = c1_contact
    ls = assembled(ctx, s, precision)
    ls.add_preconditioner(abi.MI_PRECOND_MAS, 0, s.n, part_ids=s.part_ids, tets=s.tets)
    ls.assemble_preconditioners()
    assert ls.spmv_layout()["layout"] == "full"  # 4,096 rows: the full BSR by size
    orr, oc, ov, _ = O.convert_sym(s.block_rows, s.rows, s.cols, s.vals)
    x = np.random.default_rng(8).standard_normal(s.n)
    yref, dref = O.spmv_sym(s.block_rows, orr, oc, ov, x)
    xd = ctx.to_device(x)
    tol = 1e-13 if precision == abi.MI_FP64 else 1e-6
    info = spmv_shapes()
    assert 0 <= info["default"] < len(info["shapes"])
    for k, sh in enumerate(info["shapes"]):
        ls.set_spmv_shape(k)
        lay = ls.spmv_layout()
        assert lay["layout"] == "window" and lay["window_rows"] == sh["rows"] and lay["threads"] == sh["threads"]
        yd = ctx.alloc(s.n, np.float64)
        d = ls.spmv_dot(xd, yd)
        assert np.abs(yd.numpy() - yref).max() <= tol * np.abs(yref).max(), k
        assert d == pytest.approx(dref, rel=tol * 10), k
        # the PCG's product, on vectors in the row order
        pos = ls.spmv_order()
        xo = np.empty((s.block_rows, 3))
        xo[pos] = x.reshape(-1, 3)
        ls.spmv_ordered(ctx.to_device(xo.ravel()), yd)
        yo = yd.numpy().reshape(-1, 3)[pos].ravel()
        assert np.abs(yo - yref).max() <= tol * np.abs(yref).max(), k
    t = tune_spmv_shape(ls, xd, reps=3)
    # an FP64 system runs a four-block-load shape as its two-block twin: not timed
    want = [k for k, sh in enumerate(info["shapes"]) if not (precision == abi.MI_FP64 and sh["vec"] == 4)]
    assert sorted(t["table"]) == [abi.MI_SPMV_SHAPE_FULL] + want
    assert ls.spmv_layout()["window_rows"] == t["table"][t["best"]]["rows"]
    with pytest.raises(abi.MiError):
        ls.set_spmv_shape(len(info["shapes"]))
    ls.close()


def test_pcg_after_a_shape_change(ctx, c1_contact):
    """a solve captured in a HIP graph with one SpMV layout is not replayed
    after the layout changes (another window size, the full BSR, the size's
    default): the next solve re-captures and gives the same result, after a
    reassembly and without one"""
    from libuipc_amd.device import spmv_shapes

    s = c1_contact
    ls = assembled(ctx, s)
    assert ls.spmv_layout()["layout"] == "full"  # 4,096 rows: the full BSR by size
    ls.set_spmv_shape(0)
    ls.add_preconditioner(abi.MI_PRECOND_MAS, 0, s.n, part_ids=s.part_ids, tets=s.tets)
    ls.assemble_preconditioners()
    shapes = spmv_shapes()["shapes"]
    k0 = ls.spmv_layout()["shape"]
    assert k0 == 0
    k1 = next(k for k, sh in enumerate(shapes) if sh["rows"] != shapes[k0]["rows"])
    b = ctx.to_device(s.rhs)
    kw = dict(criterion=abi.MI_CRIT_REL_RESIDUAL, tol=1e-6, true_residual=1, use_graph=1)
    x0 = ctx.alloc(s.n, np.float64)
    st0 = ls.solve(b, x0, **kw)
    for k, reassemble in ((k1, True), (abi.MI_SPMV_SHAPE_FULL, False), (k0, True), (abi.MI_SPMV_SHAPE_AUTO, False),
                          (k1, False)):
        ls.set_spmv_shape(k)
        want = "window" if k >= 0 else "full"
        assert ls.spmv_layout()["layout"] == want, k
        if reassemble:
            ls.assemble(*ls._trip)
            ls.assemble_preconditioners()
        assert ls.spmv_layout()["layout"] == want, k
        x = ctx.alloc(s.n, np.float64)
        st = ls.solve(b, x, **kw)
        assert st["converged"] == 1 and st["rel_res_true"] <= 2e-6, (k, st)
        assert abs(st["iters_exact"] - st0["iters_exact"]) <= max(2, st0["iters_exact"] // 50), (k, st, st0)
        assert rel_err(x.numpy(), x0.numpy()) < 1e-4
    ls.close()


@pytest.mark.parametrize("precision", [abi.MI_FP64, abi.MI_FP32])
@pytest.mark.parametrize("unpart", [False, True])
def test_spmv_ordered(ctx, c1_contact, precision, unpart):
    """mi_spmv_ordered, the PCG's own product on vectors in the SpMV's row
    order, is the oracle's product permuted by mi_sys_spmv_order: the order
    is a permutation, other than the identity with a MAS attached, that puts
    the unpartitioned vertices after the partitioned ones; x^T A x is the
    oracle's"""
    s = c1_contact
    parts = s.part_ids.copy()
    if unpart:
        parts[::7] = -1
    ls = assembled(ctx, s, precision)
    ls.set_spmv_shape(0)  # 4,096 rows take the full BSR by size
    assert np.array_equal(ls.spmv_order(), np.arange(s.block_rows))  # no MAS: the identity
    ls.add_preconditioner(abi.MI_PRECOND_MAS, 0, s.n, part_ids=parts, tets=s.tets)
    ls.assemble_preconditioners()
    pos = ls.spmv_order()
    assert np.array_equal(np.sort(pos), np.arange(s.block_rows))
    assert not np.array_equal(pos, np.arange(s.block_rows))
    if unpart:
        assert pos[parts < 0].min() > pos[parts >= 0].max()
    orr, oc, ov, _ = O.convert_sym(s.block_rows, s.rows, s.cols, s.vals)
    x = np.random.default_rng(31).standard_normal(s.n)
    yref, dref = O.spmv_sym(s.block_rows, orr, oc, ov, x)
    xo = np.empty((s.block_rows, 3))
    xo[pos] = x.reshape(-1, 3)
    yd = ctx.alloc(s.n, np.float64)
    dd = ctx.alloc(1, np.float64)
    ls.spmv_ordered(ctx.to_device(xo.ravel()), yd, dd)
    y = yd.numpy().reshape(-1, 3)[pos].ravel()
    tol = 1e-13 if precision == abi.MI_FP64 else 1e-6
    assert np.abs(y - yref).max() <= tol * np.abs(yref).max()
    assert dd.numpy()[0] == pytest.approx(dref, rel=tol * 10)
    ls.close()


@pytest.mark.parametrize("precision", [abi.MI_FP64, abi.MI_FP32])
def test_pcg_row_order_and_caller_order(ctx, c1_contact, precision):
    """the PCG on vectors in the SpMV's row order (the default with a MAS
    attached) and on the caller's vertex order (keep_order = 1) run the same
    iteration up to rounding: iterations within 2%, the same solution to the
    tolerance, in both directions of a switch between captured graphs, and
    from a given x0"""
    s = c1_contact
    ls = assembled(ctx, s, precision)
    ls.set_spmv_shape(0)  # 4,096 rows take the full BSR by size
    ls.add_preconditioner(abi.MI_PRECOND_MAS, 0, s.n, part_ids=s.part_ids, tets=s.tets)
    ls.assemble_preconditioners()
    b = ctx.to_device(s.rhs)
    # FP32 value planes: the true residual levels off near 1e-5 (the float
    # products), so the FP32 solve stops there
    tol, true_tol, x_tol = (1e-8, 2e-8, 1e-6) if precision == abi.MI_FP64 else (1e-5, 1e-4, 1e-3)
    kw = dict(criterion=abi.MI_CRIT_REL_RESIDUAL, tol=tol, true_residual=1, use_graph=1)
    res = []
    for keep in (0, 1, 0, 1):
        x = ctx.alloc(s.n, np.float64)
        st = ls.solve(b, x, keep_order=keep, **kw)
        assert st["converged"] == 1 and st["rel_res_true"] <= true_tol, (keep, st)
        res.append((st, x.numpy()))
    it0, x0 = res[0][0]["iters_exact"], res[0][1]
    for st, x in res[1:]:
        assert abs(st["iters_exact"] - it0) <= max(2, it0 // 50)
        assert rel_err(x, x0) < x_tol
    # from x0 = a perturbed solution, in either order
    rng = np.random.default_rng(4)
    xg = x0 + 1e-3 * np.abs(x0).max() * rng.standard_normal(s.n)
    for keep in (0, 1):
        x = ctx.to_device(xg)
        st = ls.solve(b, x, x_init=1, keep_order=keep, **kw)
        assert st["converged"] == 1 and st["rel_res_true"] <= true_tol, (keep, st)
        assert st["iters_exact"] < it0 and rel_err(x.numpy(), x0) < x_tol
    ls.close()


@pytest.mark.parametrize("precision", [abi.MI_FP64, abi.MI_FP32])
def test_spmv_irregular_rows(ctx, precision):
    """Rows the SpMV's 64-block windows must get right: runs of one-block rows
    (64 row starts in one window), empty rows (a zero block is stored for
    them), a row of 1,900 blocks that carries across many windows, and short
    random rows; checked against the oracle's expansion of U + U^T - D."""
    rng = np.random.default_rng(11)
    B = 3000
    rows, cols = [np.arange(1000)], [np.arange(1000)]          # one-block rows 0..999
    rows.append(np.full(1900, 1100)); cols.append(np.arange(1100, 3000))  # long row 1100
    r = rng.integers(1101, 3000, 9000)
    c = np.minimum(B - 1, r + rng.integers(0, 40, 9000))
    rows.append(r); cols.append(c)                               # rows 1000..1099 stay empty
    rows = np.concatenate(rows).astype(np.int32)
    cols = np.concatenate(cols).astype(np.int32)
    vals = rng.standard_normal((rows.size, 9))
    ls = LinearSystem(ctx, B, precision)
    rd, cd, vd = ctx.to_device(rows), ctx.to_device(cols), ctx.to_device(vals)
    ls.assemble(rd, cd, vd)
    orr, oc, ov, _ = O.convert_sym(B, rows, cols, vals)
    x = rng.standard_normal(3 * B)
    yref, dref = O.spmv_sym(B, orr, oc, ov, x)
    xd = ctx.to_device(x)
    yd = ctx.alloc(3 * B, np.float64)
    d = ls.spmv_dot(xd, yd)
    y = yd.numpy()
    tol = 1e-13 if precision == abi.MI_FP64 else 1e-6
    assert np.abs(y - yref).max() <= tol * np.abs(yref).max()
    assert not y[3000:3300].any()                   # the empty rows
    assert d == pytest.approx(dref, rel=tol * 10)
    if precision == abi.MI_FP64:
        w = rng.standard_normal(3 * B)
        wd = ctx.to_device(w)
        ls.spmv(1.5, xd, -1.0, wd)
        assert np.abs(wd.numpy() - (1.5 * yref - w)).max() <= 1e-12 * np.abs(yref).max()
    ls.close()


def test_spmv_symmetry_full_size(ctx):
    """size-independent property at 64^3 + contacts: x^T (A w) == w^T (A x)"""
    s = synth.fem_system(64, 200_000, seed=2)
    ls = assembled(ctx, s)
    rng = np.random.default_rng(6)
    x, w = rng.standard_normal(s.n), rng.standard_normal(s.n)
    xd, wd = ctx.to_device(x), ctx.to_device(w)
    yx, yw = ctx.alloc(s.n, np.float64), ctx.alloc(s.n, np.float64)
    dx = ls.spmv_dot(xd, yx)
    ls.spmv_dot(wd, yw)
    a, b = x @ yw.numpy(), w @ yx.numpy()
    assert a == pytest.approx(b, rel=1e-11)
    assert dx == pytest.approx(x @ yx.numpy(), rel=1e-11)


# ------------------------------------------------------------------ preconditioners
def test_diag3_apply(ctx, c1_contact):
    s = c1_contact
    ls = assembled(ctx, s)
    ls.add_preconditioner(abi.MI_PRECOND_DIAG3, 0, s.n)
    ls.assemble_preconditioners()
    orr, oc, ov, _ = O.convert_sym(s.block_rows, s.rows, s.cols, s.vals)
    D = O.diag3(orr, oc, ov, 0, s.block_rows)
    r = np.random.default_rng(7).standard_normal(s.n)
    zd = ctx.alloc(s.n, np.float64)
    ls.apply_preconditioner(ctx.to_device(r), zd)
    zref = O.diag3_apply(D, r)
    # 3x3 inverses of contact-stiffened blocks: rounding x block condition number
    assert rel_err(zd.numpy(), zref) < 1e-10


def test_diag12_and_identity_ranges(ctx):
    """ABD-like bodies in [0, 84) with a 12x12 block Jacobi, an uncovered tail -> z = r"""
    from tests.test_oracle import abd_like_system

    rng = np.random.default_rng(5)
    rows, cols, vals = abd_like_system(rng, 7)
    B = 28 + 5
    rows = np.concatenate([rows, np.arange(28, B, dtype=np.int32)])
    cols = np.concatenate([cols, np.arange(28, B, dtype=np.int32)])
    extra = np.zeros((5, 9))
    extra[:, [0, 4, 8]] = 2.0
    vals = np.concatenate([vals, extra])

    class S:
        pass

    s = S()
    s.block_rows, s.rows, s.cols, s.vals = B, rows, cols, vals
    ls = assembled(ctx, s)
    ls.add_preconditioner(abi.MI_PRECOND_DIAG12, 0, 84)
    ls.assemble_preconditioners()
    orr, oc, ov, _ = O.convert_sym(B, rows, cols, vals)
    H = O.diag12(orr, oc, ov, 0, 7)
    r = rng.standard_normal(3 * B)
    zd = ctx.alloc(3 * B, np.float64)
    ls.apply_preconditioner(ctx.to_device(r), zd)
    z = zd.numpy()
    assert rel_err(z[:84], O.diag12_apply(H, r[:84])) < 1e-12
    assert np.array_equal(z[84:], r[84:])


# ------------------------------------------------------------------ PCG
def oracle_pcg(s, criterion, tol, precond=True, **kw):
    orr, oc, ov, _ = O.convert_sym(s.block_rows, s.rows, s.cols, s.vals)
    pc = []
    if precond:
        pc = [dict(kind=1, blk_off=0, blk_cnt=s.block_rows, data=O.diag3(orr, oc, ov, 0, s.block_rows))]
    return O.pcg(s.block_rows, orr, oc, ov, s.rhs, pc, criterion=criterion, tol=tol, **kw)


@pytest.mark.parametrize("criterion,tol", [(abi.MI_CRIT_RZ_RATIO, 1e-3), (abi.MI_CRIT_REL_RESIDUAL, 1e-4),
                                           (abi.MI_CRIT_REL_RESIDUAL, 1e-10)])
@pytest.mark.parametrize("use_graph", [0, 1])
def test_pcg_diag3(ctx, c1_contact, criterion, tol, use_graph):
    s = c1_contact
    ls = assembled(ctx, s)
    ls.add_preconditioner(abi.MI_PRECOND_DIAG3, 0, s.n)
    ls.assemble_preconditioners()
    b = ctx.to_device(s.rhs)
    x = ctx.alloc(s.n, np.float64)
    st = ls.solve(b, x, criterion=criterion, tol=tol, use_graph=use_graph, true_residual=1)
    xo, ro = oracle_pcg(s, criterion, tol)
    assert st["converged"] == 1 and ro["converged"] == 1
    assert abs(st["iters_exact"] - ro["iters_exact"]) <= max(1, ro["iters_exact"] // 50)
    assert st["iters"] % 5 == 0
    if criterion == abi.MI_CRIT_REL_RESIDUAL:
        assert st["rel_res"] <= tol
        assert st["rel_res_true"] <= 2 * tol
    # same algorithm in FP64: same iterate up to the round-off of the atomics
    assert rel_err(x.numpy(), xo) < max(1e-6, 100 * tol)


@pytest.mark.parametrize("precond", ["diag3", "mas"])
def test_pcg_initial_guess(ctx, c1_contact, precond):
    """x_init = 1: x0 = 0 given explicitly reproduces the default solve (to
    the rounding of the SpMV's LDS atomics: iterations within 2%, the same
    solution to the tolerance); x0 = the solution converges at once; x0 = the
    solution plus a perturbation needs fewer iterations than x0 = 0 and still
    meets ||b - A x|| <= tol ||b||"""
    s = c1_contact
    ls = assembled(ctx, s)
    if precond == "diag3":
        ls.add_preconditioner(abi.MI_PRECOND_DIAG3, 0, s.n)
    else:
        ls.add_preconditioner(abi.MI_PRECOND_MAS, 0, s.n, part_ids=s.part_ids, tets=s.tets)
    ls.assemble_preconditioners()
    b = ctx.to_device(s.rhs)
    kw = dict(criterion=abi.MI_CRIT_REL_RESIDUAL, tol=1e-6, true_residual=1)
    x = ctx.alloc(s.n, np.float64)
    cold = ls.solve(b, x, **kw)
    x_cold = x.numpy()
    x.zero()
    st = ls.solve(b, x, x_init=1, **kw)
    assert abs(st["iters_exact"] - cold["iters_exact"]) <= max(2, cold["iters_exact"] // 50)
    assert st["rel_res_true"] <= 2e-6 and rel_err(x.numpy(), x_cold) < 1e-4
    # the solution of a tighter solve as x0
    xt = ctx.alloc(s.n, np.float64)
    ls.solve(b, xt, criterion=abi.MI_CRIT_REL_RESIDUAL, tol=1e-10)
    x_star = xt.numpy()
    x.upload(x_star)
    st = ls.solve(b, x, x_init=1, **kw)
    assert st["converged"] == 1 and st["iters_exact"] <= 1 and st["rel_res_true"] <= 1e-6
    rng = np.random.default_rng(9)
    x.upload(x_star + 1e-3 * np.abs(x_star).max() * rng.standard_normal(s.n))
    st = ls.solve(b, x, x_init=1, **kw)
    assert st["converged"] == 1 and st["rel_res_true"] <= 2e-6
    assert st["iters_exact"] < cold["iters_exact"]
    assert rel_err(x.numpy(), x_star) < 1e-4


def test_pcg_bypass_is_plain_cg(ctx, c1):
    ls = assembled(ctx, c1)
    ls.add_preconditioner(abi.MI_PRECOND_DIAG3, 0, c1.n)
    ls.assemble_preconditioners()
    b = ctx.to_device(c1.rhs)
    x = ctx.alloc(c1.n, np.float64)
    st = ls.solve(b, x, criterion=abi.MI_CRIT_REL_RESIDUAL, tol=1e-6, force_bypass=1)
    xo, ro = oracle_pcg(c1, abi.MI_CRIT_REL_RESIDUAL, 1e-6, precond=False)
    assert st["bypassed"] == 1
    assert abs(st["iters_exact"] - ro["iters_exact"]) <= 1
    assert rel_err(x.numpy(), xo) < 1e-5


class _Sys:
    pass


def with_decoupled_vertex(s, block, rhs3):
    """s plus one vertex coupled to nothing: its diagonal block and rhs given"""
    t = _Sys()
    B = s.block_rows
    t.block_rows, t.n = B + 1, 3 * (B + 1)
    t.rows = np.concatenate([s.rows, [B]]).astype(np.int32)
    t.cols = np.concatenate([s.cols, [B]]).astype(np.int32)
    t.vals = np.concatenate([np.asarray(s.vals).reshape(-1, 9), np.asarray(block, np.float64).reshape(1, 9)])
    t.rhs = np.concatenate([s.rhs, rhs3])
    return t


# the PSD guard's natural triggers (linear_fused_pcg.cu:1843-1875): a singular
# diagonal block makes the 3x3 block Jacobi non-finite on a finite r (the
# block's null direction has a zero rhs, so plain CG still converges); a
# negative diagonal block with a large rhs makes r0.z0 < 0
PSD_CASES = {
    "singular": (np.diag([1.0, 1.0, 0.0]).reshape(-1, order="F"), [0.5, -0.25, 0.0]),
    "negative": ((-np.eye(3)).reshape(-1), [30.0, 30.0, 30.0]),
}


@pytest.mark.parametrize("case", sorted(PSD_CASES))
def test_pcg_psd_guard_natural(ctx, c1, case):
    """the guard bypasses the preconditioner by itself: bypassed = 1 and the
    same plain-CG iterations and solution as the oracle's FusedPCG"""
    t = with_decoupled_vertex(c1, *PSD_CASES[case])
    ls = assembled(ctx, t)
    ls.add_preconditioner(abi.MI_PRECOND_DIAG3, 0, t.n)
    ls.assemble_preconditioners()
    x = ctx.alloc(t.n, np.float64)
    st = ls.solve(ctx.to_device(t.rhs), x, criterion=abi.MI_CRIT_REL_RESIDUAL, tol=1e-6, true_residual=1)
    xo, ro = oracle_pcg(t, abi.MI_CRIT_REL_RESIDUAL, 1e-6)
    assert ro["bypassed"] == 1 and st["bypassed"] == 1
    assert st["converged"] == 1 and ro["converged"] == 1
    assert abs(st["iters_exact"] - ro["iters_exact"]) <= 1
    assert st["rz0"] == pytest.approx(ro["rz0"], rel=1e-12)
    assert rel_err(x.numpy(), xo) < 1e-6
    # without the fallback a non-finite r0.z0 is an error (check_init_rz_nan_inf)
    if case == "singular":
        with pytest.raises(abi.MiError) as e:
            ls.solve(ctx.to_device(t.rhs), x, psd_fallback=0)
        assert e.value.code == abi.MI_ERR_NONFINITE
    ls.close()


def test_pcg_nonfinite_rhs_is_an_error(ctx, c1):
    """a NaN in b makes r non-finite: no bypass, MI_ERR_NONFINITE (the
    oracle's FusedPCG stops without iterating)"""
    rhs = c1.rhs.copy()
    rhs[5] = np.nan
    ls = assembled(ctx, c1)
    ls.add_preconditioner(abi.MI_PRECOND_DIAG3, 0, c1.n)
    ls.assemble_preconditioners()
    x = ctx.alloc(c1.n, np.float64)
    with pytest.raises(abi.MiError) as e:
        ls.solve(ctx.to_device(rhs), x)
    assert e.value.code == abi.MI_ERR_NONFINITE
    t = _Sys()
    t.block_rows, t.rows, t.cols, t.vals, t.rhs = c1.block_rows, c1.rows, c1.cols, c1.vals, rhs
    _, ro = oracle_pcg(t, abi.MI_CRIT_RZ_RATIO, 1e-3)
    assert ro["bypassed"] == 0 and ro["converged"] == 0 and ro["iters"] == 0
    # a NaN in a given x0 makes r0 = b - A x0 non-finite: the same error, not a bypass
    x0 = np.zeros(c1.n)
    x0[7] = np.nan
    x.upload(x0)
    with pytest.raises(abi.MiError) as e:
        ls.solve(ctx.to_device(c1.rhs), x, x_init=1)
    assert e.value.code == abi.MI_ERR_NONFINITE
    ls.close()


@pytest.mark.parametrize("use_graph", [0, 1])
def test_pcg_iteration_cap(ctx, c1_contact, use_graph):
    """a tiny max_iter_ratio: max_iter - 1 iterations, converged = 0 and
    iters = max_iter (linear_fused_pcg.cu:1994), the oracle's iterate"""
    s = c1_contact
    ls = assembled(ctx, s)
    ls.add_preconditioner(abi.MI_PRECOND_DIAG3, 0, s.n)
    ls.assemble_preconditioners()
    ratio = 12.5 / s.n  # max_iter = 12
    x = ctx.alloc(s.n, np.float64)
    st = ls.solve(ctx.to_device(s.rhs), x, criterion=abi.MI_CRIT_REL_RESIDUAL, tol=1e-12, max_iter_ratio=ratio,
                  use_graph=use_graph)
    xo, ro = oracle_pcg(s, abi.MI_CRIT_REL_RESIDUAL, 1e-12, max_iter_ratio=ratio)
    assert ro["converged"] == 0 and ro["iters"] == 12
    assert st["converged"] == 0 and st["iters"] == 12 and st["iters_exact"] == -1
    assert st["rel_res"] == pytest.approx(ro["rel_res"], rel=1e-8)
    assert rel_err(x.numpy(), xo) < 1e-10
    ls.close()


def test_pcg_fp32_matrix(ctx, c1_contact):
    """precision arm: FP32 matrix stream, FP64 vectors and dots"""
    s = c1_contact
    ls = assembled(ctx, s, abi.MI_FP32)
    ls.add_preconditioner(abi.MI_PRECOND_DIAG3, 0, s.n)
    ls.assemble_preconditioners()
    x = ctx.alloc(s.n, np.float64)
    st = ls.solve(ctx.to_device(s.rhs), x, criterion=abi.MI_CRIT_REL_RESIDUAL, tol=1e-4, true_residual=1)
    assert st["converged"] == 1 and st["rel_res_true"] < 1e-3
    xo, ro = oracle_pcg(s, abi.MI_CRIT_REL_RESIDUAL, 1e-8)
    assert rel_err(x.numpy(), xo) < 1e-2


def test_pcg_zero_rhs(ctx, c1):
    ls = assembled(ctx, c1)
    x = ctx.alloc(c1.n, np.float64)
    st = ls.solve(ctx.to_device(np.zeros(c1.n)), x)
    assert st["iters"] == 0 and st["converged"] == 1
    assert not np.any(x.numpy())


def test_errors_are_reported(ctx, c1):
    ls = LinearSystem(ctx, c1.block_rows)
    x = ctx.alloc(c1.n, np.float64)
    with pytest.raises(abi.MiError) as e:
        ls.solve(ctx.to_device(c1.rhs), x)
    assert e.value.code == abi.MI_ERR_STATE
    with pytest.raises(abi.MiError):
        ls.add_preconditioner(abi.MI_PRECOND_DIAG3, 3, c1.n)  # outside the system
