"""mi_spmv_dot_noalloc -- the SpMV with x^T A x whose partials live in a fixed
device array -- against mi_spmv_dot and the oracle, through the C ABI: two
calls in a row, the second on a smaller system, FP64 and FP32 value planes,
accumulation into a nonzero y, a NULL dot, and the FP32 size limit."""

import numpy as np
import pytest

from libuipc_amd import abi, synth
from libuipc_amd.device import LinearSystem
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _system(ctx, s, precision):
    ls = LinearSystem(ctx, s.block_rows, precision)
    keep = (ctx.to_device(s.rows), ctx.to_device(s.cols), ctx.to_device(s.vals))
    ls.assemble(*keep)
    ls._keep = keep
    return ls


def _call(ls, fn, xd, y0):
    yd = ls.ctx.to_device(y0)
    d = ls.ctx.alloc(1, np.float64)
    abi.check(getattr(ls.lib, fn)(ls.h, xd.ptr, yd.ptr, d.ptr, None))
    ls.ctx.synchronize()
    return yd.numpy(), float(d.numpy()[0])


@pytest.mark.parametrize("precision", [abi.MI_FP64, abi.MI_FP32])
def test_noalloc_matches_spmv_dot(ctx, precision):
    tol = 1e-13 if precision == abi.MI_FP64 else 1e-6
    # the second call's system is the smaller one
    for k, (grid, contacts, seed) in enumerate([(16, 20000, 7), (8, 2000, 3)]):
        s = synth.fem_system(grid, contacts, seed=seed)
        ls = _system(ctx, s, precision)
        rng = np.random.default_rng(10 + k)
        x = rng.standard_normal(s.n)
        xd = ctx.to_device(x)
        zero = np.zeros(s.n)
        y_ref, d_ref = _call(ls, "mi_spmv_dot", xd, zero)
        y_na, d_na = _call(ls, "mi_spmv_dot_noalloc", xd, zero)
        scale = np.abs(y_ref).max()
        assert np.abs(y_na - y_ref).max() <= 1e-13 * scale, (grid, precision)
        assert d_na == pytest.approx(d_ref, rel=1e-12)
        # against the oracle's U + U^T - D product
        r, c, v, _ = O.convert_sym(s.block_rows, s.rows, s.cols, s.vals)
        yo, do = O.spmv_sym(s.block_rows, r, c, v, x)
        assert np.abs(y_na - yo).max() <= tol * np.abs(yo).max()
        assert d_na == pytest.approx(do, rel=10 * tol)
        # y is accumulated into, as in mi_spmv_dot
        y0 = rng.standard_normal(s.n)
        y_acc, _ = _call(ls, "mi_spmv_dot_noalloc", xd, y0)
        y_acc_ref, _ = _call(ls, "mi_spmv_dot", xd, y0)
        assert np.abs(y_acc - y_acc_ref).max() <= 1e-13 * np.abs(y_acc_ref).max()
        # a NULL dot is allowed
        yd = ctx.to_device(zero)
        abi.check(ls.lib.mi_spmv_dot_noalloc(ls.h, xd.ptr, yd.ptr, None, None))
        assert np.abs(yd.numpy() - y_ref).max() <= 1e-13 * scale
        ls.close()


def test_noalloc_fp32_size_limit(ctx):
    """above 2^18 block rows: on the window layout (the default) both
    precisions run and give A x and x^T A x; the full-BSR layout's FP32 kernel
    needs a float4 copy of x that would not fit the fixed array, so there an
    FP32 system is refused with MI_ERR_ARG"""
    B = (1 << 18) + 1
    idx = np.arange(B, dtype=np.int32)
    vals = np.tile(np.eye(3).reshape(1, 9), (B, 1))
    x = np.arange(3 * B, dtype=np.float64) % 7.0
    for precision in (abi.MI_FP32, abi.MI_FP64):
        ls = LinearSystem(ctx, B, precision)
        keep = (ctx.to_device(idx), ctx.to_device(idx), ctx.to_device(vals))
        ls.assemble(*keep)
        xd = ctx.to_device(x)
        yd = ctx.to_device(np.zeros(3 * B))
        d = ctx.alloc(1, np.float64)
        rc = ls.lib.mi_spmv_dot_noalloc(ls.h, xd.ptr, yd.ptr, d.ptr, None)
        if precision == abi.MI_FP32 and ls.spmv_layout()["layout"] == "full":
            assert rc == abi.MI_ERR_ARG
            assert b"262144" in ls.lib.mi_last_error()
        else:
            assert rc == abi.MI_OK
            assert np.array_equal(yd.numpy(), x) and float(d.numpy()[0]) == pytest.approx(float(x @ x))
        ls.close()
