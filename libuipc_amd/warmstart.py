"""Learned initial guess for MAS-PCG (a message-passing network over the
block graph of the assembled system; no counterpart in the reference, whose
solve starts from x0 = 0).

Per block row i the inputs are b_i / s_b, the 3x3 Jacobi guess (D_i^-1 b_i)
scaled by its RMS, the row's neighbour count and weight sum, and 1.  The
network (width 64, BF16 matrix products with FP32 accumulation,
include/mi_warmstart.h) is

    H_0 = relu(X W_in + b_in)
    H_l = relu(H_{l-1} W_s^l + (G H_{l-1}) W_n^l + b^l),    l = 1..3
    Y   = H_3 W_out + b_out
    x0  = D^-1 b + s_b D^-1 Y

with G the block graph weighted by |A_ij|_F / sqrt(|A_ii|_F |A_jj|_F), and D
the 3x3 diagonal blocks: with Y = 0 the guess is the block-Jacobi one.  Given
the MAS hierarchy's first coarse level (each block row's level-1 node), a
coarse-level correction is added:

    Hc  = Pool H_3                  (mean over each coarse node's rows)
    Yc  = Hc W_c + b_c
    x0 += s_j P Yc                  (P: injection to the node's rows)

with s_j the RMS of D^-1 b; W_c starts at zero, so untrained it adds
nothing.  The
training loss is the relative residual ||b - A x0||^2 / ||b||^2 (gradient
-2 A r / ||b||^2) or the energy (x0^T A x0 / 2 - b^T x0) / (b^T D^-1 b)
(gradient -r / (b^T D^-1 b)), through the library's SpMV, then the network's
backward pass on the device; Adam on FP32 master weights.  At solve time the
guess is used only when its residual is below ||b|| (else x0 = 0, the plain
method)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from .abi import check
from .device import Context, DeviceArray, LinearSystem

F_IN, WIDTH, LAYERS, F_OUT = 8, 64, 3, 3


def _layout():
    """offsets of the parameters in one flat FP32 buffer"""
    off, out = 0, {}
    for name, shape in [("W_in", (F_IN, WIDTH)), ("b_in", (WIDTH,))] + \
            [(f"{p}{l}", s) for l in range(1, LAYERS + 1)
             for p, s in (("W_s", (WIDTH, WIDTH)), ("W_n", (WIDTH, WIDTH)), ("b", (WIDTH,)))] + \
            [("W_out", (WIDTH, F_OUT)), ("b_out", (F_OUT,)), ("W_c", (WIDTH, F_OUT)), ("b_c", (F_OUT,))]:
        out[name] = (off, shape)
        off += int(np.prod(shape))
    return out, off


LAYOUT, N_PARAMS = _layout()


@dataclass
class Graph:
    """One assembled system's block graph, inputs and scales on the device."""
    n: int
    ptr: DeviceArray
    idx: DeviceArray
    w: DeviceArray
    dinv: DeviceArray  # float64[n, 9], column-major 3x3 inverses of the diagonal blocks
    x_jacobi: DeviceArray  # float64[3n]
    X: DeviceArray  # float32[n, F_IN]
    b: DeviceArray  # float64[3n]
    s_b: float
    bb: float
    bxj: float  # b . D^-1 b, the scale of the energy loss
    # the coarse level (nc = 0: none): per coarse node its rows (cptr, cidx;
    # weights 1 / size and 1), per row its node (pidx; weights 1 and 1 / size)
    nc: int = 0
    s_c: float = 1.0
    cptr: DeviceArray | None = None
    cidx: DeviceArray | None = None
    c_mean: DeviceArray | None = None
    c_one: DeviceArray | None = None
    pptr: DeviceArray | None = None
    pidx: DeviceArray | None = None
    p_one: DeviceArray | None = None
    p_inv: DeviceArray | None = None
    eye: DeviceArray | None = None  # float64[n, 9] identity blocks


def coarse_of_mas(ls: LinearSystem, handle: int) -> np.ndarray:
    """each block row's level-1 node of a MAS preconditioner over the whole
    system (-1: unpartitioned)"""
    return ls.mas_export(handle, matrices=False)["coarse_tables"][:, 0].copy()


def build_graph(ctx: Context, ls: LinearSystem, rhs: np.ndarray, coarse: np.ndarray | None = None) -> Graph:
    """The graph and input features of an assembled system (host construction
    from the exported upper BCOO: a few ms for 4,096 block rows).  `coarse`:
    each block row's coarse node (any ids; -1: a node of its own), for the
    coarse-level correction."""
    rows, cols, vals, _ = ls.export_bcoo()
    n = ls.B
    blocks = vals.reshape(-1, 3, 3).transpose(0, 2, 1)  # column-major -> [r][c]
    fro = np.linalg.norm(vals, axis=1)
    diag = rows == cols
    dnorm = np.ones(n)
    dnorm[rows[diag]] = fro[diag]
    D = np.tile(np.eye(3), (n, 1, 1))
    D[rows[diag]] = blocks[diag]
    Dinv = np.linalg.inv(D)
    off = ~diag
    r, c = rows[off], cols[off]
    wv = (fro[off] / np.sqrt(dnorm[r] * dnorm[c])).astype(np.float32)
    src = np.concatenate([r, c])
    dst = np.concatenate([c, r])
    ww = np.concatenate([wv, wv])
    order = np.lexsort((dst, src))
    src, dst, ww = src[order], dst[order], ww[order]
    ptr = np.zeros(n + 1, np.int32)
    np.add.at(ptr, src + 1, 1)
    ptr = np.cumsum(ptr).astype(np.int32)
    b3 = rhs.reshape(n, 3)
    xj = np.einsum("nij,nj->ni", Dinv, b3)
    s_b = float(np.sqrt(np.mean(rhs * rhs))) or 1.0
    s_j = float(np.sqrt(np.mean(xj * xj))) or 1.0
    deg = np.diff(ptr).astype(np.float64)
    wsum = np.zeros(n)
    np.add.at(wsum, src, ww)
    X = np.zeros((n, F_IN), np.float32)
    X[:, 0:3] = b3 / s_b
    X[:, 3:6] = xj / s_j
    X[:, 6] = deg / 16.0
    X[:, 7] = wsum / max(1.0, float(wsum.mean()))
    dinv_cm = Dinv.transpose(0, 2, 1).reshape(n, 9)  # [r][c] -> column-major
    g = Graph(n, ctx.to_device(ptr), ctx.to_device(dst.astype(np.int32)), ctx.to_device(ww),
              ctx.to_device(np.ascontiguousarray(dinv_cm)), ctx.to_device(xj.reshape(-1)), ctx.to_device(X),
              ctx.to_device(rhs), s_b, float(rhs @ rhs), float(b3.reshape(-1) @ xj.reshape(-1)))
    if coarse is not None:
        c = np.asarray(coarse, np.int64).copy()
        if c.shape != (n,):
            raise ValueError("build_graph: one coarse node per block row")
        own = c < 0
        c[own] = c.max(initial=-1) + 1 + np.arange(int(own.sum()))
        _, node = np.unique(c, return_inverse=True)
        nc = int(node.max()) + 1
        size = np.bincount(node, minlength=nc).astype(np.float64)
        order = np.argsort(node, kind="stable")
        cptr = np.zeros(nc + 1, np.int32)
        cptr[1:] = np.cumsum(size).astype(np.int32)
        g.nc, g.s_c = nc, s_j
        g.cptr, g.cidx = ctx.to_device(cptr), ctx.to_device(order.astype(np.int32))
        g.c_mean = ctx.to_device((1.0 / size[node[order]]).astype(np.float32))
        g.c_one = ctx.to_device(np.ones(n, np.float32))
        g.pptr, g.pidx = ctx.to_device(np.arange(n + 1, dtype=np.int32)), ctx.to_device(node.astype(np.int32))
        g.p_one = ctx.to_device(np.ones(n, np.float32))
        g.p_inv = ctx.to_device((1.0 / size[node]).astype(np.float32))
        g.eye = ctx.to_device(np.tile(np.eye(3).reshape(1, 9), (n, 1)))
    return g


class WarmStartNet:
    def __init__(self, ctx: Context, seed: int = 0):
        self.ctx, self.lib, self.st = ctx, ctx.lib, ctx.stream()
        rng = np.random.default_rng(seed)
        p = np.zeros(N_PARAMS, np.float32)
        for name, (o, shape) in LAYOUT.items():
            if name.startswith("W"):
                fan_in = shape[0]
                scale = 0.0 if name in ("W_out", "W_c") else np.sqrt(2.0 / fan_in)
                p[o:o + int(np.prod(shape))] = (scale * rng.standard_normal(int(np.prod(shape)))).astype(np.float32)
        self.params = ctx.to_device(p)
        self.grads = ctx.alloc(N_PARAMS, np.float32)
        self.m = ctx.to_device(np.zeros(N_PARAMS, np.float32))
        self.v = ctx.to_device(np.zeros(N_PARAMS, np.float32))
        self.step = 0
        self._n = -1

    # ---- parameters
    def p(self, name: str, grad: bool = False) -> int:
        o, _ = LAYOUT[name]
        base = (self.grads if grad else self.params).ptr.value
        return base + 4 * o

    def numpy(self) -> np.ndarray:
        return self.params.numpy()

    def load(self, p: np.ndarray) -> None:
        self.params.upload(np.ascontiguousarray(p, np.float32))

    # ---- scratch
    def _scratch(self, n: int) -> None:
        if n == self._n:
            return
        a = lambda *s: self.ctx.alloc(s, np.float32)  # noqa: E731
        self.H = [a(n, WIDTH) for _ in range(LAYERS + 1)]
        self.M = [a(n, WIDTH) for _ in range(LAYERS + 1)]
        self.Y = a(n, F_OUT)
        self.dH, self.dH2, self.G = a(n, WIDTH), a(n, WIDTH), a(n, WIDTH)
        self.dY = a(n, F_OUT)
        self.Yp, self.dYp = a(n, F_OUT), a(n, F_OUT)
        self.x0 = self.ctx.alloc(3 * n, np.float64)
        self.r = self.ctx.alloc(3 * n, np.float64)
        self.gx = self.ctx.alloc(3 * n, np.float64)
        self._n = n
        self._nc = -1

    def _coarse_scratch(self, nc: int) -> None:
        if nc == getattr(self, "_nc", -1):
            return
        a = lambda *s: self.ctx.alloc(s, np.float32)  # noqa: E731
        self.Hc, self.dHc = a(max(nc, 1), WIDTH), a(max(nc, 1), WIDTH)
        self.Yc, self.dYc = a(max(nc, 1), F_OUT), a(max(nc, 1), F_OUT)
        self._nc = nc

    def _agg(self, n, F, ptr, idx, w, H, Z) -> None:
        check(self.lib.mi_ws_aggregate(self.ctx.h, n, F, ptr.ptr, idx.ptr, w.ptr, H.ptr, Z.ptr, self.st))

    def _gemm(self, M, N, K, A, sa_m, sa_k, B, sb_k, sb_n, Cp, ldc, acc=0, bias=None, relu=0):
        check(self.lib.mi_ws_gemm(self.ctx.h, M, N, K, A, sa_m, sa_k, B, sb_k, sb_n, Cp, ldc, acc, bias, relu,
                                  self.st))

    def _ptr(self, d: DeviceArray) -> int:
        return d.ptr.value

    # ---- forward: x0 of the graph's system (device, float64[3n])
    def forward(self, g: Graph) -> DeviceArray:
        self._scratch(g.n)
        n, P = g.n, self._ptr
        self._gemm(n, WIDTH, F_IN, P(g.X), F_IN, 1, self.p("W_in"), WIDTH, 1, P(self.H[0]), WIDTH,
                   bias=self.p("b_in"), relu=1)
        for l in range(1, LAYERS + 1):
            check(self.lib.mi_ws_aggregate(self.ctx.h, n, WIDTH, g.ptr.ptr, g.idx.ptr, g.w.ptr, self.H[l - 1].ptr,
                                           self.M[l].ptr, self.st))
            self._gemm(n, WIDTH, WIDTH, P(self.H[l - 1]), WIDTH, 1, self.p(f"W_s{l}"), WIDTH, 1, P(self.H[l]), WIDTH)
            self._gemm(n, WIDTH, WIDTH, P(self.M[l]), WIDTH, 1, self.p(f"W_n{l}"), WIDTH, 1, P(self.H[l]), WIDTH,
                       acc=1, bias=self.p(f"b{l}"), relu=1)
        self._gemm(n, F_OUT, WIDTH, P(self.H[LAYERS]), WIDTH, 1, self.p("W_out"), F_OUT, 1, P(self.Y), F_OUT,
                   bias=self.p("b_out"))
        check(self.lib.mi_ws_block3(self.ctx.h, n, g.dinv.ptr, g.s_b, self.Y.ptr, self.x0.ptr, g.x_jacobi.ptr, None,
                                    None, self.st))
        if g.nc:
            # the coarse-level correction: pool, a linear map, inject
            self._coarse_scratch(g.nc)
            self._agg(g.nc, WIDTH, g.cptr, g.cidx, g.c_mean, self.H[LAYERS], self.Hc)
            self._gemm(g.nc, F_OUT, WIDTH, P(self.Hc), WIDTH, 1, self.p("W_c"), F_OUT, 1, P(self.Yc), F_OUT,
                       bias=self.p("b_c"))
            self._agg(n, F_OUT, g.pptr, g.pidx, g.p_one, self.Yc, self.Yp)
            check(self.lib.mi_ws_block3(self.ctx.h, n, g.eye.ptr, g.s_c, self.Yp.ptr, self.x0.ptr, self.x0.ptr,
                                        None, None, self.st))
        return self.x0

    # ---- the loss and the parameter gradient
    def loss_and_grad(self, ls: LinearSystem, g: Graph, loss: str = "residual") -> float:
        """loss "residual": ||b - A x0||^2 / ||b||^2, gradient -2 A r / ||b||^2
        (two SpMVs); loss "energy": (x0^T A x0 / 2 - b^T x0) / (b^T D^-1 b),
        which differs from ||x0 - A^-1 b||_A^2 / 2 (the error CG's iteration
        count depends on) by a constant, gradient -r / (b^T D^-1 b) (one
        SpMV)."""
        n, P = g.n, self._ptr
        x0 = self.forward(g)
        self.ctx.lib.mi_memcpy_d2d(self.ctx.h, self.r.ptr, g.b.ptr, 8 * 3 * n)
        ls.spmv(-1.0, x0, 1.0, self.r)
        # dY = s D^-1 gx (D^-1 symmetric), s = s_b times the gradient's scale
        if loss == "residual":
            value = self.ctx.dot(self.r, self.r) / g.bb
            ls.spmv(-2.0 / g.bb, self.r, 0.0, self.gx)
            gvec, s = self.gx, g.s_b
        elif loss == "energy":
            # x0^T A x0 = x0^T (b - r)  =>  E = -x0^T (b + r) / 2
            value = -0.5 * (self.ctx.dot(x0, g.b) + self.ctx.dot(x0, self.r)) / g.bxj
            gvec, s = self.r, -g.s_b / g.bxj
        else:
            raise ValueError(f"unknown loss {loss!r}")
        check(self.lib.mi_ws_block3(self.ctx.h, n, g.dinv.ptr, s, None, None, None, gvec.ptr, self.dY.ptr,
                                    self.st))
        H3 = self.H[LAYERS]
        self._gemm(WIDTH, F_OUT, n, P(H3), 1, WIDTH, P(self.dY), F_OUT, 1, self.p("W_out", True), F_OUT)
        check(self.lib.mi_ws_colsum(self.ctx.h, n, F_OUT, self.dY.ptr, self.p("b_out", True), 0, self.st))
        if g.nc:
            # d x0 / d Yp = s_j, so dYp = s_j (s / s_b) gvec; dYc = P^T dYp;
            # dH_3 gets Pool^T (dYc W_c^T)
            check(self.lib.mi_ws_block3(self.ctx.h, n, g.eye.ptr, s * g.s_c / g.s_b, None, None, None, gvec.ptr,
                                        self.dYp.ptr, self.st))
            self._agg(g.nc, F_OUT, g.cptr, g.cidx, g.c_one, self.dYp, self.dYc)
            self._gemm(WIDTH, F_OUT, g.nc, P(self.Hc), 1, WIDTH, P(self.dYc), F_OUT, 1, self.p("W_c", True), F_OUT)
            check(self.lib.mi_ws_colsum(self.ctx.h, g.nc, F_OUT, self.dYc.ptr, self.p("b_c", True), 0, self.st))
            self._gemm(g.nc, WIDTH, F_OUT, P(self.dYc), F_OUT, 1, self.p("W_c"), 1, F_OUT, P(self.dHc), WIDTH)
            self._agg(n, WIDTH, g.pptr, g.pidx, g.p_inv, self.dHc, self.dH)
        else:
            for name in ("W_c", "b_c"):
                o, sh = LAYOUT[name]
                check(self.lib.mi_memset(self.ctx.h, self.p(name, True), 0, 4 * int(np.prod(sh))))
        self._gemm(n, WIDTH, F_OUT, P(self.dY), F_OUT, 1, self.p("W_out"), 1, F_OUT, P(self.dH), WIDTH,
                   acc=1 if g.nc else 0)
        for l in range(LAYERS, 0, -1):
            check(self.lib.mi_ws_relu_backward(self.ctx.h, n * WIDTH, self.H[l].ptr, self.dH.ptr, self.st))
            self._gemm(WIDTH, WIDTH, n, P(self.H[l - 1]), 1, WIDTH, P(self.dH), WIDTH, 1, self.p(f"W_s{l}", True),
                       WIDTH)
            self._gemm(WIDTH, WIDTH, n, P(self.M[l]), 1, WIDTH, P(self.dH), WIDTH, 1, self.p(f"W_n{l}", True), WIDTH)
            check(self.lib.mi_ws_colsum(self.ctx.h, n, WIDTH, self.dH.ptr, self.p(f"b{l}", True), 0, self.st))
            # dH_{l-1} = dA W_s^T + (G dA) W_n^T
            check(self.lib.mi_ws_aggregate(self.ctx.h, n, WIDTH, g.ptr.ptr, g.idx.ptr, g.w.ptr, self.dH.ptr,
                                           self.G.ptr, self.st))
            self._gemm(n, WIDTH, WIDTH, P(self.dH), WIDTH, 1, self.p(f"W_s{l}"), 1, WIDTH, P(self.dH2), WIDTH)
            self._gemm(n, WIDTH, WIDTH, P(self.G), WIDTH, 1, self.p(f"W_n{l}"), 1, WIDTH, P(self.dH2), WIDTH, acc=1)
            self.dH, self.dH2 = self.dH2, self.dH
        check(self.lib.mi_ws_relu_backward(self.ctx.h, n * WIDTH, self.H[0].ptr, self.dH.ptr, self.st))
        self._gemm(F_IN, WIDTH, n, P(g.X), 1, F_IN, P(self.dH), WIDTH, 1, self.p("W_in", True), WIDTH)
        check(self.lib.mi_ws_colsum(self.ctx.h, n, WIDTH, self.dH.ptr, self.p("b_in", True), 0, self.st))
        return value

    def adam(self, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> None:
        self.step += 1
        check(self.lib.mi_ws_adam(self.ctx.h, N_PARAMS, self.params.ptr, self.grads.ptr, self.m.ptr, self.v.ptr,
                                  C.c_float(lr), C.c_float(beta1), C.c_float(beta2), C.c_float(eps), self.step,
                                  self.st))

    # ---- the solve: learned x0 when its residual is below ||b||, else x0 = 0
    def solve(self, ls: LinearSystem, g: Graph, x: DeviceArray, **kw) -> dict:
        x0 = self.forward(g)
        self.ctx.lib.mi_memcpy_d2d(self.ctx.h, self.r.ptr, g.b.ptr, 8 * 3 * g.n)
        ls.spmv(-1.0, x0, 1.0, self.r)
        rel0 = float(np.sqrt(self.ctx.dot(self.r, self.r) / g.bb))
        if rel0 < 1.0:
            self.ctx.lib.mi_memcpy_d2d(self.ctx.h, x.ptr, x0.ptr, 8 * 3 * g.n)
            st = ls.solve(g.b, x, x_init=1, **kw)
        else:
            st = ls.solve(g.b, x, **kw)
        st["warm_rel_res0"] = rel0
        st["warm_used"] = int(rel0 < 1.0)
        return st
