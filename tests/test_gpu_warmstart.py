"""The learned initial guess (libuipc_amd/warmstart.py, include/mi_warmstart.h):
the device forward pass and parameter gradient (residual and energy losses)
against a float64 numpy restatement of the same network with every
matrix-product operand rounded to BF16 as on the device, training lowers the
residual of the guess, and a solve from the guess meets the same tolerance as
from x0 = 0."""
import numpy as np
import pytest

from libuipc_amd import abi, synth
from libuipc_amd.device import LinearSystem, rel_err
from libuipc_amd.warmstart import LAYERS, LAYOUT, WarmStartNet, build_graph, coarse_of_mas

pytestmark = pytest.mark.gpu


def system(ctx, grid, contacts, seed):
    s = synth.fem_system(grid, contacts, seed=seed)
    ls = LinearSystem(ctx, s.block_rows)
    ls.assemble(ctx.to_device(s.rows), ctx.to_device(s.cols), ctx.to_device(s.vals))
    return s, ls


def unpack(p):
    return {k: p[o:o + int(np.prod(sh))].reshape(sh).astype(np.float64) for k, (o, sh) in LAYOUT.items()}


def bf16(a):
    """round to nearest even BF16 (the device rounds every matrix-core operand)"""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32).astype(np.float64)


def mm(a, b):
    return bf16(a) @ bf16(b)


def np_forward_backward(P, X, ptr, idx, w, dinv, xj, s_b, A, b, loss_kind, node=None):
    """node: each row's coarse node 0..nc-1 (the coarse-level correction), or None"""
    n = X.shape[0]
    G = np.zeros((n, n))
    for i in range(n):
        for k in range(ptr[i], ptr[i + 1]):
            G[i, idx[k]] += w[k]
    relu = lambda v: np.maximum(v, 0.0)  # noqa: E731
    H = [relu(mm(X, P["W_in"]) + P["b_in"])]
    M = [None]
    for l in range(1, LAYERS + 1):
        M.append(G @ H[-1])
        H.append(relu(mm(H[-1], P[f"W_s{l}"]) + mm(M[-1], P[f"W_n{l}"]) + P[f"b{l}"]))
    Y = mm(H[-1], P["W_out"]) + P["b_out"]
    D = dinv.reshape(n, 3, 3).transpose(0, 2, 1)  # column-major -> [r][c]
    x0 = xj + s_b * np.einsum("nij,nj->ni", D, Y).reshape(-1)
    if node is not None:
        nc = int(node.max()) + 1
        Pro = np.zeros((n, nc))
        Pro[np.arange(n), node] = 1.0
        Pool = Pro.T / Pro.sum(0)[:, None]
        s_c = float(np.sqrt(np.mean(xj * xj)))
        Hc = Pool @ H[-1]
        Yc = mm(Hc, P["W_c"]) + P["b_c"]
        x0 = x0 + s_c * (Pro @ Yc).reshape(-1)
    r = b - A @ x0
    if loss_kind == "residual":
        loss = r @ r / (b @ b)
        gx = -2.0 * (A @ r) / (b @ b)
    else:
        bxj = b @ xj
        loss = (0.5 * x0 @ (A @ x0) - b @ x0) / bxj
        gx = -r / bxj
    dY = s_b * np.einsum("nji,nj->ni", D, gx.reshape(n, 3))
    dY = dY.astype(np.float32).astype(np.float64)
    g = {"W_out": mm(H[-1].T, dY), "b_out": dY.sum(0)}
    dH = mm(dY, P["W_out"].T)
    if node is not None:
        dYc = Pro.T @ (s_c * gx.reshape(n, 3)).astype(np.float32).astype(np.float64)
        g["W_c"], g["b_c"] = mm(Hc.T, dYc), dYc.sum(0)
        dH = dH + Pool.T @ mm(dYc, P["W_c"].T)
    else:
        g["W_c"], g["b_c"] = np.zeros_like(P["W_c"]), np.zeros_like(P["b_c"])
    for l in range(LAYERS, 0, -1):
        dA = dH * (H[l] > 0)
        g[f"W_s{l}"] = mm(H[l - 1].T, dA)
        g[f"W_n{l}"] = mm(M[l].T, dA)
        g[f"b{l}"] = dA.sum(0)
        dH = mm(dA, P[f"W_s{l}"].T) + mm(G.T @ dA, P[f"W_n{l}"].T)
    dA = dH * (H[0] > 0)
    g["W_in"] = mm(X.T, dA)
    g["b_in"] = dA.sum(0)
    return x0, loss, g


def dense(s):
    from oracle import oracle as O

    r, c, v, _ = O.convert_sym(s.block_rows, s.rows, s.cols, s.vals)
    n = 3 * s.block_rows
    A = np.zeros((n, n))
    for k in range(len(r)):
        blk = v[k].reshape(3, 3).T
        A[3 * r[k]:3 * r[k] + 3, 3 * c[k]:3 * c[k] + 3] += blk
        if r[k] != c[k]:
            A[3 * c[k]:3 * c[k] + 3, 3 * r[k]:3 * r[k] + 3] += blk.T
    return A


@pytest.mark.parametrize("coarse", [False, True], ids=["fine", "coarse"])
@pytest.mark.parametrize("loss_kind", ["residual", "energy"])
def test_forward_and_gradient_match_numpy(ctx, loss_kind, coarse):
    """with coarse: the MAS level-1 nodes as the coarse level of the
    correction head (unpartitioned rows a node each)"""
    s, ls = system(ctx, 5, 40, 31)
    node = None
    if coarse:
        parts = s.part_ids.copy()
        parts[::11] = -1
        h = ls.add_preconditioner(abi.MI_PRECOND_MAS, 0, s.n, part_ids=parts, tets=s.tets)
        cmap = coarse_of_mas(ls, h)
        assert (cmap < 0).sum() == (parts < 0).sum()
        g = build_graph(ctx, ls, s.rhs, coarse=cmap)
        c = cmap.astype(np.int64).copy()
        c[c < 0] = c.max() + 1 + np.arange(int((c < 0).sum()))
        node = np.unique(c, return_inverse=True)[1]
        assert g.nc == node.max() + 1 and g.nc < s.block_rows
    else:
        g = build_graph(ctx, ls, s.rhs)
    net = WarmStartNet(ctx, seed=3)
    p = net.numpy()
    rng = np.random.default_rng(4)
    for name in ("W_out", "W_c"):
        o, sh = LAYOUT[name]
        p[o:o + int(np.prod(sh))] = 0.05 * rng.standard_normal(int(np.prod(sh)))
    net.load(p)
    loss = net.loss_and_grad(ls, g, loss=loss_kind)
    x0 = net.x0.numpy()
    grads = unpack(net.grads.numpy())
    X = g.X.numpy().astype(np.float64)
    xw, lw, gw = np_forward_backward(unpack(p), X, g.ptr.numpy(), g.idx.numpy(), g.w.numpy().astype(np.float64),
                                     g.dinv.numpy(), g.x_jacobi.numpy(), g.s_b, dense(s), s.rhs, loss_kind, node)
    # with the operands rounded to BF16 as the device rounds them, what is left
    # is FP32 accumulation against float64
    assert rel_err(x0 - g.x_jacobi.numpy(), xw - g.x_jacobi.numpy()) < 1e-3
    assert abs(loss - lw) <= 1e-3 * abs(lw)
    for k, want in gw.items():
        got = grads[k]
        assert np.abs(got - want).max() <= 1e-2 * np.abs(want).max() + 1e-12, k


def test_training_lowers_the_residual_and_solves_match(ctx):
    systems = [system(ctx, 8, 150, 40 + k) for k in range(4)]
    graphs = [build_graph(ctx, ls, s.rhs) for s, ls in systems]
    net = WarmStartNet(ctx, seed=1)
    first = [net.loss_and_grad(ls, g) for (s, ls), g in zip(systems, graphs)]
    for step in range(60):
        k = step % len(systems)
        net.loss_and_grad(systems[k][1], graphs[k])
        net.adam(lr=2e-3)
    last = [net.loss_and_grad(ls, g) for (s, ls), g in zip(systems, graphs)]
    assert np.mean(last) < np.mean(first)
    s, ls = systems[0]
    ls.add_preconditioner(abi.MI_PRECOND_MAS, 0, s.n, part_ids=s.part_ids, tets=s.tets)
    ls.assemble_preconditioners()
    kw = dict(criterion=abi.MI_CRIT_REL_RESIDUAL, tol=1e-4, true_residual=1)
    x = ctx.alloc(s.n, np.float64)
    cold = ls.solve(graphs[0].b, x, **kw)
    warm = net.solve(ls, graphs[0], x, **kw)
    assert warm["warm_used"] == 1 and warm["warm_rel_res0"] < 1.0
    assert warm["converged"] == 1 and warm["rel_res_true"] < 2e-4 and cold["converged"] == 1
