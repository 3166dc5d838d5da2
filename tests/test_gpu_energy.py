"""Line-search side of the Newton step (include/mi_producers.h): the trial step
x = x0 + alpha dx and the kinetic and stable Neo-Hookean energies, against
numpy restatements of the reference's energy kernels, the closed form at
rest (F = I: Psi = mu^2 / (2 lambda)), and the Hessian producer: along any
direction d the second difference of the energy equals d^T H d where the
PSD projection is inactive (a uniformly stretched mesh)."""
import numpy as np
import pytest

from libuipc_amd import synth
from libuipc_amd.abi import check

pytestmark = pytest.mark.gpu


def snh_energy_np(tets, X, x, mu, lam, dt):
    X = X.reshape(-1, 3)
    x = x.reshape(-1, 3)
    Dm = np.stack([X[tets[:, c]] - X[tets[:, 0]] for c in (1, 2, 3)], axis=2)
    Ds = np.stack([x[tets[:, c]] - x[tets[:, 0]] for c in (1, 2, 3)], axis=2)
    F = Ds @ np.linalg.inv(Dm)
    vol = np.abs(np.linalg.det(Dm)) / 6.0
    J = np.linalg.det(F)
    Ic = (F * F).sum(axis=(1, 2))
    psi = 0.5 * (mu * (Ic - 3.0) + lam * (J - 1.0 - mu / lam) ** 2)
    return float((psi * vol * dt * dt).sum())


class Energies:
    def __init__(self, ctx, p):
        self.ctx, self.p = ctx, p
        self.tets = ctx.to_device(p.tets)
        self.X = ctx.to_device(p.rest_positions)
        self.mass = ctx.to_device(p.mass)
        self.out = ctx.alloc(1, np.float64)

    def snh(self, x):
        xd = self.ctx.to_device(x)
        check(self.ctx.lib.mi_energy_snh(self.ctx.h, len(self.p.tets), self.tets.ptr, self.X.ptr, xd.ptr, self.p.mu,
                                         self.p.lam, self.p.dt, self.out.ptr, None))
        return float(self.out.numpy()[0])

    def kinetic(self, x, xt):
        xd, td = self.ctx.to_device(x), self.ctx.to_device(xt)
        check(self.ctx.lib.mi_energy_kinetic(self.ctx.h, len(self.p.mass), xd.ptr, td.ptr, self.mass.ptr,
                                             self.out.ptr, None))
        return float(self.out.numpy()[0])


@pytest.fixture(scope="module")
def p8():
    return synth.producer_inputs(8, 0, seed=4)


def test_snh_energy_matches_numpy_and_is_reproducible(ctx, p8):
    e = Energies(ctx, p8)
    got = e.snh(p8.positions)
    want = snh_energy_np(p8.tets, p8.rest_positions, p8.positions, p8.mu, p8.lam, p8.dt)
    assert abs(got - want) <= 1e-12 * abs(want)
    assert e.snh(p8.positions) == got  # same bits


def test_snh_energy_at_rest_closed_form(ctx, p8):
    e = Energies(ctx, p8)
    X = p8.rest_positions.reshape(-1, 3)
    vol = np.abs(np.linalg.det(np.stack([X[p8.tets[:, c]] - X[p8.tets[:, 0]] for c in (1, 2, 3)], axis=2))) / 6
    want = p8.dt ** 2 * vol.sum() * p8.mu ** 2 / (2 * p8.lam)
    assert abs(e.snh(p8.rest_positions) - want) <= 1e-12 * want


def test_kinetic_energy(ctx, p8):
    e = Energies(ctx, p8)
    rng = np.random.default_rng(2)
    xt = p8.positions + 1e-3 * rng.standard_normal(p8.positions.shape)
    d = (p8.positions - xt).reshape(-1, 3)
    want = float((0.5 * p8.mass * (d * d).sum(1)).sum())
    assert abs(e.kinetic(p8.positions, xt) - want) <= 1e-12 * want
    assert e.kinetic(xt, xt) == 0.0


def test_step_forward(ctx):
    rng = np.random.default_rng(3)
    n = 100_003
    x0, dx = rng.standard_normal(n), rng.standard_normal(n)
    a, b, out = ctx.to_device(x0), ctx.to_device(dx), ctx.alloc(n, np.float64)
    for alpha in (1.0, 0.5, 0.125, 0.0):
        check(ctx.lib.mi_step_forward(ctx.h, a.ptr, b.ptr, alpha, out.ptr, n, None))
        assert np.allclose(out.numpy(), x0 + alpha * dx, rtol=1e-15, atol=1e-15)
    check(ctx.lib.mi_step_forward(ctx.h, a.ptr, b.ptr, 0.0, out.ptr, n, None))
    assert np.array_equal(out.numpy(), x0)


def test_energy_second_difference_is_the_produced_hessian(ctx):
    """stretched 1.1x (all SNH eigenvalues > 0, so no clamping): the second
    difference of E_snh + E_kin along random d equals d^T H d of the
    kinetic + SNH triplets that mi_produce_* write"""
    p = synth.producer_inputs(5, 0, seed=8)
    B, nt = len(p.mass), len(p.tets)
    x = 1.1 * p.rest_positions + 1e-3 * np.random.default_rng(1).standard_normal(3 * B)
    e = Energies(ctx, p)
    # the produced Hessian at x
    T = B + 10 * nt
    rows, cols, vals = ctx.alloc(T, np.int32), ctx.alloc(T, np.int32), ctx.alloc((T, 9), np.float64)
    xd = ctx.to_device(x)
    check(ctx.lib.mi_produce_kinetic(ctx.h, B, e.mass.ptr, 0, rows.ptr, cols.ptr, vals.ptr, None))
    check(ctx.lib.mi_produce_snh(ctx.h, nt, e.tets.ptr, e.X.ptr, xd.ptr, p.mu, p.lam, p.dt, 0,
                                 rows.ptr.value + 4 * B, cols.ptr.value + 4 * B, vals.ptr.value + 72 * B, None))
    r, c, v = rows.numpy(), cols.numpy(), vals.numpy().reshape(-1, 3, 3, order="F")
    d = np.random.default_rng(5).standard_normal(3 * B)
    dv = d.reshape(-1, 3)
    q = np.einsum("ti,tij,tj->t", dv[r], v, dv[c])
    dHd = float(np.where(r == c, q, 2 * q).sum())
    xt = p.rest_positions  # kinetic target: any fixed x_tilde

    def E(y):
        return e.snh(y) + e.kinetic(y, xt)

    h = 1e-4 * np.abs(x).max() / np.abs(d).max()
    fd = (E(x + h * d) - 2 * E(x) + E(x - h * d)) / (h * h)
    assert abs(fd - dHd) <= 2e-5 * abs(dHd), (fd, dHd)
