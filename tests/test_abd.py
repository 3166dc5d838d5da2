"""Pins the oracle's affine-body restatement (oracle/abd.py) against what the
mathematics fixes: the mass-matrix layout equals sum_k m_k J(xbar_k)^T J(xbar_k)
over point masses, the ortho-potential Hessian equals central differences of
its energy, closed forms at rotations and uniform scalings, and the PSD
projection.  CPU only."""
import numpy as np

from libuipc_amd import synth
from oracle import abd as OA


def test_mass_matrix_is_integrated_jacobian():
    rng = np.random.default_rng(1)
    xbar = rng.standard_normal((20, 3))
    m = rng.uniform(0.1, 2.0, 20)
    M = OA.mass_matrix(OA.dyadic_from_points(xbar, m))
    want = sum(mk * OA.jacobian(xk).T @ OA.jacobian(xk) for xk, mk in zip(xbar, m))
    assert np.allclose(M, want, rtol=1e-13, atol=1e-13)


def test_ortho_hessian_is_the_second_derivative_of_the_energy():
    rng = np.random.default_rng(2)
    for trial in range(4):
        q = rng.standard_normal(12)
        H = OA.ortho_hessian(q, 3.0)
        h = 1e-4
        fd = np.zeros((9, 9))
        for a in range(9):
            for b in range(9):
                e = np.zeros(12)
                f = np.zeros(12)
                e[3 + a] = h
                f[3 + b] = h
                fd[a, b] = (OA.ortho_energy(q + e + f, 3.0) - OA.ortho_energy(q + e - f, 3.0)
                            - OA.ortho_energy(q - e + f, 3.0) + OA.ortho_energy(q - e - f, 3.0)) / (4 * h * h)
        assert np.abs(fd - H).max() <= 1e-5 * np.abs(H).max()


def test_ortho_closed_forms():
    R, _ = np.linalg.qr(np.random.default_rng(3).standard_normal((3, 3)))
    q = np.concatenate([np.zeros(3), R.reshape(-1)])
    assert abs(OA.ortho_energy(q, 5.0)) < 1e-24
    assert np.linalg.eigvalsh(OA.ortho_hessian(q, 5.0)).min() > -1e-12
    s = 1.3
    q = np.concatenate([np.zeros(3), (s * np.eye(3)).reshape(-1)])
    assert np.isclose(OA.ortho_energy(q, 5.0), 5.0 * 3 * (s * s - 1) ** 2)


def test_psd_projection():
    rng = np.random.default_rng(4)
    B = rng.standard_normal((9, 9))
    H = B + B.T
    P = OA.psd(H)
    assert np.linalg.eigvalsh(P).min() > -1e-12
    assert np.allclose(OA.psd(B @ B.T), B @ B.T, atol=1e-12)


def test_generated_bodies_layout():
    a = synth.abd_inputs(5, seed=1)
    r, c, v = OA.triplets(a.q, a.dyadic, a.volume, a.kappa, a.dt, body_offset=7)
    assert np.array_equal(r, 7 + np.repeat(4 * np.arange(5), 16) + np.tile(np.repeat(np.arange(4), 4), 5))
    assert np.array_equal(c, 7 + np.repeat(4 * np.arange(5), 16) + np.tile(np.arange(4), 20))
    lower = (r > c)
    assert not v[lower].any()
    # every body block is symmetric PSD and mass-dominant at kappa = 1e7, dt = 0.01
    H = OA.body_hessian(a.q[0], a.dyadic[0], a.volume[0], a.kappa[0], a.dt)
    assert np.allclose(H, H.T) and np.linalg.eigvalsh(H).min() > 0
