"""TEST INFRASTRUCTURE ONLY.  numpy restatement of the affine-body Hessian
producer (mi_produce_abd): the BDF1 affine mass matrix
(affine_body/bdf/affine_body_bdf1_kinetic.cu:38-71, abd_jacobi_matrix.cu:115-135),
the orthogonality potential (affine_body/constitutions/ortho_potential.cu:36-70,
sym/ortho_potential.inl:3-20) with its PSD projection (utils/make_spd.h:24-46),
and the 4x4 block layout with the lower blocks zeroed
(abd_linear_subsystem.cu:21-31, 80-113)."""
from __future__ import annotations

import numpy as np


def jacobian(xbar: np.ndarray) -> np.ndarray:
    """J(xbar): x = p + A xbar = J q, q = (p, rows of A)"""
    J = np.zeros((3, 12))
    J[:, :3] = np.eye(3)
    for r in range(3):
        J[r, 3 + 3 * r:6 + 3 * r] = xbar
    return J


def dyadic_from_points(xbar: np.ndarray, m: np.ndarray) -> np.ndarray:
    """(m, m xbar, m <xbar xbar^T> as xx xy xz yy yz zz) of point masses"""
    S = (m[:, None, None] * xbar[:, :, None] * xbar[:, None, :]).sum(0)
    return np.array([m.sum(), *(m[:, None] * xbar).sum(0), S[0, 0], S[0, 1], S[0, 2], S[1, 1], S[1, 2], S[2, 2]])


def mass_matrix(d: np.ndarray) -> np.ndarray:
    m, mx = d[0], d[1:4]
    S = np.array([[d[4], d[5], d[6]], [d[5], d[7], d[8]], [d[6], d[8], d[9]]])
    M = np.zeros((12, 12))
    for r in range(3):
        M[r, r] += m
        M[r, 3 + 3 * r:6 + 3 * r] += mx
        M[3 + 3 * r:6 + 3 * r, r] += mx
        M[3 + 3 * r:6 + 3 * r, 3 + 3 * r:6 + 3 * r] += S
    return M


def ortho_energy(q: np.ndarray, kappa: float) -> float:
    A = q[3:].reshape(3, 3)
    G = A @ A.T - np.eye(3)
    return float(kappa * (G * G).sum())


def ortho_hessian(q: np.ndarray, kappa: float) -> np.ndarray:
    """9x9, blocks H_kl = 4 kappa (delta_kl A^T A + a_l a_k^T + G_kl I)"""
    A = q[3:].reshape(3, 3)
    G = A @ A.T - np.eye(3)
    AtA = A.T @ A
    H = np.zeros((9, 9))
    for k in range(3):
        for l in range(3):
            B = np.outer(A[l], A[k]) + G[k, l] * np.eye(3)
            if k == l:
                B = B + AtA
            H[3 * k:3 * k + 3, 3 * l:3 * l + 3] = 4 * kappa * B
    return H


def psd(H: np.ndarray) -> np.ndarray:
    w, V = np.linalg.eigh(H)
    return (V * np.maximum(w, 0.0)) @ V.T


def body_hessian(q, d, volume, kappa, dt, fixed=False) -> np.ndarray:
    if fixed:
        return np.eye(12)
    H = mass_matrix(d)
    H[3:, 3:] += volume * dt * dt * psd(ortho_hessian(q, kappa))
    return H


def triplets(qs, ds, volumes, kappas, dt, fixed=None, body_offset=0):
    """rows, cols, vals (column-major 3x3) as mi_produce_abd writes them"""
    n = len(qs)
    rows = np.empty(16 * n, np.int32)
    cols = np.empty(16 * n, np.int32)
    vals = np.zeros((16 * n, 9))
    for I in range(n):
        H = body_hessian(qs[I], ds[I], volumes[I], kappas[I], dt, bool(fixed[I]) if fixed is not None else False)
        for ii in range(4):
            for jj in range(4):
                t = 16 * I + 4 * ii + jj
                rows[t] = body_offset + 4 * I + ii
                cols[t] = body_offset + 4 * I + jj
                if ii <= jj:
                    vals[t] = H[3 * ii:3 * ii + 3, 3 * jj:3 * jj + 3].reshape(-1, order="F")
    return rows, cols, vals
