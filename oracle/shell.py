"""TEST INFRASTRUCTURE ONLY.  numpy restatement of the cloth Hessian producer
(mi_produce_bw_shell): the strain-limiting Baraff-Witkin shell of
finite_element/constitutions/strain_limiting_baraff_witkin_shell_2d.cu:78-148 and
strain_limiting_baraff_witkin_shell_2d.h (Dm2x2 :35-73, dFdX :76-105, E
:108-137, ddEddF :180-248), written as 3-node half blocks
(utils/matrix_assembler.h:312-437)."""
from __future__ import annotations

import numpy as np


def rest_dm(x0, x1, x2) -> np.ndarray:
    """the rest triangle's edges in its own plane (rotated so its normal is +z)"""
    n = np.cross(x1 - x0, x2 - x0)
    n = n / np.linalg.norm(n)
    c = n[2]
    if c + 1.0 == 0.0:
        R = np.diag([-1.0, -1.0, 1.0])
    else:
        v = np.cross(n, [0.0, 0.0, 1.0])
        K = np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])
        R = np.eye(3) + K + K @ K / (1.0 + c)
    u = [(R @ p)[:2] for p in (x0, x1, x2)]
    return np.column_stack([u[1] - u[0], u[2] - u[0]])


def energy_F(F, stretch, shear, rate) -> float:
    su, sv = np.linalg.norm(F[:, 0]), np.linalg.norm(F[:, 1])
    I6 = F[:, 0] @ F[:, 1]
    e = (su - 1) ** 2 + (rate * (su - 1) ** 3 if su > 1 else 0.0) + (sv - 1) ** 2 + (rate * (sv - 1) ** 3 if sv > 1 else 0.0)
    return stretch * e + shear * I6 * I6


def hessian_F(F, stretch, shear, rate) -> np.ndarray:
    """6x6 in vec(F) = (f_u, f_v), the reference's closed-form PSD Hessian"""
    fu, fv = F[:, 0], F[:, 1]
    I5u, I5v = max(fu @ fu, 1e-12), max(fv @ fv, 1e-12)
    su, sv = np.sqrt(I5u), np.sqrt(I5v)
    Hs = np.zeros((6, 6))
    if su > 1:
        Hs[:3, :3] += np.eye(3) * (su - 1) * (3 * su * rate - 3 * rate + 2) / su
    if sv > 1:
        Hs[3:, 3:] += np.eye(3) * (sv - 1) * (3 * sv * rate - 3 * rate + 2) / sv
    cu = (3 * I5u * rate - 3 * rate + 2) / su if su > 1 else 2.0
    cv = (3 * I5v * rate - 3 * rate + 2) / sv if sv > 1 else 2.0
    Hs[:3, :3] += cu * np.outer(fu / su, fu / su)
    Hs[3:, 3:] += cv * np.outer(fv / sv, fv / sv)
    S = np.zeros((6, 6))
    S[np.arange(3), 3 + np.arange(3)] = 1
    S[3 + np.arange(3), np.arange(3)] = 1
    I6 = fu @ fv
    sg = 1.0 if I6 >= 0 else -1.0
    g = np.concatenate([fv, fu])
    I2 = fu @ fu + fv @ fv
    lam0 = 0.5 * (I2 + np.sqrt(I2 * I2 + 12 * I6 * I6))
    q0 = I6 * S @ g + lam0 * g
    q0 = q0 / np.linalg.norm(q0)
    T = 0.5 * (np.eye(6) + sg * S)
    Tq = T @ q0
    Hh = 2 * (abs(I6) * (T - np.outer(Tq, Tq) / (Tq @ Tq)) + lam0 * np.outer(q0, q0))
    return stretch * Hs + shear * Hh


def dFdx(B) -> np.ndarray:
    """6x9: vec(F) = (f_u, f_v) against (x0, x1, x2)"""
    D = np.zeros((6, 9))
    for c in range(2):
        w = [-(B[0, c] + B[1, c]), B[0, c], B[1, c]]
        for k in range(3):
            D[3 * c:3 * c + 3, 3 * k:3 * k + 3] = w[k] * np.eye(3)
    return D


def triangle(Xr, xc, stretch, shear, rate, dt):
    """(9x9 Hessian, energy) of one triangle"""
    Dm = rest_dm(*Xr)
    B = np.linalg.inv(Dm)
    area = 0.5 * np.linalg.norm(np.cross(Xr[1] - Xr[0], Xr[2] - Xr[0]))
    F = np.column_stack([xc[1] - xc[0], xc[2] - xc[0]]) @ B
    D = dFdx(B)
    H = area * dt * dt * D.T @ hessian_F(F, stretch, shear, rate) @ D
    return H, area * dt * dt * energy_F(F, stretch, shear, rate)


def triplets(tris, X, x, stretch, shear, rate, dt, vert_offset=0):
    X = np.asarray(X).reshape(-1, 3)
    x = np.asarray(x).reshape(-1, 3)
    n = len(tris)
    rows = np.empty(6 * n, np.int32)
    cols = np.empty(6 * n, np.int32)
    vals = np.empty((6 * n, 9))
    for t, tri in enumerate(tris):
        H, _ = triangle(X[tri], x[tri], stretch, shear, rate, dt)
        idx = tri + vert_offset
        s = 6 * t
        for a in range(3):
            for b in range(a, 3):
                K = H[3 * a:3 * a + 3, 3 * b:3 * b + 3]
                if idx[a] > idx[b]:
                    rows[s], cols[s], K = idx[b], idx[a], K.T
                else:
                    rows[s], cols[s] = idx[a], idx[b]
                vals[s] = K.reshape(-1, order="F")
                s += 1
    return rows, cols, vals


# ---- discrete shell bending (finite_element/constitutions/discrete_shell_bending.cu:84-156)
def dihedral_angle(v0, v1, v2, v3) -> float:
    """utils/dihedral_angle.h:14-30: hinge V1-V2, wings V0 and V3"""
    n1 = np.cross(v1 - v0, v2 - v0)
    n2 = np.cross(v2 - v3, v1 - v3)
    th = np.arccos(np.clip(n1 @ n2 / np.sqrt((n1 @ n1) * (n2 @ n2)), -1.0, 1.0))
    return -th if np.cross(n2, n1) @ (v1 - v2) < 0 else th


def dihedral_gradient(V0, V1, V2, V3) -> np.ndarray:
    e0, e1, e2, e3, e4 = V2 - V1, V0 - V1, V3 - V1, V0 - V2, V3 - V2
    n1, n2 = np.cross(e0, e1), np.cross(e2, e0)
    q1, q2, l0 = n1 @ n1, n2 @ n2, np.linalg.norm(e0)
    return np.concatenate([-l0 / q1 * n1,
                           -(e0 @ e3) / (l0 * q1) * n1 - (e0 @ e4) / (l0 * q2) * n2,
                           (e0 @ e1) / (l0 * q1) * n1 + (e0 @ e2) / (l0 * q2) * n2,
                           -l0 / q2 * n2])


def bending_rest(X0, X1, X2, X3):
    """(L0, h_bar, theta_bar) of discrete_shell_bending_reference.h:8-27"""
    L0 = np.linalg.norm(X2 - X1)
    A = 0.5 * (np.linalg.norm(np.cross(X1 - X0, X2 - X0)) + np.linalg.norm(np.cross(X2 - X3, X1 - X3)))
    return L0, A / 3.0 / L0, dihedral_angle(X0, X1, X2, X3)


def bending_energy(Xh, xh, kappa, dt) -> float:
    L0, hb, tb = bending_rest(*Xh)
    return dt * dt * L0 * kappa * (dihedral_angle(*xh) - tb) ** 2 / hb


def bending_hessian(Xh, xh, kappa, dt) -> np.ndarray:
    """Gauss-Newton: dt^2 (2 L0 kappa / h_bar) grad grad^T"""
    L0, hb, _ = bending_rest(*Xh)
    g = dihedral_gradient(*xh)
    return dt * dt * 2 * L0 * kappa / hb * np.outer(g, g)


def bending_triplets(hinges, X, x, kappa, dt, vert_offset=0):
    X = np.asarray(X).reshape(-1, 3)
    x = np.asarray(x).reshape(-1, 3)
    n = len(hinges)
    rows = np.empty(10 * n, np.int32)
    cols = np.empty(10 * n, np.int32)
    vals = np.empty((10 * n, 9))
    for h, st in enumerate(hinges):
        H = bending_hessian(X[st], x[st], kappa, dt)
        idx = st + vert_offset
        s = 10 * h
        for a in range(4):
            for b in range(a, 4):
                K = H[3 * a:3 * a + 3, 3 * b:3 * b + 3]
                if idx[a] > idx[b]:
                    rows[s], cols[s], K = idx[b], idx[a], K.T
                else:
                    rows[s], cols[s] = idx[a], idx[b]
                vals[s] = K.reshape(-1, order="F")
                s += 1
    return rows, cols, vals
