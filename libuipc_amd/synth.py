"""Seeded synthetic systems (host generator in libmi_synth.so, include/mi_synth.h).

fem_system(grid, n_contacts, seed) returns the triplets of
    A = M + dt^2 (K_snh_psd + sum kappa g g^T)
in libuipc's producer order, the right-hand side, and the topology/partition
the MAS preconditioner needs.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import abi


@dataclass
class FemSystem:
    grid: int
    block_rows: int
    rows: np.ndarray  # int32[T]
    cols: np.ndarray  # int32[T]
    vals: np.ndarray  # float64[T, 9], column-major 3x3
    rhs: np.ndarray  # float64[3B]
    tets: np.ndarray  # int32[n_tets, 4]
    part_ids: np.ndarray  # int32[B]
    n_contacts: int

    @property
    def n(self) -> int:
        return 3 * self.block_rows

    @property
    def n_triplets(self) -> int:
        return int(self.rows.size)


def fem_desc(grid: int, n_contacts: int = 0, seed: int = 1, *, jitter: float = 0.05, youngs: float = 1e5,
             poisson: float = 0.4, density: float = 1e3, dt: float = 0.01, contact_cells: int = 2,
             kappa_lo: float = 1e4, kappa_hi: float = 1e8, threads: int = 0) -> abi.SynthFemDesc:
    return abi.SynthFemDesc(grid, jitter, seed, youngs, poisson, density, dt, n_contacts, contact_cells, kappa_lo,
                            kappa_hi, threads)


def fem_system(grid: int, n_contacts: int = 0, seed: int = 1, **kw) -> FemSystem:
    lib = abi.synth()
    d = fem_desc(grid, n_contacts, seed, **kw)
    sz = abi.SynthFemSizes()
    if lib.mi_synth_fem_sizes_of(C.byref(d), C.byref(sz)) != 0:
        raise RuntimeError(lib.mi_synth_last_error().decode())
    T, B, nt = sz.n_triplets, sz.block_rows, sz.n_tets
    rows = np.empty(T, np.int32)
    cols = np.empty(T, np.int32)
    vals = np.empty((T, 9), np.float64)
    rhs = np.empty(3 * B, np.float64)
    tets = np.empty((nt, 4), np.int32)
    parts = np.empty(B, np.int32)
    rc = lib.mi_synth_fem_fill(C.byref(d), rows.ctypes.data_as(abi.c_i32p), cols.ctypes.data_as(abi.c_i32p),
                               vals.ctypes.data_as(abi.c_f64p), rhs.ctypes.data_as(abi.c_f64p),
                               tets.ctypes.data_as(abi.c_i32p), parts.ctypes.data_as(abi.c_i32p), None)
    if rc != 0:
        raise RuntimeError(lib.mi_synth_last_error().decode())
    return FemSystem(grid, B, rows, cols, vals, rhs, tets, parts, n_contacts)


@dataclass
class Topology:
    """Mesh and MAS partition of fem_system(grid, ..., seed), without the matrix."""
    block_rows: int
    tets: np.ndarray  # int32[n_tets, 4]
    part_ids: np.ndarray  # int32[B]


def generate_topology(grid: int, seed: int = 1, **kw) -> Topology:
    lib = abi.synth()
    d = fem_desc(grid, 0, seed, **kw)
    sz = abi.SynthFemSizes()
    if lib.mi_synth_fem_sizes_of(C.byref(d), C.byref(sz)) != 0:
        raise RuntimeError(lib.mi_synth_last_error().decode())
    tets = np.empty((sz.n_tets, 4), np.int32)
    parts = np.empty(sz.block_rows, np.int32)
    if lib.mi_synth_fem_fill(C.byref(d), None, None, None, None, tets.ctypes.data_as(abi.c_i32p),
                             parts.ctypes.data_as(abi.c_i32p), None) != 0:
        raise RuntimeError(lib.mi_synth_last_error().decode())
    return Topology(sz.block_rows, tets, parts)


def sizes(grid: int, n_contacts: int = 0) -> dict:
    lib = abi.synth()
    d = fem_desc(grid, n_contacts)
    sz = abi.SynthFemSizes()
    if lib.mi_synth_fem_sizes_of(C.byref(d), C.byref(sz)) != 0:
        raise RuntimeError(lib.mi_synth_last_error().decode())
    return {"block_rows": sz.block_rows, "n_tets": sz.n_tets, "n_triplets": sz.n_triplets, "n_parts": sz.n_parts}


@dataclass
class ProducerInputs:
    """What the device producers (include/mi_producers.h) need for the system
    fem_system(grid, n_contacts, seed) generates on the host."""
    rest_positions: np.ndarray  # float64[3B]
    positions: np.ndarray  # float64[3B], deformed
    mass: np.ndarray  # float64[B]
    tets: np.ndarray  # int32[n_tets, 4]
    contact_idx: np.ndarray  # int32[n_contacts, 4]
    contact_g: np.ndarray  # float64[n_contacts, 12]
    contact_kappa: np.ndarray  # float64[n_contacts]
    mu: float
    lam: float
    dt: float
    part_ids: np.ndarray | None = None  # int32[B], the MAS partition of fem_system


def producer_inputs(grid: int, n_contacts: int = 0, seed: int = 1, **kw) -> ProducerInputs:
    lib = abi.synth()
    d = fem_desc(grid, n_contacts, seed, **kw)
    sz = abi.SynthFemSizes()
    if lib.mi_synth_fem_sizes_of(C.byref(d), C.byref(sz)) != 0:
        raise RuntimeError(lib.mi_synth_last_error().decode())
    B, nt = sz.block_rows, sz.n_tets
    X = np.empty(3 * B)
    x = np.empty(3 * B)
    m = np.empty(B)
    tets = np.empty((nt, 4), np.int32)
    ci = np.empty((max(n_contacts, 1), 4), np.int32)
    cg = np.empty((max(n_contacts, 1), 12))
    ck = np.empty(max(n_contacts, 1))
    parts = np.empty(B, np.int32)
    if lib.mi_synth_fem_fill(C.byref(d), None, None, None, None, tets.ctypes.data_as(abi.c_i32p),
                             parts.ctypes.data_as(abi.c_i32p), x.ctypes.data_as(abi.c_f64p)) != 0:
        raise RuntimeError(lib.mi_synth_last_error().decode())
    if lib.mi_synth_fem_aux(C.byref(d), X.ctypes.data_as(abi.c_f64p), m.ctypes.data_as(abi.c_f64p),
                            ci.ctypes.data_as(abi.c_i32p), cg.ctypes.data_as(abi.c_f64p),
                            ck.ctypes.data_as(abi.c_f64p)) != 0:
        raise RuntimeError(lib.mi_synth_last_error().decode())
    E, nu = d.youngs, d.poisson
    mu = E / (2.0 * (1.0 + nu))
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    return ProducerInputs(X, x, m, tets, ci[:n_contacts], cg[:n_contacts], ck[:n_contacts], mu, lam, d.dt, parts)


@dataclass
class AbdInputs:
    """Affine bodies for mi_produce_abd: cubes of random size and position,
    q near a rotation (A = R + noise), kappa = 1e7, density 1e3."""
    q: np.ndarray  # float64[n, 12]
    dyadic: np.ndarray  # float64[n, 10]: m, m xbar, m <xbar xbar^T> (xx xy xz yy yz zz)
    volume: np.ndarray  # float64[n]
    kappa: np.ndarray  # float64[n]
    dt: float


def abd_inputs(n: int, seed: int = 1, *, kappa: float = 1e7, density: float = 1e3, dt: float = 0.01,
               noise: float = 1e-2) -> AbdInputs:
    rng = np.random.default_rng(seed)
    side = rng.uniform(0.05, 0.2, n)
    c = rng.uniform(-1.0, 1.0, (n, 3))  # centre of each cube in its rest frame
    m = density * side ** 3
    S = m[:, None, None] * (c[:, :, None] * c[:, None, :] + (side ** 2 / 12.0)[:, None, None] * np.eye(3))
    dyadic = np.column_stack([m, m[:, None] * c, S[:, 0, 0], S[:, 0, 1], S[:, 0, 2], S[:, 1, 1], S[:, 1, 2],
                              S[:, 2, 2]])
    R, _ = np.linalg.qr(rng.standard_normal((n, 3, 3)))
    A = R + noise * rng.standard_normal((n, 3, 3))
    q = np.concatenate([rng.uniform(-5.0, 5.0, (n, 3)), A.reshape(n, 9)], axis=1)
    return AbdInputs(q, dyadic, side ** 3, np.full(n, kappa), dt)


@dataclass
class ClothInputs:
    """A square cloth grid for mi_produce_bw_shell: the rest grid in the z = 0
    plane, deformed positions (a stretch or compression per direction, a shear,
    a rigid rotation and per-vertex noise, so that every branch of the
    constitution is taken)."""
    tris: np.ndarray  # int32[T, 3]
    rest_positions: np.ndarray  # float64[V, 3]
    positions: np.ndarray  # float64[V, 3]
    stretch: float
    shear: float
    strain_rate: float
    dt: float
    bending: float  # discrete shell bending stiffness E t^3 / (12 (1 - nu^2))


def cloth_inputs(n: int, seed: int = 1, *, youngs: float = 1e5, poisson: float = 0.3, thickness: float = 1e-3,
                 strain_rate: float = 100.0, dt: float = 0.01, noise: float = 0.05) -> ClothInputs:
    """n x n vertices, 2 (n - 1)^2 triangles; stiffness as
    StrainLimitingBaraffWitkin::apply_to: stretch = (lambda + 2 mu) 2 thickness,
    shear = mu (src/constitution/strain_limiting_baraff_witkin.cpp:35-44)"""
    rng = np.random.default_rng(seed)
    h = 1.0 / (n - 1)
    gx, gy = np.meshgrid(np.arange(n) * h, np.arange(n) * h, indexing="xy")
    X = np.column_stack([gx.ravel(), gy.ravel(), np.zeros(n * n)])
    idx = np.arange(n * n, dtype=np.int32).reshape(n, n)
    a, b, c, d = idx[:-1, :-1].ravel(), idx[:-1, 1:].ravel(), idx[1:, :-1].ravel(), idx[1:, 1:].ravel()
    tris = np.concatenate([np.stack([a, b, d], 1), np.stack([a, d, c], 1)]).astype(np.int32)
    Fd = np.array([[1.05, 0.08, 0.0], [-0.03, 0.97, 0.0], [0.0, 0.0, 1.0]])
    R, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    x = X @ Fd.T @ R.T + noise * h * rng.standard_normal(X.shape)
    mu = youngs / (2 * (1 + poisson))
    lam = youngs * poisson / (1 - poisson * poisson)  # plane-stress lambda
    bending = youngs * thickness ** 3 / (12 * (1 - poisson * poisson))
    return ClothInputs(tris, X, x, (lam + 2 * mu) * 2 * thickness, mu, strain_rate, dt, bending)


def hinges_of(tris: np.ndarray) -> np.ndarray:
    """Bending stencils (X0, X1, X2, X3) of the interior edges, as
    DiscreteShellBending::do_init builds them (discrete_shell_bending.cu:258-290):
    per edge (X1 < X2) shared by exactly two triangles, X0 and X3 the lower and
    the higher opposite vertex, edges in ascending (X1, X2) order."""
    T = np.sort(np.asarray(tris, np.int64), axis=1)
    e = np.concatenate([T[:, [0, 1]], T[:, [0, 2]], T[:, [1, 2]]])
    o = np.concatenate([T[:, 2], T[:, 1], T[:, 0]])
    order = np.lexsort((o, e[:, 1], e[:, 0]))
    e, o = e[order], o[order]
    key = e[:, 0] * (int(T.max()) + 1) + e[:, 1]
    start = np.flatnonzero(np.r_[True, key[1:] != key[:-1]])
    count = np.diff(np.r_[start, len(key)])
    two = start[count == 2]
    return np.column_stack([o[two], e[two, 0], e[two, 1], o[two + 1]]).astype(np.int32)


def cloth_contacts(n: int, n_contacts: int, seed: int = 1, *, cells: int = 2, kappa_lo: float = 1e4,
                   kappa_hi: float = 1e8) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Vertex-triangle stencils on the n x n cloth grid of cloth_inputs (the
    recipe of the FEM generator's contacts): a random vertex p and a random
    triangle of a cell within `cells` cells of it that does not contain p;
    g ~ N(0, 1)^12 and kappa log-uniform in [kappa_lo, kappa_hi].
    Returns (stencils int32[m, 4], g float64[m, 12], kappa float64[m])."""
    rng = np.random.default_rng(seed)
    got = []
    have = 0
    while have < n_contacts:
        k = 2 * (n_contacts - have) + 16
        pi = rng.integers(0, n, k)
        pj = rng.integers(0, n, k)
        ci = np.clip(pi + rng.integers(-cells, cells, k), 0, n - 2)
        cj = np.clip(pj + rng.integers(-cells, cells, k), 0, n - 2)
        upper = rng.integers(0, 2, k).astype(bool)
        a = ci * n + cj
        tri = np.where(upper[:, None], np.stack([a, a + 1, a + n + 1], 1), np.stack([a, a + n + 1, a + n], 1))
        p = pi * n + pj
        keep = ~(tri == p[:, None]).any(1)  # the vertex is not a corner of its triangle
        got.append(np.column_stack([p, tri])[keep])
        have += int(keep.sum())
    st = np.concatenate(got)[:n_contacts].astype(np.int32)
    m = len(st)
    g = rng.standard_normal((m, 12))
    kappa = np.exp(rng.uniform(np.log(kappa_lo), np.log(kappa_hi), m))
    return st, g, kappa


@dataclass
class Batch:
    """Independent FEM + contact systems on consecutive block rows (config 4's
    recipe): per system a grid of N = round((DoF / 3)^(1/3)) vertices per side
    for DoF log-uniform in [dof_lo, dof_hi], contacts at the 64^3 system's
    density (2M per 262,144 vertices)."""
    systems: list  # FemSystem per system
    seg_rows: np.ndarray  # int32[nseg + 1]
    rows: np.ndarray
    cols: np.ndarray
    vals: np.ndarray
    rhs: np.ndarray
    tets: np.ndarray  # global vertex ids
    part_ids: np.ndarray  # global, never crossing a system

    @property
    def block_rows(self) -> int:
        return int(self.seg_rows[-1])


def batch_systems(n_systems: int, seed: int = 1, *, dof_lo: float = 3e3, dof_hi: float = 6e5) -> Batch:
    rng = np.random.default_rng(seed)
    dofs = np.exp(rng.uniform(np.log(dof_lo), np.log(dof_hi), n_systems))
    grids = np.maximum(4, np.round((dofs / 3.0) ** (1.0 / 3.0)).astype(int))
    density = 2_000_000 / 64 ** 3
    systems, seg = [], [0]
    for k, g in enumerate(grids):
        s = fem_system(int(g), int(round(density * g ** 3)), seed=seed * 1000 + k)
        systems.append(s)
        seg.append(seg[-1] + s.block_rows)
    seg = np.array(seg, np.int32)
    rows = np.concatenate([s.rows + seg[k] for k, s in enumerate(systems)]).astype(np.int32)
    cols = np.concatenate([s.cols + seg[k] for k, s in enumerate(systems)]).astype(np.int32)
    vals = np.concatenate([s.vals for s in systems])
    rhs = np.concatenate([s.rhs for s in systems])
    tets = np.concatenate([s.tets + seg[k] for k, s in enumerate(systems)]).astype(np.int32)
    parts, off = [], 0
    for s in systems:
        parts.append(np.where(s.part_ids >= 0, s.part_ids + off, -1))
        off += int(s.part_ids.max()) + 1
    return Batch(systems, seg, rows, cols, vals, rhs, tets, np.concatenate(parts).astype(np.int32))
