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
    if lib.mi_synth_fem_fill(C.byref(d), None, None, None, None, tets.ctypes.data_as(abi.c_i32p), None,
                             x.ctypes.data_as(abi.c_f64p)) != 0:
        raise RuntimeError(lib.mi_synth_last_error().decode())
    if lib.mi_synth_fem_aux(C.byref(d), X.ctypes.data_as(abi.c_f64p), m.ctypes.data_as(abi.c_f64p),
                            ci.ctypes.data_as(abi.c_i32p), cg.ctypes.data_as(abi.c_f64p),
                            ck.ctypes.data_as(abi.c_f64p)) != 0:
        raise RuntimeError(lib.mi_synth_last_error().decode())
    E, nu = d.youngs, d.poisson
    mu = E / (2.0 * (1.0 + nu))
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    return ProducerInputs(X, x, m, tets, ci[:n_contacts], cg[:n_contacts], ck[:n_contacts], mu, lam, d.dt)
