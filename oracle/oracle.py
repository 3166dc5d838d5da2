"""ctypes wrapper of the CPU oracle (oracle.c).  TEST INFRASTRUCTURE ONLY:
imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the checker, never by the product package."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "_build", "liboracle.so")

i32p = C.POINTER(C.c_int32)
u32p = C.POINTER(C.c_uint32)
f64p = C.POINTER(C.c_double)


class OrPrecond(C.Structure):
    _fields_ = [("kind", C.c_int32), ("blk_off", C.c_int32), ("blk_cnt", C.c_int32), ("data", f64p),
                ("mas", C.c_void_p), ("mas_fp32", C.c_int32), ("mas_active_levels", C.c_int32)]


class OrPcgResult(C.Structure):
    _fields_ = [("iters", C.c_int64), ("iters_exact", C.c_int64), ("rz0", C.c_double), ("rz_final", C.c_double),
                ("rel_res", C.c_double), ("rel_res_true", C.c_double), ("converged", C.c_int32),
                ("bypassed", C.c_int32), ("seconds", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def build() -> str:
    # `ref` builds oracle/_ref/ only where the reference sources are present
    r = subprocess.run(["make", "-s", "-j8", "-C", HERE, "all", "ref"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    return SO


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(SO):
            build()
        _lib = _bind(C.CDLL(SO))
    return _lib


def mutant(name: str) -> C.CDLL:
    """A planted-mistake build of the oracle (-DOR_MUTANT_<name>, see
    oracle.c), for tests that show a check detects the mistake."""
    so = os.path.join(HERE, "_build", f"liboracle_mutant_{name}.so")
    r = subprocess.run(["make", "-s", "-C", HERE, so], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"mutant build failed:\n{r.stdout}\n{r.stderr}")
    return _bind(C.CDLL(so))


def _bind(L: C.CDLL) -> C.CDLL:
    L.or_key_bits.restype = C.c_int
    L.or_key_bits.argtypes = [C.c_uint64]
    L.or_convert_sym.restype = C.c_int64
    L.or_convert_sym.argtypes = [C.c_int32, C.c_int64, i32p, i32p, f64p, i32p, i32p, f64p, i32p]
    L.or_spmv_sym.restype = C.c_double
    L.or_spmv_sym.argtypes = [C.c_int32, C.c_int64, i32p, i32p, f64p, f64p, f64p]
    L.or_dot.restype = C.c_double
    L.or_dot.argtypes = [f64p, f64p, C.c_int64]
    L.or_norm.restype = C.c_double
    L.or_norm.argtypes = [f64p, C.c_int64]
    L.or_diag3_assemble.restype = C.c_int
    L.or_diag3_assemble.argtypes = [C.c_int64, i32p, i32p, f64p, C.c_int32, C.c_int32, f64p]
    L.or_diag3_apply.restype = None
    L.or_diag3_apply.argtypes = [C.c_int32, f64p, f64p, f64p]
    L.or_diag12_assemble.restype = C.c_int
    L.or_diag12_assemble.argtypes = [C.c_int64, i32p, i32p, f64p, C.c_int32, C.c_int32, f64p]
    L.or_diag12_apply.restype = None
    L.or_diag12_apply.argtypes = [C.c_int32, f64p, f64p, f64p]
    L.or_mas_create.restype = C.c_void_p
    L.or_mas_create.argtypes = [C.c_int32, i32p, i32p, C.c_int64, i32p, C.c_int64, i32p, C.c_int64]
    L.or_mas_destroy.restype = None
    L.or_mas_destroy.argtypes = [C.c_void_p]
    L.or_mas_info.restype = None
    L.or_mas_info.argtypes = [C.c_void_p, i32p]
    L.or_mas_tables.restype = None
    L.or_mas_tables.argtypes = [C.c_void_p, i32p, i32p, i32p, i32p, i32p, u32p, i32p]
    L.or_mas_going_next_size.restype = C.c_int64
    L.or_mas_going_next_size.argtypes = [C.c_void_p]
    L.or_mas_assemble.restype = C.c_int
    L.or_mas_assemble.argtypes = [C.c_void_p, C.c_int64, i32p, i32p, f64p, C.c_int32, C.c_int]
    L.or_mas_num_clusters.restype = C.c_int64
    L.or_mas_num_clusters.argtypes = [C.c_void_p]
    L.or_mas_cluster_hessians.restype = None
    L.or_mas_cluster_hessians.argtypes = [C.c_void_p, f64p]
    L.or_mas_cluster_inverses.restype = None
    L.or_mas_cluster_inverses.argtypes = [C.c_void_p, f64p]
    L.or_mas_apply.restype = None
    L.or_mas_apply.argtypes = [C.c_void_p, f64p, f64p, C.c_int, C.c_int]
    L.or_pcg.restype = C.c_int
    L.or_pcg.argtypes = [C.c_int32, C.c_int64, i32p, i32p, f64p, C.POINTER(OrPrecond), C.c_int32, f64p, f64p,
                         C.c_int32, C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_int32, C.c_int64,
                         C.POINTER(OrPcgResult)]
    return L


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def key_bits(max_key: int) -> int:
    return lib().or_key_bits(max_key)


def convert_sym(B, rows, cols, vals):
    rows, cols, vals = _i(rows), _i(cols), _d(vals).reshape(-1, 9)
    T = rows.size
    orow = np.empty(max(T, 1), np.int32)
    ocol = np.empty(max(T, 1), np.int32)
    oval = np.empty((max(T, 1), 9), np.float64)
    rp = np.empty(B + 1, np.int32)
    U = lib().or_convert_sym(B, T, rows.ctypes.data_as(i32p), cols.ctypes.data_as(i32p), vals.ctypes.data_as(f64p),
                             orow.ctypes.data_as(i32p), ocol.ctypes.data_as(i32p), oval.ctypes.data_as(f64p),
                             rp.ctypes.data_as(i32p))
    if U < 0:
        raise ValueError("or_convert_sym failed")
    return orow[:U].copy(), ocol[:U].copy(), oval[:U].copy(), rp


def spmv_sym(B, rows, cols, vals, x):
    rows, cols, vals, x = _i(rows), _i(cols), _d(vals).reshape(-1, 9), _d(x)
    y = np.empty(3 * B, np.float64)
    d = lib().or_spmv_sym(B, rows.size, rows.ctypes.data_as(i32p), cols.ctypes.data_as(i32p),
                          vals.ctypes.data_as(f64p), x.ctypes.data_as(f64p), y.ctypes.data_as(f64p))
    return y, d


def dot(x, y):
    x, y = _d(x), _d(y)
    return lib().or_dot(x.ctypes.data_as(f64p), y.ctypes.data_as(f64p), x.size)


def norm(x):
    x = _d(x)
    return lib().or_norm(x.ctypes.data_as(f64p), x.size)


def diag3(rows, cols, vals, blk_off, blk_cnt):
    rows, cols, vals = _i(rows), _i(cols), _d(vals).reshape(-1, 9)
    D = np.empty((blk_cnt, 9), np.float64)
    lib().or_diag3_assemble(rows.size, rows.ctypes.data_as(i32p), cols.ctypes.data_as(i32p),
                            vals.ctypes.data_as(f64p), blk_off, blk_cnt, D.ctypes.data_as(f64p))
    return D


def diag3_apply(D, r):
    D, r = _d(D), _d(r)
    z = np.empty_like(r)
    lib().or_diag3_apply(D.shape[0], D.ctypes.data_as(f64p), r.ctypes.data_as(f64p), z.ctypes.data_as(f64p))
    return z


def diag12(rows, cols, vals, blk_off, bodies):
    rows, cols, vals = _i(rows), _i(cols), _d(vals).reshape(-1, 9)
    H = np.empty((bodies, 144), np.float64)
    rc = lib().or_diag12_assemble(rows.size, rows.ctypes.data_as(i32p), cols.ctypes.data_as(i32p),
                                  vals.ctypes.data_as(f64p), blk_off, bodies, H.ctypes.data_as(f64p))
    if rc != 0:
        raise ValueError("singular body Hessian")
    return H


def diag12_apply(H, r):
    H, r = _d(H), _d(r)
    z = np.empty_like(r)
    lib().or_diag12_apply(H.shape[0], H.ctypes.data_as(f64p), r.ctypes.data_as(f64p), z.ctypes.data_as(f64p))
    return z


class Mas:
    """MASPreconditionerEngine restated on the CPU."""

    def __init__(self, vert_num, part_ids, tets=None, tris=None, edges=None, *, library=None):
        self._keep = []
        self._L = library if library is not None else lib()

        def arr(a, k):
            if a is None:
                return None, 0
            a = _i(a).reshape(-1)
            self._keep.append(a)
            return a.ctypes.data_as(i32p), a.size // k

        pp, _ = arr(part_ids, 1)
        tp, nt = arr(tets, 4)
        trp, ntr = arr(tris, 3)
        ep, ne = arr(edges, 2)
        self.h = self._L.or_mas_create(vert_num, pp, tp, nt, trp, ntr, ep, ne)
        if not self.h:
            raise ValueError("or_mas_create failed (partition larger than 16?)")
        info = np.empty(5, np.int32)
        self._L.or_mas_info(self.h, info.ctypes.data_as(i32p))
        self.V, self.N, self.L, self.total, self.level1_begin = (int(v) for v in info)

    def __del__(self):
        try:
            if self.h:
                self._L.or_mas_destroy(self.h)
        except Exception:
            pass

    def tables(self):
        gsz = self._L.or_mas_going_next_size(self.h)
        p2r = np.empty(self.N, np.int32)
        r2p = np.empty(self.V, np.int32)
        gn = np.empty(gsz, np.int32)
        ls = np.empty(2 * (self.L + 1), np.int32)
        ct = np.empty(self.V * 6, np.int32)
        fc = np.empty(self.V, np.uint32)
        po = np.empty(max(self.N // 16, 1), np.int32)
        self._L.or_mas_tables(self.h, p2r.ctypes.data_as(i32p), r2p.ctypes.data_as(i32p), gn.ctypes.data_as(i32p),
                            ls.ctypes.data_as(i32p), ct.ctypes.data_as(i32p), fc.ctypes.data_as(u32p),
                            po.ctypes.data_as(i32p))
        return {"part_to_real": p2r, "real_to_part": r2p, "going_next": gn, "level_sizes": ls.reshape(-1, 2),
                "coarse_tables": ct.reshape(self.V, 6), "fine_connect": fc, "prefix_orig": po[:self.N // 16]}

    def assemble(self, rows, cols, vals, blk_off=0, inv_fp32=False):
        rows, cols, vals = _i(rows), _i(cols), _d(vals).reshape(-1, 9)
        self._A = (rows, cols, vals)
        self._L.or_mas_assemble(self.h, rows.size, rows.ctypes.data_as(i32p), cols.ctypes.data_as(i32p),
                              vals.ctypes.data_as(f64p), blk_off, 1 if inv_fp32 else 0)

    def num_clusters(self):
        return self._L.or_mas_num_clusters(self.h)

    def hessians(self):
        C_ = self.num_clusters()
        out = np.empty((C_, 136, 9), np.float64)
        self._L.or_mas_cluster_hessians(self.h, out.ctypes.data_as(f64p))
        return out

    def inverses(self):
        C_ = self.num_clusters()
        out = np.empty((C_, 136, 9), np.float64)
        self._L.or_mas_cluster_inverses(self.h, out.ctypes.data_as(f64p))
        return out

    def apply(self, r, fp32=False, active_levels=0):
        r = _d(r)
        z = np.empty_like(r)
        self._L.or_mas_apply(self.h, r.ctypes.data_as(f64p), z.ctypes.data_as(f64p), 1 if fp32 else 0, active_levels)
        return z


def pcg(B, rows, cols, vals, b, precond=None, *, criterion=0, tol=1e-3, max_iter_ratio=2.0, check_interval=5,
        psd_fallback=1, threads=1, max_iter_cap=0):
    """precond: list of dicts {kind, blk_off, blk_cnt, data (ndarray) | mas (Mas), mas_fp32, active_levels}"""
    rows, cols, vals, b = _i(rows), _i(cols), _d(vals).reshape(-1, 9), _d(b)
    precond = precond or []
    arr = (OrPrecond * max(1, len(precond)))()
    keep = []
    for k, p in enumerate(precond):
        data = p.get("data")
        if data is not None:
            data = _d(data)
            keep.append(data)
        arr[k].kind = p["kind"]
        arr[k].blk_off = p.get("blk_off", 0)
        arr[k].blk_cnt = p["blk_cnt"]
        arr[k].data = data.ctypes.data_as(f64p) if data is not None else None
        arr[k].mas = p["mas"].h if p.get("mas") is not None else None
        arr[k].mas_fp32 = 1 if p.get("mas_fp32") else 0
        arr[k].mas_active_levels = p.get("active_levels", 0)
    x = np.zeros(3 * B, np.float64)
    res = OrPcgResult()
    lib().or_pcg(B, rows.size, rows.ctypes.data_as(i32p), cols.ctypes.data_as(i32p), vals.ctypes.data_as(f64p),
                 arr, len(precond), b.ctypes.data_as(f64p), x.ctypes.data_as(f64p), criterion, tol, max_iter_ratio,
                 check_interval, psd_fallback, threads, max_iter_cap, C.byref(res))
    return x, res.as_dict()


# ---- the reference's own partition (oracle/_ref/libref_partition.so: the
# reference's METIS sources compiled in place, mesh_partition's control flow)
REF_PART_SO = os.path.join(HERE, "_ref", "libref_partition.so")
_ref_part = None


def ref_partition_available() -> bool:
    return os.path.exists(REF_PART_SO)


def ref_mesh_partition(V: int, part_max_size: int = 16, *, tets=None, tris=None, edges=None) -> np.ndarray:
    """geometry::mesh_partition (src/geometry/mesh_partition.cpp:77-230) as
    the reference computes it, METIS included."""
    global _ref_part
    if _ref_part is None:
        L = C.CDLL(REF_PART_SO)
        L.ref_mesh_partition.restype = C.c_int
        L.ref_mesh_partition.argtypes = [C.c_int32, i32p, C.c_int64, i32p, C.c_int64, i32p, C.c_int64, C.c_int32,
                                         i32p]
        _ref_part = L
    keep = []

    def arr(a, k):
        if a is None:
            return None, 0
        a = np.ascontiguousarray(a, dtype=np.int32).reshape(-1)
        keep.append(a)
        return a.ctypes.data_as(i32p), a.size // k

    tp, nt = arr(tets, 4)
    rp, nr = arr(tris, 3)
    ep, ne = arr(edges, 2)
    out = np.empty(max(V, 1), np.int32)
    rc = _ref_part.ref_mesh_partition(V, tp, nt, rp, nr, ep, ne, part_max_size, out.ctypes.data_as(i32p))
    if rc != 0:
        raise ValueError("mesh_partition: part_max_size must be greater than zero")
    return out[:V]
