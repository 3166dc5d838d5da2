"""Thin host-side helpers over the C ABI: a device context, device arrays that
round-trip numpy arrays, and a LinearSystem object whose methods mirror the
reference's GlobalLinearSystem / LinearFusedPCG / LocalPreconditioner hooks.
Every operation runs in libmi_linsys.so; nothing here computes."""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

from . import abi
from .abi import check


class Context:
    def __init__(self, device: int = 0):
        self.lib = abi.linsys()
        h = C.c_void_p()
        check(self.lib.mi_ctx_create(device, C.byref(h)))
        self.h = h

    def stream(self) -> C.c_void_p:
        s = C.c_void_p()
        check(self.lib.mi_ctx_stream(self.h, C.byref(s)))
        return s

    def synchronize(self) -> None:
        check(self.lib.mi_ctx_synchronize(self.h))

    def close(self) -> None:
        if self.h:
            check(self.lib.mi_ctx_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # device memory
    def alloc(self, shape, dtype) -> "DeviceArray":
        return DeviceArray(self, shape, dtype)

    def to_device(self, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = DeviceArray(self, a.shape, a.dtype)
        d.upload(a)
        return d

    def dot(self, x: "DeviceArray", y: "DeviceArray") -> float:
        if x.size != y.size:
            raise ValueError("dot: size mismatch")  # LinearSystemContext::dot throws std::invalid_argument
        if x.dtype == np.float32:
            out = C.c_float()
            check(self.lib.mi_dot_f32(self.h, x.ptr, y.ptr, x.size, C.byref(out)))
        else:
            out = C.c_double()
            check(self.lib.mi_dot_f64(self.h, x.ptr, y.ptr, x.size, C.byref(out)))
        return out.value

    def norm(self, x: "DeviceArray") -> float:
        if x.dtype == np.float32:
            out = C.c_float()
            check(self.lib.mi_norm_f32(self.h, x.ptr, x.size, C.byref(out)))
        else:
            out = C.c_double()
            check(self.lib.mi_norm_f64(self.h, x.ptr, x.size, C.byref(out)))
        return out.value

    # utils/matrix_market.h:120-158 (export_vector_market)
    def write_vector_mtx(self, path: str, x: "DeviceArray") -> None:
        check(self.lib.mi_vec_write_mtx(self.h, os.fsencode(path), x.ptr, x.size))

    # geometry::mesh_partition (src/geometry/mesh_partition.cpp:77-230), on the device
    def mesh_partition(self, vert_num: int, part_max_size: int = 16, *, tets=None, tris=None,
                       edges=None) -> tuple[np.ndarray, int]:
        keep = []

        def dev(a, k):
            if a is None:
                return None, 0
            a = np.ascontiguousarray(a, dtype=np.int32).reshape(-1)
            d = self.to_device(a) if a.size else None
            keep.append(d)
            return (d.ptr if d is not None else None), a.size // k

        tp, nt = dev(tets, 4)
        rp, nr = dev(tris, 3)
        ep, ne = dev(edges, 2)
        out = self.alloc(max(vert_num, 1), np.int32)
        n = C.c_int32(0)
        try:
            check(self.lib.mi_mesh_partition(self.h, vert_num, tp, nt, rp, nr, ep, ne, part_max_size, out.ptr,
                                             C.byref(n), None))
            return out.numpy()[:vert_num], int(n.value)
        finally:
            out.free()
            for d in keep:
                if d is not None:
                    d.free()


class DeviceArray:
    def __init__(self, ctx: Context, shape, dtype, borrow: int | None = None):
        """a device buffer of this shape; `borrow` = the address of memory the
        library owns (not freed here) instead of a new allocation"""
        self.ctx = ctx
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.size = int(np.prod(self.shape)) if self.shape else 1
        self.nbytes = self.size * self.dtype.itemsize
        self.owned = borrow is None
        if borrow is not None:
            self.ptr = C.c_void_p(borrow)
            return
        p = C.c_void_p()
        check(ctx.lib.mi_malloc(ctx.h, C.byref(p), max(1, self.nbytes)))
        self.ptr = p

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.size
        check(self.ctx.lib.mi_memcpy_h2d(self.ctx.h, self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes))

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        check(self.ctx.lib.mi_memcpy_d2h(self.ctx.h, out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes))
        return out

    def zero(self) -> None:
        check(self.ctx.lib.mi_memset(self.ctx.h, self.ptr, 0, self.nbytes))

    def free(self) -> None:
        if self.ptr and self.owned:
            check(self.ctx.lib.mi_free(self.ctx.h, self.ptr))
        self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.owned and self.ctx.h:
                self.ctx.lib.mi_free(self.ctx.h, self.ptr)
        except Exception:
            pass


def _i32p(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32).ctypes.data_as(abi.c_i32p)


def spmv_shapes(lib=None) -> dict:
    """the SpMV's launch shapes: {"default": k, "shapes": [{threads, rows, per}, ...]}"""
    lib = lib or abi.linsys()
    n, d = C.c_int32(), C.c_int32()
    check(lib.mi_spmv_shapes(C.byref(n), C.byref(d), None, None, None, None))
    t, r, p, v = ((C.c_int32 * n.value)() for _ in range(4))
    check(lib.mi_spmv_shapes(C.byref(n), C.byref(d), t, r, p, v))
    return {"default": d.value,
            "shapes": [{"threads": t[k], "rows": r[k], "per": p[k], "vec": v[k]} for k in range(n.value)]}


class LinearSystem:
    """GlobalLinearSystem-shaped handle over one mi_sys."""

    def __init__(self, ctx: Context, block_rows: int, precision: int = abi.MI_FP64, reserve_ratio: float = 1.1):
        self.ctx = ctx
        self.lib = ctx.lib
        self.B = int(block_rows)
        self.n = 3 * self.B
        self.precision = int(precision)
        d = abi.SysDesc(self.B, reserve_ratio, precision)
        h = C.c_void_p()
        check(self.lib.mi_sys_create(ctx.h, C.byref(d), C.byref(h)))
        self.h = h
        self._keep = []

    def close(self):
        if self.h:
            check(self.lib.mi_sys_destroy(self.h))
            self.h = None

    # MatrixConverter::convert_sym
    def assemble(self, rows: DeviceArray, cols: DeviceArray, vals: DeviceArray, stream=None) -> None:
        """mi_sys_assemble, or mi_sys_assemble_f32 for float32 values"""
        T = rows.size
        fn = self.lib.mi_sys_assemble_f32 if vals.dtype == np.float32 else self.lib.mi_sys_assemble
        check(fn(self.h, T, rows.ptr, cols.ptr, vals.ptr, stream))

    def nnz(self) -> int:
        v = C.c_int64()
        check(self.lib.mi_sys_nnz(self.h, C.byref(v)))
        return v.value

    def export_bcoo(self):
        U = self.nnz()
        r = self.ctx.alloc(max(U, 1), np.int32)
        c = self.ctx.alloc(max(U, 1), np.int32)
        v = self.ctx.alloc((max(U, 1), 9), np.float64)
        p = self.ctx.alloc(self.B + 1, np.int32)
        check(self.lib.mi_sys_export_bcoo(self.h, r.ptr, c.ptr, v.ptr, p.ptr, None))
        self.ctx.synchronize()
        return r.numpy()[:U], c.numpy()[:U], v.numpy()[:U], p.numpy()

    # extras/debug/dump_linear_system (global_linear_system.cu:42-70)
    def write_mtx(self, path: str, one_based: bool = True) -> None:
        check(self.lib.mi_sys_write_mtx(self.h, os.fsencode(path), int(one_based)))

    # Spmv::rbk_sym_spmv_dot
    def spmv_dot(self, x: DeviceArray, y: DeviceArray, accumulate: bool = False) -> float:
        """y = A x (y += A x when accumulate); returns x^T A x"""
        if not accumulate:
            y.zero()
        d = self.ctx.alloc(1, np.float64)
        check(self.lib.mi_spmv_dot(self.h, x.ptr, y.ptr, d.ptr, None))
        return float(d.numpy()[0])

    # GlobalLinearSystem::Impl::_update_subsystem_extent / _assemble_linear_system
    def update_extents(self, extents: list[dict]) -> tuple[list[dict], int, int]:
        """extents: dicts with is_diag, dof_count, block_count, rl_block_count,
        l, r; returns (views, total triplets, triplet capacity)"""
        n = len(extents)
        ext = (abi.Extent * max(1, n))()
        for k, e in enumerate(extents):
            ext[k] = abi.Extent(int(e.get("is_diag", 1)), int(e.get("dof_count", 0)), int(e.get("block_count", 0)),
                                int(e.get("rl_block_count", 0)), int(e.get("l", -1)), int(e.get("r", -1)))
        views = (abi.ExtentView * max(1, n))()
        tot, cap = C.c_int64(), C.c_int64()
        check(self.lib.mi_sys_update_extents(self.h, n, ext, views, C.byref(tot), C.byref(cap)))
        out = [{f: getattr(views[k], f) for f, _ in abi.ExtentView._fields_} for k in range(n)]
        self.B = sum(int(e.get("dof_count", 0)) for e in extents if e.get("is_diag", 1)) // 3
        self.n = 3 * self.B
        return out, tot.value, cap.value

    def begin_assembly(self, stream=None) -> dict:
        """fills the owned triplet rows/cols with -1 and b with 0; returns the
        device addresses of rows, cols, vals and b"""
        ptrs = [C.c_void_p() for _ in range(4)]
        check(self.lib.mi_sys_begin_assembly(self.h, *(C.byref(p) for p in ptrs), stream))
        return dict(zip(("rows", "cols", "vals", "b"), (p.value for p in ptrs)))

    def assemble_owned(self, stream=None) -> None:
        check(self.lib.mi_sys_assemble_owned(self.h, stream))

    def set_spmv_shape(self, shape: int, stream=None) -> None:
        """the window kernel's launch shape (an assembled layout is rebuilt)"""
        check(self.lib.mi_sys_set_spmv_shape(self.h, int(shape), stream))

    def spmv_layout(self) -> dict:
        """the layout the SpMV streams: window (1) or full BSR (0), window rows,
        workgroup size and blocks per product"""
        lay, wr, nt, nb, sh = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int64(), C.c_int32()
        check(self.lib.mi_sys_spmv_layout(self.h, C.byref(lay), C.byref(wr), C.byref(nt), C.byref(nb), C.byref(sh)))
        out = {"layout": "window" if lay.value == 1 else "full", "window_rows": wr.value, "threads": nt.value,
               "blocks": nb.value, "shape": sh.value}
        if sh.value >= 0:
            info = spmv_shapes(self.lib)["shapes"][sh.value]
            out["per"], out["vec"] = info["per"], info["vec"]
        return out

    def spmv(self, a: float, x: DeviceArray, b: float, y: DeviceArray) -> None:
        check(self.lib.mi_spmv(self.h, a, x.ptr, b, y.ptr, None))

    def spmv_order(self) -> np.ndarray:
        """mi_sys_spmv_order: the row position of every block row"""
        pos = np.empty(max(self.B, 1), np.int32)
        check(self.lib.mi_sys_spmv_order(self.h, pos.ctypes.data_as(C.c_void_p)))
        return pos[: self.B]

    def spmv_ordered(self, x: DeviceArray, y: DeviceArray, d_dot: DeviceArray | None = None, stream=None) -> None:
        """mi_spmv_ordered: y = A x (and x^T A x) on vectors in the SpMV's row order"""
        check(self.lib.mi_spmv_ordered(self.h, x.ptr, y.ptr, d_dot.ptr if d_dot is not None else None, stream))

    def add_preconditioner(self, kind: int, dof_offset: int, dof_count: int, *, part_ids=None, tets=None,
                           tris=None, edges=None, mas_precision: int = abi.MI_FP64, active_levels: int = 0,
                           build: int = abi.MI_MAS_BUILD_HOST, device_arrays: dict | None = None,
                           reinit: int | None = None, invert: int = abi.MI_MAS_INVERT_MFMA) -> int:
        """`device_arrays` (MI_MAS_BUILD_DEVICE only) maps "part_ids", "tets",
        "tris", "edges" to device pointers read instead of the host arrays;
        their counts still come from the host arrays' shapes.  `reinit` =
        a handle rebuilds that preconditioner in place (mi_precond_reinit).
        `invert` picks the cluster inverse kernel (mi_mas_invert)."""
        keep = []

        def arr(a, k):
            if a is None:
                return None, 0
            a = np.ascontiguousarray(a, dtype=np.int32).reshape(-1)
            keep.append(a)
            return a.ctypes.data_as(abi.c_i32p), a.size // k

        pp, _ = arr(part_ids, 1)
        tp, nt = arr(tets, 4)
        trp, ntr = arr(tris, 3)
        ep, ne = arr(edges, 2)
        dev = {k: (v.value if isinstance(v, C.c_void_p) else v) for k, v in (device_arrays or {}).items()}
        d = abi.PrecondDesc(kind, dof_offset, dof_count, dof_count // 3, pp, tp, nt, trp, ntr, ep, ne,
                            mas_precision, active_levels, build, dev.get("part_ids"), dev.get("tets"),
                            dev.get("tris"), dev.get("edges"), invert)
        if reinit is not None:
            check(self.lib.mi_precond_reinit(self.h, reinit, C.byref(d)))
            return reinit
        h = C.c_int()
        check(self.lib.mi_precond_add(self.h, C.byref(d), C.byref(h)))
        return h.value

    # MASPreconditionerEngine::set_active_level_num
    def mas_set_active_levels(self, handle: int, levels: int) -> None:
        check(self.lib.mi_precond_mas_set_active_levels(self.h, handle, levels))

    def mas_set_system_levels(self, handle: int, seg_rows, levels) -> None:
        """mi_precond_mas_set_system_levels: levels[g] for the batched system
        on block rows seg_rows[g] .. seg_rows[g + 1]; empty lists clear"""
        seg = np.ascontiguousarray(seg_rows, dtype=np.int32)
        lv = np.ascontiguousarray(levels, dtype=np.int32)
        nseg = max(0, seg.size - 1) if lv.size else 0
        if nseg and lv.size != nseg:
            raise ValueError("one level count per system")
        check(self.lib.mi_precond_mas_set_system_levels(self.h, handle, nseg, seg.ctypes.data_as(C.c_void_p),
                                                        lv.ctypes.data_as(C.c_void_p)))

    def mas_export(self, handle: int, matrices: bool = True) -> dict:
        info = np.zeros(6, np.int32)
        check(self.lib.mi_precond_mas_export(self.h, handle, info.ctypes.data_as(abi.c_i32p), None, None, None, None,
                                             None, None, None))
        V, N, L, total, l1b, Cn = (int(v) for v in info)
        padded = (max(N, V) + 15) // 16 * 16
        out = {"V": V, "N": N, "L": L, "total": total, "level1_begin": l1b, "clusters": Cn,
               "part_to_real": np.zeros(max(N, 1), np.int32), "real_to_part": np.zeros(max(V, 1), np.int32),
               "going_next": np.zeros(max(padded * L, 1), np.int32), "level_sizes": np.zeros(2 * (L + 1), np.int32),
               "coarse_tables": np.zeros(max(6 * V, 1), np.int32)}
        hess = inv = None
        if matrices:
            hess = np.zeros((max(Cn, 1), 136, 9))
            inv = np.zeros((max(Cn, 1), 136, 9))
        p = lambda a: None if a is None else a.ctypes.data_as(abi.c_f64p)  # noqa: E731
        q = lambda a: a.ctypes.data_as(abi.c_i32p)  # noqa: E731
        check(self.lib.mi_precond_mas_export(self.h, handle, None, q(out["part_to_real"]), q(out["real_to_part"]),
                                             q(out["going_next"]), q(out["level_sizes"]), q(out["coarse_tables"]),
                                             p(hess), p(inv)))
        out["part_to_real"] = out["part_to_real"][:N]
        out["level_sizes"] = out["level_sizes"].reshape(-1, 2)
        out["coarse_tables"] = out["coarse_tables"].reshape(V, 6)
        if matrices:
            out["hess"], out["inv"] = hess[:Cn], inv[:Cn]
        return out

    # MASPreconditionerEngine::dump_cluster_matrices_debug
    def mas_dump(self, handle: int, output_dir: str, frame: int = 0, newton_iter: int = 0) -> None:
        check(self.lib.mi_precond_mas_dump(self.h, handle, os.fsencode(output_dir), frame, newton_iter))

    def assemble_preconditioners(self) -> None:
        check(self.lib.mi_precond_assemble(self.h, None))

    def apply_preconditioner(self, r: DeviceArray, z: DeviceArray) -> None:
        check(self.lib.mi_precond_apply(self.h, r.ptr, z.ptr, None, None))

    def solve(self, b: DeviceArray, x: DeviceArray, **kw) -> dict:
        o = abi.PcgOpts()
        check(self.lib.mi_pcg_default_opts(C.byref(o)))
        for k, v in kw.items():
            setattr(o, k, v)
        st = abi.PcgStats()
        check(self.lib.mi_pcg_solve(self.h, b.ptr, x.ptr, C.byref(o), C.byref(st), None))
        return st.as_dict()

    def solve_batched(self, seg_rows, b: DeviceArray, x: DeviceArray, **kw) -> list[dict]:
        """mi_pcg_solve_batched: independent systems on the block-row ranges
        seg_rows[k] .. seg_rows[k + 1]; one stats dict per system"""
        o = abi.PcgOpts()
        check(self.lib.mi_pcg_default_opts(C.byref(o)))
        for k, v in kw.items():
            setattr(o, k, v)
        seg = np.ascontiguousarray(seg_rows, dtype=np.int32)
        nseg = seg.size - 1
        st = (abi.PcgStats * max(nseg, 1))()
        check(self.lib.mi_pcg_solve_batched(self.h, nseg, seg.ctypes.data_as(abi.c_i32p), b.ptr, x.ptr, C.byref(o),
                                            st, None))
        return [st[k].as_dict() for k in range(nseg)]


def read_mtx_block_triplets(path: str, one_based: bool = True):
    """A coordinate Matrix Market file as 3x3 block triplets (duplicates summed,
    ascending by (block row, block col)): (block_rows, rows, cols, vals[T, 9]
    column-major)."""
    lib = abi.linsys()
    B, T = C.c_int32(), C.c_int64()
    fp = os.fsencode(path)
    check(lib.mi_mtx_read_block_triplets(fp, int(one_based), C.byref(B), C.byref(T), None, None, None))
    rows = np.empty(T.value, np.int32)
    cols = np.empty(T.value, np.int32)
    vals = np.empty((T.value, 9), np.float64)
    check(lib.mi_mtx_read_block_triplets(fp, int(one_based), C.byref(B), C.byref(T), _i32p(rows), _i32p(cols),
                                         vals.ctypes.data_as(abi.c_f64p)))
    return B.value, rows, cols, vals


def read_vector_mtx(path: str) -> np.ndarray:
    lib = abi.linsys()
    n = C.c_int64()
    fp = os.fsencode(path)
    check(lib.mi_vec_read_mtx(fp, C.byref(n), None))
    x = np.empty(n.value, np.float64)
    check(lib.mi_vec_read_mtx(fp, C.byref(n), x.ctypes.data_as(abi.c_f64p)))
    return x


def rel_err(a: np.ndarray, b: np.ndarray) -> float:
    nb = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b)) / (nb if nb > 0 else 1.0)


__all__ = ["Context", "DeviceArray", "LinearSystem", "read_mtx_block_triplets", "read_vector_mtx", "rel_err"]
