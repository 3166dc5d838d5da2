"""ctypes binding of the C ABI in include/mi_linsys.h and include/mi_synth.h.

Loading fails loudly when the native library is missing: there is no Python or
CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LINSYS_SO = os.path.join(PKG, "libmi_linsys.so")
SYNTH_SO = os.path.join(PKG, "libmi_synth.so")

MI_OK = 0
MI_ERR_ARG = -1
MI_ERR_HIP = -2
MI_ERR_STATE = -3
MI_ERR_NOMEM = -4
MI_ERR_UNSUPPORTED = -5
MI_ERR_NONFINITE = -6

MI_FP64 = 0
MI_FP32 = 1

MI_PRECOND_DIAG3 = 1
MI_PRECOND_DIAG12 = 2
MI_PRECOND_MAS = 3

MI_MAS_BUILD_HOST = 0
MI_MAS_BUILD_DEVICE = 1
MI_MAS_INVERT_MFMA = 0
MI_MAS_INVERT_SWEEP = 1

MI_CRIT_RZ_RATIO = 0
MI_CRIT_REL_RESIDUAL = 1

c_i32p = C.POINTER(C.c_int32)
c_u32p = C.POINTER(C.c_uint32)
c_f64p = C.POINTER(C.c_double)


class SysDesc(C.Structure):
    _fields_ = [("block_rows", C.c_int32), ("reserve_ratio", C.c_double), ("value_precision", C.c_int32)]


class Extent(C.Structure):
    _fields_ = [("is_diag", C.c_int32), ("dof_count", C.c_int64), ("block_count", C.c_int64),
                ("rl_block_count", C.c_int64), ("l", C.c_int32), ("r", C.c_int32)]


class ExtentView(C.Structure):
    _fields_ = [("dof_offset", C.c_int64), ("triplet_offset", C.c_int64), ("rl_triplet_offset", C.c_int64),
                ("row_offset", C.c_int32), ("col_offset", C.c_int32)]


class PrecondDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("dof_offset", C.c_int64),
        ("dof_count", C.c_int64),
        ("vert_num", C.c_int32),
        ("part_ids_host", c_i32p),
        ("tets_host", c_i32p),
        ("n_tets", C.c_int64),
        ("tris_host", c_i32p),
        ("n_tris", C.c_int64),
        ("edges_host", c_i32p),
        ("n_edges", C.c_int64),
        ("mas_precision", C.c_int32),
        ("mas_active_levels", C.c_int32),
        ("mas_build", C.c_int32),
        ("part_ids_dev", C.c_void_p),
        ("tets_dev", C.c_void_p),
        ("tris_dev", C.c_void_p),
        ("edges_dev", C.c_void_p),
        ("mas_invert", C.c_int32),
    ]


MI_SPMV_SHAPE_AUTO, MI_SPMV_SHAPE_FULL = -1, -2


class PcgOpts(C.Structure):
    _fields_ = [
        ("criterion", C.c_int32),
        ("tol", C.c_double),
        ("max_iter_ratio", C.c_double),
        ("check_interval", C.c_int32),
        ("use_graph", C.c_int32),
        ("psd_fallback", C.c_int32),
        ("force_bypass", C.c_int32),
        ("true_residual", C.c_int32),
        ("graph_checks", C.c_int32),
        ("x_init", C.c_int32),
        ("keep_order", C.c_int32),
    ]


class PcgStats(C.Structure):
    _fields_ = [
        ("iters", C.c_int64),
        ("iters_exact", C.c_int64),
        ("rz0", C.c_double),
        ("rz_final", C.c_double),
        ("rel_res", C.c_double),
        ("rel_res_true", C.c_double),
        ("converged", C.c_int32),
        ("bypassed", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class SynthFemDesc(C.Structure):
    _fields_ = [
        ("grid", C.c_int32),
        ("jitter", C.c_double),
        ("seed", C.c_uint64),
        ("youngs", C.c_double),
        ("poisson", C.c_double),
        ("density", C.c_double),
        ("dt", C.c_double),
        ("n_contacts", C.c_int64),
        ("contact_cells", C.c_int32),
        ("kappa_lo", C.c_double),
        ("kappa_hi", C.c_double),
        ("threads", C.c_int32),
    ]


class SynthFemSizes(C.Structure):
    _fields_ = [("block_rows", C.c_int32), ("n_tets", C.c_int64), ("n_triplets", C.c_int64), ("n_parts", C.c_int32)]


# every symbol include/mi_linsys.h declares, with its prototype
LINSYS_PROTOS = {
    "mi_last_error": (C.c_char_p, []),
    "mi_version": (C.c_int, []),
    "mi_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "mi_ctx_destroy": (C.c_int, [C.c_void_p]),
    "mi_ctx_synchronize": (C.c_int, [C.c_void_p]),
    "mi_ctx_stream": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "mi_malloc": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_size_t]),
    "mi_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mi_memcpy_h2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    "mi_memcpy_d2h": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    "mi_memcpy_d2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    "mi_memset": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_size_t]),
    "mi_dot_f64": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, c_f64p]),
    "mi_norm_f64": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, c_f64p]),
    "mi_dot_f32": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.c_float)]),
    "mi_norm_f32": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.c_float)]),
    "mi_sys_create": (C.c_int, [C.c_void_p, C.POINTER(SysDesc), C.POINTER(C.c_void_p)]),
    "mi_sys_destroy": (C.c_int, [C.c_void_p]),
    "mi_sys_assemble": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mi_sys_assemble_f32": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mi_sys_nnz": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "mi_sys_export_bcoo": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mi_spmv_dot": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mi_spmv_dot_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mi_sys_update_extents": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.POINTER(C.c_int64),
                                        C.POINTER(C.c_int64)]),
    "mi_sys_begin_assembly": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                        C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_void_p]),
    "mi_sys_assemble_owned": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mi_spmv_shapes": (C.c_int, [C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_void_p]),
    "mi_sys_set_spmv_shape": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    "mi_sys_spmv_layout": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                     C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    "mi_spmv": (C.c_int, [C.c_void_p, C.c_double, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p]),
    "mi_sys_spmv_order": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mi_spmv_ordered": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mi_precond_add": (C.c_int, [C.c_void_p, C.POINTER(PrecondDesc), C.POINTER(C.c_int)]),
    "mi_precond_reinit": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(PrecondDesc)]),
    "mi_precond_mas_set_active_levels": (C.c_int, [C.c_void_p, C.c_int, C.c_int32]),
    "mi_precond_mas_set_system_levels": (C.c_int, [C.c_void_p, C.c_int, C.c_int32, C.c_void_p, C.c_void_p]),
    "mi_precond_assemble": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mi_precond_apply": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mi_precond_mas_export": (C.c_int, [C.c_void_p, C.c_int, c_i32p, c_i32p, c_i32p, c_i32p, c_i32p, c_i32p, c_f64p,
                                        c_f64p]),
    "mi_sys_write_mtx": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int]),
    "mi_subsystem_fix_vertices": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                            C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "mi_mesh_partition": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p,
                                    C.c_int64, C.c_int32, C.c_void_p, C.POINTER(C.c_int32), C.c_void_p]),
    "mi_subsystem_distribute": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p]),
    "mi_precond_mas_dump": (C.c_int, [C.c_void_p, C.c_int, C.c_char_p, C.c_int64, C.c_int64]),
    "mi_vec_write_mtx": (C.c_int, [C.c_void_p, C.c_char_p, C.c_void_p, C.c_int64]),
    "mi_mtx_read_block_triplets": (C.c_int, [C.c_char_p, C.c_int, c_i32p, C.POINTER(C.c_int64), c_i32p, c_i32p,
                                             c_f64p]),
    "mi_vec_read_mtx": (C.c_int, [C.c_char_p, C.POINTER(C.c_int64), c_f64p]),
    "mi_produce_abd": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_double, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mi_produce_bw_shell": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                      C.c_double, C.c_double, C.c_double, C.c_int32, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p]),
    "mi_produce_bending": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                     C.c_double, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mi_step_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_int64,
                                  C.c_void_p]),
    "mi_energy_kinetic": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p]),
    "mi_energy_snh": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                C.c_double, C.c_void_p, C.c_void_p]),
    "mi_produce_kinetic": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_void_p]),
    "mi_produce_snh": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                 C.c_double, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mi_produce_contact": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                     C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mi_ws_gemm": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_int64, C.c_int64,
                             C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_int64, C.c_int32, C.c_void_p,
                             C.c_int32, C.c_void_p]),
    "mi_ws_aggregate": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p]),
    "mi_ws_relu_backward": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mi_ws_colsum": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    "mi_ws_adam": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_float,
                             C.c_float, C.c_float, C.c_float, C.c_int32, C.c_void_p]),
    "mi_ws_block3": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_void_p, C.c_void_p, C.c_void_p]),
    "mi_pcg_default_opts": (C.c_int, [C.POINTER(PcgOpts)]),
    "mi_pcg_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(PcgOpts), C.POINTER(PcgStats),
                               C.c_void_p]),
    "mi_pcg_solve_batched": (C.c_int, [C.c_void_p, C.c_int32, c_i32p, C.c_void_p, C.c_void_p, C.POINTER(PcgOpts),
                                       C.POINTER(PcgStats), C.c_void_p]),
}

SYNTH_PROTOS = {
    "mi_synth_fem_sizes_of": (C.c_int, [C.POINTER(SynthFemDesc), C.POINTER(SynthFemSizes)]),
    "mi_synth_fem_fill": (C.c_int, [C.POINTER(SynthFemDesc), c_i32p, c_i32p, c_f64p, c_f64p, c_i32p, c_i32p,
                                    c_f64p]),
    "mi_synth_fem_aux": (C.c_int, [C.POINTER(SynthFemDesc), c_f64p, c_f64p, c_i32p, c_f64p, c_f64p]),
    "mi_synth_last_error": (C.c_char_p, []),
}

_LIBS: dict[str, C.CDLL] = {}


def _load(path: str, protos: dict) -> C.CDLL:
    if path in _LIBS:
        return _LIBS[path]
    if not os.path.exists(path):
        raise RuntimeError(
            f"native library {path} is missing: build it first with `python -m libuipc_amd.build` "
            "(there is no fallback path)")
    lib = C.CDLL(path)
    for name, (res, args) in protos.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIBS[path] = lib
    return lib


def linsys() -> C.CDLL:
    return _load(LINSYS_SO, LINSYS_PROTOS)


def synth() -> C.CDLL:
    return _load(SYNTH_SO, SYNTH_PROTOS)


class MiError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def check(code: int) -> None:
    if code != MI_OK:
        raise MiError(code, linsys().mi_last_error().decode(errors="replace"))
