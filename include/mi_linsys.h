/*
 * mi_linsys.h -- C ABI of the MI355X (gfx950) per-Newton global linear solve.
 *
 * The library replaces four pieces of libuipc's CUDA backend
 * (src/backends/cuda/ in the reference tree):
 *   (a) triplet -> sorted, de-duplicated upper-triangular 3x3 BCOO conversion
 *       (MatrixConverter<Float,3>::convert_sym),
 *   (b) the block-symmetric SpMV (+ fused p^T A p),
 *   (c) the local preconditioners (3x3 / 12x12 block Jacobi and the MAS
 *       multi-level additive Schwarz preconditioner, build and apply),
 *   (d) the PCG driver and its dot / axpy / convergence kernels.
 *
 * Conventions
 *   - Plain C types only; every pointer is a DEVICE pointer unless its name
 *     ends in `_host`.
 *   - A 3x3 block value is 9 doubles in column-major order (Eigen::Matrix3d's
 *     layout), i.e. element (r, c) at index r + 3 c.
 *   - Every function returns MI_OK (0) or a negative MI_ERR_* code; the text
 *     of the last error of the calling thread is returned by mi_last_error().
 *     No C++ exception crosses this boundary.
 *   - `stream` is a hipStream_t passed as void*; NULL selects the context's
 *     own stream.  Calls on one mi_sys must be serialised by the caller
 *     (the reference is single-threaded per GlobalLinearSystem as well).
 */
#ifndef MI_LINSYS_H
#define MI_LINSYS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MI_LINSYS_VERSION 1

enum mi_status
{
    MI_OK              = 0,
    MI_ERR_ARG         = -1, /* invalid argument (reference: UIPC_ASSERT / std::invalid_argument) */
    MI_ERR_HIP         = -2, /* a HIP runtime call failed (reference: CUDA_TOOL_CHECK)            */
    MI_ERR_STATE       = -3, /* call out of order (e.g. apply before assemble)                   */
    MI_ERR_NOMEM       = -4, /* device allocation failed                                        */
    MI_ERR_UNSUPPORTED = -5, /* option not supported by this build                               */
    MI_ERR_NONFINITE   = -6  /* NaN / Inf in rz (reference: check_init_rz_nan_inf)               */
};

typedef struct mi_ctx mi_ctx;
typedef struct mi_sys mi_sys;

/* text of the calling thread's last error ("" if none) */
const char* mi_last_error(void);
int         mi_version(void);

/* ------------------------------------------------------------------ context */
int mi_ctx_create(int device, mi_ctx** out);
int mi_ctx_destroy(mi_ctx* ctx);
int mi_ctx_synchronize(mi_ctx* ctx);
/* the context's stream as a hipStream_t */
int mi_ctx_stream(mi_ctx* ctx, void** stream_out);

/* Device-memory helpers for hosts that have no allocator of their own
 * (test harnesses, ctypes callers).  Synchronous. */
int mi_malloc(mi_ctx* ctx, void** dptr, size_t bytes);
int mi_free(mi_ctx* ctx, void* dptr);
int mi_memcpy_h2d(mi_ctx* ctx, void* dst, const void* src_host, size_t bytes);
int mi_memcpy_d2h(mi_ctx* ctx, void* dst_host, const void* src, size_t bytes);
int mi_memcpy_d2d(mi_ctx* ctx, void* dst, const void* src, size_t bytes);
int mi_memset(mi_ctx* ctx, void* dst, int value, size_t bytes);

/* ------------------------------------------------------------------ vectors */
/* LinearSystemContext::dot / ::norm
 * (cuda_tool/linear_system.h:436-502, cuda_tool/linear_reduction.h:14-60).
 * The norm is the overflow-safe scaled 2-norm.  Results land on the host;
 * the calls synchronise the context stream. */
int mi_dot_f64(mi_ctx* ctx, const double* x, const double* y, int64_t n, double* out_host);
int mi_norm_f64(mi_ctx* ctx, const double* x, int64_t n, double* out_host);
int mi_dot_f32(mi_ctx* ctx, const float* x, const float* y, int64_t n, float* out_host);
int mi_norm_f32(mi_ctx* ctx, const float* x, int64_t n, float* out_host);

/* ------------------------------------------------------------------ system */
enum mi_precision
{
    MI_FP64 = 0, /* plain method: every stream in double                        */
    MI_FP32 = 1  /* precision arm: float matrix / MAS streams, double dots       */
};

typedef struct mi_sys_desc
{
    int32_t block_rows;      /* B: number of 3x3 block rows (n = 3 B)                     */
    double  reserve_ratio;   /* capacity growth factor (reference GlobalLinearSystem: 1.1) */
    int32_t value_precision; /* mi_precision of the SpMV's matrix stream                  */
} mi_sys_desc;

int mi_sys_create(mi_ctx* ctx, const mi_sys_desc* desc, mi_sys** out);
int mi_sys_destroy(mi_sys* sys);

/* MatrixConverter<Float,3>::convert_sym
 * (algorithm/details/matrix_converter.inl:1549-1699).
 * Keeps the triplets with 0 <= row <= col (others are dropped, as the
 * reference's flag kernel at matrix_converter.inl:129-143 does), sorts them by
 * (row, col), sums duplicates and builds the BSR row offsets
 * (matrix_converter.inl:1041-1096).  Duplicates are summed in ascending source
 * order, so the values are deterministic.  The triplet arrays are read, never
 * written.  One host synchronisation: the unique-block count is read back
 * to size the outputs (the reference reads two counts, matrix_converter.inl). */
int mi_sys_assemble(mi_sys*        sys,
                    int64_t        n_triplets,
                    const int32_t* rows,
                    const int32_t* cols,
                    const double*  vals,
                    void*          stream);
/* The same conversion from FP32 triplet values (the precision arm: producers
 * that write their Hessian blocks in float).  Duplicates are summed in FP64 in
 * ascending source order, so the blocks equal mi_sys_assemble's on the same
 * values widened to double.  FP32 systems only (value_precision = MI_FP32,
 * MI_ERR_ARG otherwise).  No counterpart in the reference, whose triplets
 * are Float = double. */
int mi_sys_assemble_f32(mi_sys*        sys,
                        int64_t        n_triplets,
                        const int32_t* rows,
                        const int32_t* cols,
                        const float*   vals,
                        void*          stream);

/* ------------------------------------------------ library-owned triplets and gradient
 * GlobalLinearSystem::Impl::_update_subsystem_extent (global_linear_system.cu:321-405).
 * The subsystems report their extents in registration order.  Diagonal ones
 * (DiagLinearSubsystem) own consecutive DoF ranges, in the order given; every
 * subsystem owns a consecutive triplet range, in the order given (an
 * off-diagonal one, OffDiagLinearSubsystem, holds its lr blocks, then its rl
 * blocks).  The library keeps the triplet buffer and the gradient b; their
 * capacity grows to reserve_ratio x the total when the total exceeds it, and
 * the contents are discarded then (reserve_triplets_discard).  A new DoF total
 * resizes the system (block_rows = DoFs / 3): assemble again, and re-create or
 * re-init preconditioners whose ranges changed. */
typedef struct mi_extent
{
    int32_t is_diag;        /* 1: diagonal subsystem, 0: off-diagonal coupling                   */
    int64_t dof_count;      /* diagonal: scalar DoFs (a multiple of 3); off-diagonal: 0           */
    int64_t block_count;    /* diagonal: triplets (DiagExtentInfo::m_block_count); off-diagonal:
                               lr triplets (OffDiagExtentInfo::m_lr_block_count)                  */
    int64_t rl_block_count; /* off-diagonal: rl triplets (m_rl_block_count); diagonal: 0         */
    int32_t l, r;           /* off-diagonal: the diagonal subsystems it couples (their indices in
                               this array); diagonal: ignored                                     */
} mi_extent;

typedef struct mi_extent_view /* per subsystem, written by mi_sys_update_extents */
{
    int64_t dof_offset;        /* diagonal: first scalar DoF (its gradient is b + dof_offset); else -1 */
    int64_t triplet_offset;    /* first triplet of its range (lr part for an off-diagonal one)         */
    int64_t rl_triplet_offset; /* off-diagonal: first rl triplet; diagonal: -1                         */
    int32_t row_offset;        /* block offsets of the (lr) submatrix: a block (i, j) of it is written */
    int32_t col_offset;        /* as row = row_offset + i, col = col_offset + j (rl: swapped offsets)  */
} mi_extent_view;

/* views[k] for ext[k]; *total_triplets (may be NULL) = the triplet total,
 * *triplet_capacity (may be NULL) = the buffer's capacity after growth. */
int mi_sys_update_extents(mi_sys*          sys,
                          int32_t          n,
                          const mi_extent* ext_host,
                          mi_extent_view*  views_host,
                          int64_t*         total_triplets,
                          int64_t*         triplet_capacity);

/* The start of GlobalLinearSystem::Impl::_assemble_linear_system
 * (global_linear_system.cu:407-420): rows and cols of every slot = -1 (a slot
 * no subsystem writes is dropped by the assembly's row >= 0 filter), b = 0;
 * returns the device buffers (valid until the next mi_sys_update_extents), for
 * the subsystems to write their slices.  Any output may be NULL. */
int mi_sys_begin_assembly(mi_sys* sys, int32_t** rows, int32_t** cols, double** vals, double** b, void* stream);

/* mi_sys_assemble of the library-owned triplet buffer (all of it). */
int mi_sys_assemble_owned(mi_sys* sys, void* stream);

/* Number of unique upper blocks of the last assembly.  Synchronises. */
int mi_sys_nnz(mi_sys* sys, int64_t* nnz_host);

/* Copies the assembled matrix out in the reference's BCOO layout
 * (rows, cols ascending by (row, col); values column-major 3x3) plus the BSR
 * row offsets (block_rows + 1 entries).  Any output may be NULL. */
int mi_sys_export_bcoo(mi_sys*  sys,
                       int32_t* rows,
                       int32_t* cols,
                       double*  vals,
                       int32_t* row_ptr,
                       void*    stream);

/* Spmv::rbk_sym_spmv_dot (linear_system/spmv.cu:1023-1159):
 * y += A x with A = U + U^T - diag(U), and *d_dot = x^T A x (device scalar,
 * overwritten).  As in the reference, y is accumulated into: pass a zeroed y
 * for y = A x (LinearFusedPCG zeroes Ap in its update kernel). */
int mi_spmv_dot(mi_sys* sys, const double* x, double* y, double* d_dot, void* stream);

/* No counterpart in the reference: mi_spmv_dot for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same operation and results as
 * mi_spmv_dot, but its workgroup partials (and, for FP32 value planes on the
 * full-BSR layout, MI_SPMV_LAYOUT=full, the float4 copy of x) live in a fixed
 * device array of 4 MiB + 64 KiB per GPU, set up by mi_ctx_create; it calls
 * no allocator.  On the full-BSR layout FP32 systems are limited to 262144
 * block rows (MI_ERR_ARG beyond); the default window layout has no limit.
 * All calls on one GPU share that array: calls that overlap on one GPU
 * (other streams, other threads) race. */
int mi_spmv_dot_noalloc(mi_sys* sys, const double* x, double* y, double* d_dot, void* stream);

/* The SpMV's launch shapes, for a tuner (the reference fits its SpMV grid to
 * each matrix, Spmv::fit_block_count, linear_system/spmv.cu:909-933).  count:
 * how many; dflt: the one systems with enough rows get; per shape (arrays of
 * count entries, any may be NULL): threads per workgroup, window rows, FP64
 * blocks per lane and stage, blocks per lane and load (2: paired 16-byte
 * FP64 / 8-byte FP32 loads; 4: 16-byte loads of four FP32 blocks, which an
 * FP64 system runs as 2). */
int mi_spmv_shapes(int32_t* count, int32_t* dflt, int32_t* threads, int32_t* rows, int32_t* per, int32_t* vec);
enum
{
    MI_SPMV_SHAPE_AUTO = -1, /* by the system's size (the default): the largest window whose count
                                fills the GPU (1.5 windows per CU), else the full BSR             */
    MI_SPMV_SHAPE_FULL = -2  /* the full BSR (both triangles, a row gather)                       */
};
/* The layout this system's SpMV uses: a window shape (an index of
 * mi_spmv_shapes), MI_SPMV_SHAPE_AUTO or MI_SPMV_SHAPE_FULL.  An assembled
 * system gets the new layout at once, on `stream`. */
int mi_sys_set_spmv_shape(mi_sys* sys, int32_t shape, void* stream);

/* Diagnostic, no counterpart in the reference: the layout the SpMV streams.
 * layout: 1 = window layout (rows in the first MAS preconditioner's part
 * order, windows of window_rows rows, a block inside a window stored once),
 * 0 = full BSR (both triangles); blocks: the 3x3 blocks one product streams;
 * threads: the SpMV kernel's workgroup size; shape: its launch shape (an
 * index of mi_spmv_shapes, -1 on the full BSR).  Any output may be NULL. */
int mi_sys_spmv_layout(mi_sys*  sys,
                       int32_t* layout,
                       int32_t* window_rows,
                       int32_t* threads,
                       int64_t* blocks,
                       int32_t* shape);

/* Spmv::rbk_sym_spmv (linear_system/spmv.cu:992-1021): y = a A x + b y. */
int mi_spmv(mi_sys* sys, double a, const double* x, double b, double* y, void* stream);

/* The SpMV's row order (no counterpart in the reference; for callers that
 * keep vectors in it, and for timing the PCG's own product).  pos (host,
 * block_rows entries): the row position of every block row; the identity
 * unless a window layout puts the rows in a MAS part order. */
int mi_sys_spmv_order(mi_sys* sys, int32_t* pos);
/* y = A x, *d_dot = x^T A x (when d_dot != NULL) with x and y in the SpMV's
 * row order (row position k holds block row pos^-1(k)): the one kernel the
 * PCG launches per iteration.  No counterpart in the reference.  Requires
 * the window layout (MI_ERR_STATE on the full BSR). */
int mi_spmv_ordered(mi_sys* sys, const double* x, double* y, double* d_dot, void* stream);

/* ------------------------------------------------------------ preconditioners */
enum mi_precond_kind
{
    MI_PRECOND_DIAG3  = 1, /* FEMDiagPreconditioner (finite_element/fem_diag_preconditioner.cu:18-117)   */
    MI_PRECOND_DIAG12 = 2, /* ABDDiagPreconditioner (affine_body/abd_diag_preconditioner.cu:15-122)      */
    MI_PRECOND_MAS    = 3  /* FEMMASPreconditioner + MASPreconditionerEngine
                              (finite_element/fem_mas_preconditioner.cu:157-469,
                               finite_element/mas_preconditioner_engine.cu)                             */
};

typedef struct mi_precond_desc
{
    int32_t kind;       /* mi_precond_kind                                                   */
    int64_t dof_offset; /* first scalar DoF of the subsystem (LocalPreconditionerAssemblyInfo) */
    int64_t dof_count;  /* scalar DoFs of the subsystem                                      */

    /* MAS only: mesh topology and partition, read at build time (host memory).
     * The adjacency is built from the simplices as FEMMASPreconditioner::do_init
     * does (tets, triangles and edges of the subsystem, vertex ids local to it). */
    int32_t        vert_num;      /* subsystem vertices (dof_count / 3)                   */
    const int32_t* part_ids_host; /* per vertex partition id, < 0 = unpartitioned;
                                     at most 16 vertices per id (mesh_partition's BANKSIZE) */
    const int32_t* tets_host;     /* 4 ids per tet (may be NULL)  */
    int64_t        n_tets;
    const int32_t* tris_host;     /* 3 ids per triangle (may be NULL) */
    int64_t        n_tris;
    const int32_t* edges_host;    /* 2 ids per edge (may be NULL) */
    int64_t        n_edges;
    int32_t        mas_precision;     /* MI_FP32: float inverses + float multi-level R/Z
                                         (the reference's storage); MI_FP64: all double  */
    int32_t        mas_active_levels; /* 0 = all levels (MASPreconditionerEngine::set_active_level_num) */

    /* Where the MAS adjacency, partition slots and level hierarchy are built
     * (the init kernels of mas_preconditioner_engine.cu:57-466 and the host
     * adjacency of fem_mas_preconditioner.cu:163-356).  Both give the same
     * tables bit for bit.  Zero-initialise the descriptor: 0 is the host build. */
    int32_t        mas_build;     /* mi_mas_build                                          */
    /* MI_MAS_BUILD_DEVICE only: device copies of the arrays above; a non-NULL one
     * is read instead of its host array (same layout and counts). */
    const int32_t* part_ids_dev;
    const int32_t* tets_dev;
    const int32_t* tris_dev;
    const int32_t* edges_dev;
    /* How the 48x48 cluster matrices are inverted (invert_cluster_matrices,
     * mas_preconditioner_engine.cu:2567-2589): both are FP64 Gauss-Jordan
     * without pivoting and differ only in rounding.  0 is the default. */
    int32_t        mas_invert;    /* mi_mas_invert                                         */
} mi_precond_desc;

enum mi_mas_build
{
    MI_MAS_BUILD_HOST   = 0,
    MI_MAS_BUILD_DEVICE = 1
};

enum mi_mas_invert
{
    MI_MAS_INVERT_MFMA  = 0, /* the default: 4-row pivot panels, the sweep in its Cholesky
                                form, the Schur updates on FP64 matrix cores            */
    MI_MAS_INVERT_SWEEP = 1  /* the reference's column sweep, one pivot at a time       */
};

/* GlobalLinearSystem's local-preconditioner registration + LocalPreconditioner::do_init.
 * Subsystem ranges must not overlap; DoFs no preconditioner covers get z = r
 * (global_linear_system.cu:546-590). */
int mi_precond_add(mi_sys* sys, const mi_precond_desc* desc, int* handle_out);

/* Rebuilds preconditioner `handle` in place from a new descriptor (any kind;
 * the same overlap rules as mi_precond_add), for a topology change: the MAS
 * adjacency, partition and hierarchy are rebuilt (on the device with
 * MI_MAS_BUILD_DEVICE).  The reference re-runs LocalPreconditioner::do_init
 * for it (fem_mas_preconditioner.cu:157-404).  Synchronises the device; call
 * mi_precond_assemble before the next apply or solve. */
int mi_precond_reinit(mi_sys* sys, int handle, const mi_precond_desc* desc);

/* MASPreconditionerEngine::set_active_level_num
 * (finite_element/mas_preconditioner_engine.h:104-110): the number of MAS
 * levels apply uses, from the finest; 1 is the 16-vertex block Jacobi, 0 or
 * more than the hierarchy has means every level.  Takes effect at the next
 * apply or solve (a captured solve is re-captured).  MI_ERR_ARG for a
 * handle that is not a MAS preconditioner. */
int mi_precond_mas_set_active_levels(mi_sys* sys, int handle, int32_t levels);
/* A level count per system for batched solves (no counterpart in the
 * reference, which holds one count per engine): system g, block rows
 * [seg_rows[g], seg_rows[g + 1]) (host arrays, nseg + 1 and nseg entries),
 * uses levels[g] levels (0 or out of range: every level).  A fine cluster
 * takes its system's count; the restriction and coarse solves run to the
 * largest.  nseg = 0, or mi_precond_mas_set_active_levels, returns to one
 * count.  MI_ERR_ARG for a handle that is not a MAS preconditioner or a
 * vertex outside the systems' rows. */
int mi_precond_mas_set_system_levels(mi_sys*        sys,
                                     int            handle,
                                     int32_t        nseg,
                                     const int32_t* seg_rows,
                                     const int32_t* levels);

/* LocalPreconditioner::do_assemble for every registered preconditioner
 * (global_linear_system.cu:499-515).  Requires a prior mi_sys_assemble. */
int mi_precond_assemble(mi_sys* sys, void* stream);

/* GlobalLinearSystem::Impl::apply_preconditioner (global_linear_system.cu:546-590):
 * z = P^{-1} r.  d_converged may be NULL; when it points at a nonzero int the
 * kernels return early, as the reference's do. */
int mi_precond_apply(mi_sys* sys, const double* r, double* z, const int32_t* d_converged, void* stream);

/* State of a MAS preconditioner for parity checks and debugging, the analogue
 * of MASPreconditionerEngine::dump_cluster_matrices_debug
 * (finite_element/mas_preconditioner_engine.cu:3154-3301).  Host outputs; any
 * may be NULL.  info_host[6] = {vert_num, map_nodes, levels, total_slots,
 * level1_begin, clusters}.  Sizes: part_to_real map_nodes, real_to_part
 * vert_num, going_next bank_align(max(map_nodes, vert_num)) * levels,
 * level_sizes 2 (levels + 1) as (count, begin) pairs, coarse_tables 6 vert_num,
 * cluster_hess / cluster_inv 136 * 9 per cluster (upper 3x3 blocks,
 * column-major; the inverse widened to double).  Synchronises. */
int mi_precond_mas_export(mi_sys*  sys,
                          int      handle,
                          int32_t* info_host,
                          int32_t* part_to_real_host,
                          int32_t* real_to_part_host,
                          int32_t* going_next_host,
                          int32_t* level_sizes_host,
                          int32_t* coarse_tables_host,
                          double*  cluster_hess_host,
                          double*  cluster_inv_host);

/* The debug dump of MASPreconditionerEngine::dump_cluster_matrices_debug
 * (finite_element/mas_preconditioner_engine.cu:3154-3301) into output_dir:
 * mas_cluster_hess.f<frame>.n<newton_iter>.mtx and mas_cluster_inv.(...).mtx
 * (coordinate real general, block diagonal of all clusters, 136 upper 3x3
 * blocks each, %.17g) and mas_cluster_meta.(...).json (sizes, levels,
 * part_to_real, going_next, real_to_part).  The reference writes its
 * allocation's 5% slack clusters too; this writes the clusters in use.
 * Synchronises. */
int mi_precond_mas_dump(mi_sys* sys, int handle, const char* output_dir, int64_t frame, int64_t newton_iter);

/* geometry::mesh_partition (src/geometry/mesh_partition.cpp:77-230), the MAS
 * partition FEMMASPreconditioner::do_init computes with BANKSIZE
 * (fem_mas_preconditioner.cu:205-249), on the device.  Same contract: the
 * adjacency of the highest-dimensional simplices given (tets, else triangles,
 * else edges); no adjacency -> part i / part_max_size; vert_num <=
 * part_max_size -> connected-component labels (numbered by each component's
 * lowest vertex); otherwise parts of at most part_max_size vertices, ids
 * 0..P-1, none empty.  The k-way step is GPU handshake matching plus packing,
 * not METIS, so the ids differ from the reference's; it is deterministic.
 * Device arrays in and out (part_ids: vert_num int32); *n_parts_out = P
 * (host, may be NULL).  Synchronises. */
int mi_mesh_partition(mi_ctx*        ctx,
                      int32_t        vert_num,
                      const int32_t* tets,
                      int64_t        n_tets,
                      const int32_t* tris,
                      int64_t        n_tris,
                      const int32_t* edges,
                      int64_t        n_edges,
                      int32_t        part_max_size,
                      int32_t*       part_ids,
                      int32_t*       n_parts_out,
                      void*          stream);

/* ------------------------------------------------ FEM subsystem, around the solve */
/* FEMLinearSubsystem_assemble_k1/k2_kernel (finite_element/fem_linear_subsystem.cu:23-53),
 * on one subsystem's slice of the triplet buffer before mi_sys_assemble: every
 * triplet whose (row - block_offset) or (col - block_offset) is a fixed vertex
 * (is_fixed[v] != 0, v < vert_num) becomes 0 off the diagonal and the identity
 * on it (per triplet, as in the reference, so k triplets on a fixed diagonal
 * sum to k I); the fixed vertices' gradient entries become 0.  gradient (3
 * vert_num doubles, the subsystem's slice) may be NULL.  Triplets with a
 * negative row are left alone. */
int mi_subsystem_fix_vertices(mi_ctx*        ctx,
                              int64_t        n_triplets,
                              const int32_t* rows,
                              const int32_t* cols,
                              double*        vals,
                              int32_t        block_offset,
                              const int32_t* is_fixed,
                              int32_t        vert_num,
                              double*        gradient,
                              void*          stream);

/* GlobalLinearSystem::Impl::distribute_solution (global_linear_system.cu:531-544)
 * -> FEMLinearSubsystem_retrieve_solution_kernel (fem_linear_subsystem.cu:119-126):
 * dxs[i] = -x[dof_offset + i] for the subsystem's 3 vert_num DoFs. */
int mi_subsystem_distribute(mi_ctx* ctx, const double* x, int64_t dof_offset, int32_t vert_num, double* dxs,
                            void* stream);

/* ------------------------------------------------------------ Matrix Market */
/* The reference's debug dumps (extras/debug/dump_linear_system,
 * linear_system/global_linear_system.cu:42-70; format utils/matrix_market.h:26-158).
 * mi_sys_write_mtx: the assembled upper BCOO as "coordinate real general", the
 * 9 scalar entries of every block including zeros, %.17g, 1-based indices
 * unless one_based == 0.  mi_vec_write_mtx: a device vector as
 * "array real general".  Both synchronise. */
int mi_sys_write_mtx(mi_sys* sys, const char* path, int one_based);
int mi_vec_write_mtx(mi_ctx* ctx, const char* path, const double* x, int64_t n);

/* Host-side readers.  mi_mtx_read_block_triplets groups the scalar entries of
 * a coordinate file into 3x3 blocks (duplicates summed), ascending by
 * (block row, block col); call it with NULL arrays to get the sizes, then
 * again with arrays of n_triplets entries (vals: 9 per block, column-major).
 * mi_vec_read_mtx reads an "array" file; x_host may be NULL to get n. */
int mi_mtx_read_block_triplets(const char* path,
                               int         one_based,
                               int32_t*    block_rows_out,
                               int64_t*    n_triplets_out,
                               int32_t*    rows_host,
                               int32_t*    cols_host,
                               double*     vals_host);
int mi_vec_read_mtx(const char* path, int64_t* n_out, double* x_host);

/* ----------------------------------------------------------------------- PCG */
enum mi_pcg_criterion
{
    MI_CRIT_RZ_RATIO     = 0, /* |r^T z| / |r0^T z0| <= tol  (linear_fused_pcg.cu:1985) */
    MI_CRIT_REL_RESIDUAL = 1  /* ||r||_2 / ||b||_2 <= tol (recurrence residual)          */
};

typedef struct mi_pcg_opts
{
    int32_t criterion;      /* mi_pcg_criterion                                            */
    double  tol;            /* reference default tol_rate = 1e-3 (scene_default_config.cpp:146) */
    double  max_iter_ratio; /* max_iter = ratio * n (reference: 2)                         */
    int32_t check_interval; /* iterations between convergence checks (reference: 5)        */
    int32_t use_graph;      /* 1: replay a captured hipGraph of graph_checks check intervals */
    int32_t psd_fallback;   /* 1: rz0 < 0 or non-finite -> plain CG (linear_fused_pcg.cu:1843-1875) */
    int32_t force_bypass;   /* 1: always bypass the preconditioner (UIPC_PCG_PSD_FALLBACK_TEST) */
    int32_t true_residual;  /* 1: compute ||b - A x|| / ||b|| after the solve (one SpMV)    */
    int32_t graph_checks;   /* check intervals per host round trip (0: 4).  The device tests
                               convergence every iteration and idles once converged, so the
                               reported counts do not depend on it; larger values cut host
                               round trips and cost at most that many idle intervals.      */
    int32_t x_init;         /* 1: x holds the initial guess, r0 = b - A x0 (a Newton step's
                               previous solution, a learned guess); 0: x0 = 0, as
                               LinearFusedPCG (linear_fused_pcg.cu:1811-1824).  The relative
                               residual stays relative to ||b||; the rz criterion to r0.z0. */
    int32_t keep_order;     /* 0 (default): when the SpMV's window layout puts the rows in a
                               MAS part order, the iteration runs on vectors in that order (b
                               gathered in, x scattered out; same iterates up to rounding);
                               1: on the caller's vertex order throughout, as the reference.
                               No counterpart in the reference.                             */
} mi_pcg_opts;

typedef struct mi_pcg_stats
{
    int64_t iters;          /* iterations the reference would report (multiple of check_interval
                               or max_iter when not converged, linear_fused_pcg.cu:1994)   */
    int64_t iters_exact;    /* first iteration whose check passed                          */
    double  rz0;
    double  rz_final;
    double  rel_res;        /* ||r_k|| / ||b|| of the recurrence at exit                   */
    double  rel_res_true;   /* ||b - A x|| / ||b|| (NaN unless opts.true_residual)          */
    int32_t converged;
    int32_t bypassed;       /* preconditioner bypassed by the PSD guard                    */
} mi_pcg_stats;

int mi_pcg_default_opts(mi_pcg_opts* opts);

/* LinearFusedPCG::do_solve + fused_pcg (linear_system/linear_fused_pcg.cu:851-937,1782-2004):
 * x0 = 0, r = b, preconditioned CG on the assembled matrix with the registered
 * preconditioners.  b and x are device vectors of n = 3 * block_rows doubles.
 * Synchronises the stream at each check (every check_interval iterations). */
int mi_pcg_solve(mi_sys*            sys,
                 const double*      b,
                 double*            x,
                 const mi_pcg_opts* opts,
                 mi_pcg_stats*      stats_host,
                 void*              stream);

/* Batched PCG: the system holds nseg independent systems on the block-row
 * ranges [seg_rows_host[k], seg_rows_host[k + 1]) (seg_rows_host[0] = 0,
 * seg_rows_host[nseg] = block_rows; no block may couple two ranges, and the
 * registered preconditioners must not either -- a MAS partition must not
 * cross them).  Every system runs FusedPCG's semantics with its own alpha,
 * beta, convergence test and cap (max_iter_ratio times its own DoFs); all of
 * them advance in the same captured iterations (one graph replay serves the
 * batch), and a system that has stopped keeps its x.  The PSD guard runs per
 * system, as the reference's FusedPCG does once per system
 * (linear_fused_pcg.cu:1843-1875): with psd_fallback, a system whose rz0 < 0,
 * or whose z is non-finite while its r is finite, is solved as plain CG
 * (stats.bypassed = 1) while the others keep the preconditioner; force_bypass
 * bypasses every system.  A non-finite r (a NaN in b) fails the whole call
 * with MI_ERR_NONFINITE.  max_iter_ratio caps each system at max_iter - 1
 * iterations, as the single-system solve does.  stats_host: nseg entries.
 * Synchronises at each block of graph_checks check intervals. */
int mi_pcg_solve_batched(mi_sys*            sys,
                         int32_t            nseg,
                         const int32_t*     seg_rows_host,
                         const double*      b,
                         double*            x,
                         const mi_pcg_opts* opts,
                         mi_pcg_stats*      stats_host,
                         void*              stream);

#ifdef __cplusplus
}
#endif

#endif /* MI_LINSYS_H */
