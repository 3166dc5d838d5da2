/*
 * oracle.h -- CPU restatement (FP64, plain C) of libuipc's per-Newton linear
 * solve, used ONLY as the checker: tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg call it.  The product (libuipc_amd) never links,
 * loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * src/backends/cuda/ of the reference tree).  Arrays are host memory.
 * 3x3 blocks are 9 doubles, column-major ((r, c) at r + 3 c).
 */
#ifndef MI_ORACLE_H
#define MI_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* number of radix bits the reference sorts on: bit_width(max_key), at least 1
 * (algorithm/details/matrix_converter.inl:726-735) */
int or_key_bits(uint64_t max_key);

/* MatrixConverter::convert_sym (matrix_converter.inl:1549-1699) + the BSR row
 * offsets (matrix_converter.inl:1041-1096).  Output arrays hold up to T entries;
 * row_ptr holds B + 1.  Returns U (unique upper blocks) or -1. */
int64_t or_convert_sym(int32_t        B,
                       int64_t        T,
                       const int32_t* rows,
                       const int32_t* cols,
                       const double*  vals,
                       int32_t*       out_rows,
                       int32_t*       out_cols,
                       double*        out_vals,
                       int32_t*       row_ptr);

/* Spmv::cpu_sym_spmv (linear_system/spmv.cu:1210-1296) with a = 1, b = 0:
 * y = (U + U^T - diag U) x;  returns x^T y (the spmv_dot scalar). */
double or_spmv_sym(int32_t        B,
                   int64_t        U,
                   const int32_t* rows,
                   const int32_t* cols,
                   const double*  vals,
                   const double*  x,
                   double*        y);

/* LinearSystemContext::dot / norm (cuda_tool/linear_system.h:436-502,
 * cuda_tool/linear_reduction.h:14-60): norm is the scaled square sum. */
double or_dot(const double* x, const double* y, int64_t n);
double or_norm(const double* x, int64_t n);

/* ---------------------------------------------------------------- diag */
/* FEMDiagPreconditioner (finite_element/fem_diag_preconditioner.cu:18-56):
 * diag_inv[v] = inverse of the diagonal block of vertex (blk_off + v). */
int or_diag3_assemble(int64_t        U,
                      const int32_t* rows,
                      const int32_t* cols,
                      const double*  vals,
                      int32_t        blk_off,
                      int32_t        blk_cnt,
                      double*        diag_inv);
void or_diag3_apply(int32_t blk_cnt, const double* diag_inv, const double* r, double* z);

/* ABDDiagPreconditioner (affine_body/abd_diag_preconditioner.cu:15-122):
 * 12x12 inverse per body of the body's 4x4 diagonal 3x3 blocks (upper blocks
 * mirrored).  body_inv is 144 doubles per body, column-major. */
int or_diag12_assemble(int64_t        U,
                       const int32_t* rows,
                       const int32_t* cols,
                       const double*  vals,
                       int32_t        blk_off,
                       int32_t        bodies,
                       double*        body_inv);
void or_diag12_apply(int32_t bodies, const double* body_inv, const double* r, double* z);

/* ------------------------------------------------------------------ MAS */
typedef struct or_mas or_mas;

/* FEMMASPreconditioner::do_init (finite_element/fem_mas_preconditioner.cu:157-404)
 * + MASPreconditionerEngine::init_neighbor / init_matrix / reorder_realtime
 * (finite_element/mas_preconditioner_engine.cu:1958-2078, kernels 57-466).
 * Adjacency from tets (4 ids), tris (3 ids) and edges (2 ids). */
or_mas* or_mas_create(int32_t        vert_num,
                      const int32_t* part_ids,
                      const int32_t* tets,
                      int64_t        n_tets,
                      const int32_t* tris,
                      int64_t        n_tris,
                      const int32_t* edges,
                      int64_t        n_edges);
void or_mas_destroy(or_mas* m);

/* sizes: [vert_num, map_nodes, level_num, total_clusters(slots), level1_begin] */
void or_mas_info(const or_mas* m, int32_t info[5]);
/* tables (any may be NULL): part_to_real[map_nodes], real_to_part[vert_num],
 * going_next[padded*L], level_sizes[2*(L+1)] (x,y pairs), coarse_tables[vert_num*6],
 * fine_connect[vert_num], prefix_orig[map_nodes/16] */
void or_mas_tables(const or_mas* m,
                   int32_t*      part_to_real,
                   int32_t*      real_to_part,
                   int32_t*      going_next,
                   int32_t*      level_sizes,
                   int32_t*      coarse_tables,
                   uint32_t*     fine_connect,
                   int32_t*      prefix_orig);
int64_t or_mas_going_next_size(const or_mas* m);

/* set_preconditioner (mas_preconditioner_engine.cu:2281-2453): scatter k1/k2
 * (530-938) then invert (1077-1175).  inv_fp32 != 0: invert in float the way
 * the reference does (sweep, float storage); else invert in double. */
int or_mas_assemble(or_mas*        m,
                    int64_t        U,
                    const int32_t* rows,
                    const int32_t* cols,
                    const double*  vals,
                    int32_t        blk_off,
                    int            inv_fp32);
/* cluster Hessians (136 blocks x 9 doubles per cluster) / inverses */
int64_t or_mas_num_clusters(const or_mas* m);
void    or_mas_cluster_hessians(const or_mas* m, double* out);
void    or_mas_cluster_inverses(const or_mas* m, double* out);

/* apply (mas_preconditioner_engine.cu:2887-2960, kernels 1603-1951):
 * z = sum_L P_L^T C_L^{-1} P_L r on the subsystem range; fp32 != 0 keeps the
 * multi-level R/Z in float as the reference does.  active_levels 0 = all. */
void or_mas_apply(const or_mas* m, const double* r, double* z, int fp32, int active_levels);

/* ------------------------------------------------------------------ PCG */
typedef struct or_precond
{
    int32_t kind; /* 0 none (z = r), 1 diag3, 2 diag12, 3 mas */
    int32_t blk_off;
    int32_t blk_cnt;      /* 3x3 block rows covered                    */
    const double* data;   /* diag3: 9/blk; diag12: 144/body            */
    const or_mas* mas;
    int32_t mas_fp32;
    int32_t mas_active_levels;
} or_precond;

typedef struct or_pcg_result
{
    int64_t iters;       /* as the reference reports (linear_fused_pcg.cu:1994) */
    int64_t iters_exact; /* iteration whose check passed                        */
    double  rz0, rz_final, rel_res, rel_res_true;
    int32_t converged, bypassed;
    double  seconds;
} or_pcg_result;

/* LinearFusedPCG::fused_pcg (linear_system/linear_fused_pcg.cu:1782-2004) with
 * the per-iteration chain of run_iteration (1374-1594): x0 = 0, r = b, z = P r,
 * p = z, PSD guard, stop when |rz|/|rz0| <= tol (criterion 0) or ||r||/||b|| <= tol
 * (criterion 1), at most max_iter = ratio * n iterations, checks every
 * check_interval.  threads > 1 runs the SpMV and vector work on that many
 * threads (the timed CPU baseline).  max_iter_cap > 0 stops after that many
 * iterations (a bounded timing sample). */
int or_pcg(int32_t           B,
           int64_t           U,
           const int32_t*    rows,
           const int32_t*    cols,
           const double*     vals,
           const or_precond* pcs,
           int32_t           n_pcs,
           const double*     b,
           double*           x,
           int32_t           criterion,
           double            tol,
           double            max_iter_ratio,
           int32_t           check_interval,
           int32_t           psd_fallback,
           int32_t           threads,
           int64_t           max_iter_cap,
           or_pcg_result*    res);

#ifdef __cplusplus
}
#endif

#endif
