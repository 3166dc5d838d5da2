// Internal C++ interfaces of the MI355X linear-solve library.  The class and
// method names follow the reference's backend (src/backends/cuda/) so that a
// maintainer can map each one to the component it replaces.
#pragma once
#include <memory>
#include <vector>

#include "common.hpp"

namespace mi
{
// ----------------------------------------------------------------- matrix
// The assembled system: sorted, de-duplicated upper-triangular 3x3 BCOO
// (the reference's bcoo_A, linear_system/global_linear_system.h:313) plus the
// BSR row offsets.  HBM layout: rows/cols int32[U]; values as 9 planes
// (plane q holds entry q of every block, column-major entry order), so that a
// wave's 64 lanes read 512 contiguous bytes per plane.
struct BsrMatrix
{
    int32_t          B      = 0;  // block rows
    int64_t          U      = 0;  // unique upper blocks (host copy, valid after assemble)
    int64_t          stride = 0;  // plane stride (>= U)
    DBuf<int32_t>    rows, cols, row_ptr;
    DBuf<double>     vals;    // 9 * stride
    DBuf<float>      vals32;  // 9 * stride, FP32 copy for the precision arm
    bool             has32 = false;
};

// ----------------------------------------------------------------- converter
// MatrixConverter<Float,3>::convert_sym (algorithm/details/matrix_converter.inl:1549-1699)
class MatrixConverter
{
  public:
    void convert_sym(int32_t        B,
                     int64_t        T,
                     const int32_t* rows,
                     const int32_t* cols,
                     const double*  vals,
                     BsrMatrix&     out,
                     bool           want32,
                     double         reserve_ratio,
                     hipStream_t    s);

  private:
    DBuf<uint64_t> keys_[2];
    DBuf<int32_t>  srcs_[2];
    DBuf<uint32_t> hist_, chunk_;
    DBuf<uint32_t> tile_cnt_;
    DBuf<int32_t>  seg_start_;
    DBuf<int64_t>  counts_;  // [0] = m (upper triplets), [1] = U
    int64_t*       pinned_ = nullptr;

  public:
    ~MatrixConverter();
};

// ----------------------------------------------------------------- SpMV
// Spmv::rbk_sym_spmv_dot (linear_system/spmv.cu:1023-1159)
struct SpmvScratch
{
    DBuf<double>   partials;
    DBuf<uint32_t> ticket;
    int            grid = 0;
};
void spmv_sym_dot(const BsrMatrix& A,
                  bool             use32,
                  const double*    x,
                  double*          y,  // accumulated into: must be zero on entry
                  double*          d_dot,
                  double*          d_alpha_num,  // when non-null: *d_alpha_out = *d_alpha_num / dot
                  double*          d_alpha_out,
                  const int32_t*   d_converged,
                  SpmvScratch&     scr,
                  hipStream_t      s);

// allocates the SpMV's reduction scratch (must run outside graph capture)
void spmv_prepare(SpmvScratch& scr);

// y = a A x + b y (Spmv::rbk_sym_spmv)
void spmv_axpby(const BsrMatrix& A, double a, const double* x, double b, double* y, SpmvScratch& scr, hipStream_t s);

// ----------------------------------------------------------------- vectors
void fill_f64(double* p, double v, int64_t n, hipStream_t s);
void copy_f64(double* dst, const double* src, int64_t n, hipStream_t s);
double dot_f64_host(const double* x, const double* y, int64_t n, hipStream_t s);
double norm_f64_host(const double* x, int64_t n, hipStream_t s);
float  dot_f32_host(const float* x, const float* y, int64_t n, hipStream_t s);
float  norm_f32_host(const float* x, int64_t n, hipStream_t s);

// ----------------------------------------------------------------- preconditioners
// LocalPreconditioner (linear_system/local_preconditioner.h:9-49)
class LocalPreconditioner
{
  public:
    virtual ~LocalPreconditioner() = default;
    int64_t dof_offset = 0;
    int64_t dof_count  = 0;
    // LocalPreconditioner::do_assemble
    virtual void assemble(const BsrMatrix& A, hipStream_t s) = 0;
    // LocalPreconditioner::do_apply on [dof_offset, dof_offset + dof_count)
    virtual void apply(const double* r, double* z, const int32_t* d_converged, hipStream_t s) = 0;
    // 3x3 diagonal-block preconditioners can be fused into the PCG update
    virtual const double* diag3_inverse() const { return nullptr; }
    // MAS only: host copies of the hierarchy and cluster matrices
    virtual bool mas_export(int32_t*, int32_t*, int32_t*, int32_t*, int32_t*, int32_t*, double*, double*)
    {
        return false;
    }
};

std::unique_ptr<LocalPreconditioner> make_diag3(int64_t dof_offset, int64_t dof_count);
std::unique_ptr<LocalPreconditioner> make_diag12(int64_t dof_offset, int64_t dof_count);
std::unique_ptr<LocalPreconditioner> make_mas(const mi_precond_desc& desc);

// ----------------------------------------------------------------- PCG
struct PcgScalars
{
    double   rz;      // r^T z of the previous iteration
    double   rz_new;  // r^T z of this iteration
    double   pAp;
    double   alpha;
    double   beta;
    double   rr;      // r^T r of this iteration
    double   rz_tol;  // tol * |rz0|          (criterion 0)
    double   rr_tol;  // tol^2 * ||b||^2      (criterion 1)
    int32_t  converged;
    int32_t  criterion;
    int32_t  iter;
    int32_t  conv_iter;
    uint32_t ticket[4];
};

void pcg_init_vectors(double* x, double* r, const double* b, int64_t n, hipStream_t s);
// p = z and PcgScalars init given rz0 (device reduction of r.z, r.r, b.b)
void pcg_dot_rz_rr(const double* r, const double* z, int64_t n, double* out3, SpmvScratch& scr, hipStream_t s);
void pcg_update_xr(PcgScalars* sc, double* x, const double* p, double* r, double* Ap, int64_t n, hipStream_t s);
void pcg_update_xr_diag3(PcgScalars*   sc,
                         double*       x,
                         const double* p,
                         double*       r,
                         double*       Ap,
                         double*       z,
                         const double* diag_inv,
                         int64_t       blk_off,
                         int64_t       blk_cnt,
                         int64_t       n,
                         SpmvScratch&  scr,
                         hipStream_t   s);
void pcg_dot_scalar(PcgScalars* sc, const double* r, const double* z, int64_t n, SpmvScratch& scr, hipStream_t s);
void pcg_update_p(PcgScalars* sc, double* p, const double* z, int64_t n, hipStream_t s);
void pcg_copy_if_active(PcgScalars* sc, double* dst, const double* src, int64_t n, hipStream_t s);
}  // namespace mi
