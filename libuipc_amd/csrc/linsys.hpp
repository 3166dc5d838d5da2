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
// The assembled system.
//   upper  sorted, de-duplicated upper-triangular 3x3 BCOO (the reference's
//          bcoo_A, linear_system/global_linear_system.h:313) with its BSR row
//          offsets; FP64 values.  Exported for parity and read by the
//          preconditioner assemblies.
//   full   both triangles as BSR (full_ptr, fcols, values), FP64 or FP32:
//          the stream the SpMV reads.
// HBM layout: indices int32.  Full-BSR values as 9 planes (plane q holds
// entry q of every block, column-major entry order), so that the lanes of a
// wave read contiguous bytes per plane; upper values per block (upper_at).
// entry q (column-major, r + 3 c) of upper block u: the upper BCOO keeps each
// block's 9 scalars together, as the reference's bcoo_A does (its consumers
// gather single blocks: the full-BSR transpose, the preconditioner assemblies)
__host__ __device__ __forceinline__ int64_t upper_at(int64_t u, int q)
{
    return 9 * u + q;
}

struct BsrMatrix
{
    int32_t       B      = 0;  // block rows
    int64_t       U      = 0;  // unique upper blocks (host copy, valid after assemble)
    DBuf<int32_t> rows, cols, row_ptr;
    DBuf<double>  vals;  // 9 per block, blocks consecutive (upper_at)
    // full BSR
    int64_t       fstride = 0;  // full plane stride (>= 2U - diagonal blocks)
    DBuf<int32_t> full_ptr, fcols;
    DBuf<double>  fvals;    // 9 * fstride (FP64 systems)
    DBuf<float>   fvals32;  // 9 * fstride (FP32 systems)
    bool          has32 = false;
    // SpMV chunks: chunk c = rows [chunk_row[c], chunk_row[c + 1]), the rows
    // whose first full-BSR block lies in [c kChunkStep, (c + 1) kChunkStep)
    DBuf<int32_t> chunk_row;  // nchunks + 1
    int32_t       nchunks = 0;

    // Window layout (the SpMV's stream).  The block rows are put in an order
    // of locality (wpos / wreal; the MAS partition's part order when a MAS
    // preconditioner covers them) and cut into windows of 2^wshift
    // consecutive positions.  An upper block whose two rows fall in one
    // window is stored once ("inner": word = bit 31 | li << 15 | lj, the two
    // rows' offsets in the window) and the SpMV adds both A x_j to y_i and
    // A^T x_i to y_j inside the window; a block whose rows fall in different
    // windows is stored with each row ("outer": word = the other row's
    // position, values transposed for the lower copy).  Rows are listed by position (wptr);
    // a row holds its upper blocks in column order, then its outer lower
    // blocks in row order.  Values as 9 planes of stride wstride.
    DBuf<int32_t> wpos, wreal;  // B: position of every row, row at every position
    DBuf<int32_t> wptr;         // B + 1, over positions
    DBuf<int32_t> wwords;
    DBuf<double>  wvals;    // 9 * wstride (FP64 systems)
    DBuf<float>   wvals32;  // 9 * wstride (FP32 systems)
    DBuf<int32_t> wchunk_lo, wchunk_hi;  // per window: lowest row and highest row + 1
    int64_t       wstride   = 0;
    int64_t       wblocks   = 0;  // host copy, valid after the build
    int32_t       wshift    = 10;
    int32_t       wshape    = 0;  // the window kernel's launch shape the layout was built for
    int32_t       wnchunks  = 0;
    bool          has_win   = false;
    bool          has_full  = false;  // the full BSR of the current assembly is built
    uint64_t      win_order = ~0ull;  // the row order the layout was built for
};
constexpr int kChunkStep = 512;  // nominal blocks per SpMV chunk
constexpr int kWinShiftMax = 15;  // window offsets are 15-bit fields of the inner words

// ----------------------------------------------------------------- converter
// Sort keys of (row, col) pairs.  Wide: row * B + col on 64 bits, the
// reference's key.  Narrow: when every kept block lies within `band` of the
// diagonal and the rows and offsets fit in 32 bits, (row << bb) | (col - row)
// on 32 bits -- the same order, sorted in fewer passes of two thirds the
// bytes.  The transposed (lower) entries of the off-diagonal upper blocks sort
// by (col, row): col * B + row, or (col << bb) | (row - col + band).
struct KeyCodec
{
    uint64_t B    = 0;
    int      bb   = 0;  // 0: wide keys
    int32_t  band = 0;
    __host__ __device__ __forceinline__ uint64_t up(int32_t r, int32_t c) const
    {
        return bb ? ((uint64_t)r << bb) | (uint64_t)(c - r) : (uint64_t)r * B + (uint64_t)c;
    }
    __host__ __device__ __forceinline__ uint64_t lo(int32_t r, int32_t c) const
    {
        return bb ? ((uint64_t)c << bb) | (uint64_t)(r - c + band) : (uint64_t)c * B + (uint64_t)r;
    }
    __host__ __device__ __forceinline__ void up_rc(uint64_t k, int32_t& r, int32_t& c) const
    {
        if(bb)
        {
            r = (int32_t)(k >> bb);
            c = r + (int32_t)(k & ((1ull << bb) - 1ull));
        }
        else
        {
            r = (int32_t)(k / B);
            c = (int32_t)(k - (uint64_t)r * B);
        }
    }
    // a lower entry's row (the upper block's column) and column (its row)
    __host__ __device__ __forceinline__ void lo_rc(uint64_t k, int32_t& lr, int32_t& lc) const
    {
        if(bb)
        {
            lr = (int32_t)(k >> bb);
            lc = lr - band + (int32_t)(k & ((1ull << bb) - 1ull));
        }
        else
        {
            lr = (int32_t)(k / B);
            lc = (int32_t)(k - (uint64_t)lr * B);
        }
    }
};


// MatrixConverter<Float,3>::convert_sym (algorithm/details/matrix_converter.inl:1549-1699)
class MatrixConverter
{
  public:
    // build_full: also the full BSR (the SpMV's other layout).  FP32 triplet
    // values (the precision arm's producers) are summed in FP64 like FP64 ones
    void convert_sym(int32_t        B,
                     int64_t        T,
                     const int32_t* rows,
                     const int32_t* cols,
                     const double*  vals,
                     BsrMatrix&     out,
                     bool           want32,
                     bool           build_full,
                     double         reserve_ratio,
                     hipStream_t    s)
    {
        convert(B, T, rows, cols, vals, false, out, want32, build_full, reserve_ratio, s);
    }
    void convert_sym(int32_t        B,
                     int64_t        T,
                     const int32_t* rows,
                     const int32_t* cols,
                     const float*   vals,
                     BsrMatrix&     out,
                     bool           want32,
                     bool           build_full,
                     double         reserve_ratio,
                     hipStream_t    s)
    {
        convert(B, T, rows, cols, vals, true, out, want32, build_full, reserve_ratio, s);
    }
    ~MatrixConverter();
    // the full BSR of `A` (assembled by the last convert_sym)
    void build_full(BsrMatrix& A, bool want32, double reserve_ratio, hipStream_t s);
    // the window layout of `A` (assembled by the last convert_sym) for the
    // row order in A.wpos / A.wreal and a window kernel shape
    void build_window(BsrMatrix& A, int shape, bool want32, double reserve_ratio, hipStream_t s);

  private:
    void convert(int32_t        B,
                 int64_t        T,
                 const int32_t* rows,
                 const int32_t* cols,
                 const void*    vals,
                 bool           vals32,
                 BsrMatrix&     out,
                 bool           want32,
                 bool           build_full,
                 double         reserve_ratio,
                 hipStream_t    s);
    template <int MODE, typename KeyT>
    int compact_sort(int64_t        n,
                     const int32_t* rows,
                     const int32_t* cols,
                     int32_t        B,
                     int64_t*       d_count,
                     double         ratio,
                     hipStream_t    s);

    DBuf<uint64_t> keys_[2];
    DBuf<int32_t>  srcs_[2];
    DBuf<uint32_t> hist_, chunk_;
    DBuf<uint32_t> tile_cnt_;
    DBuf<int32_t>  seg_start_, lower_rows_, lower_ptr_;
    DBuf<int64_t>  counts_;  // [0] kept triplets, [1] U, [2] off-diagonal upper blocks, [3] full, [4] window blocks
    int            tcur_ = 0;  // buffer of the transposed (lower) sort
    KeyCodec       codec_{};
    int            key_bits_ = 0;
    bool           counted_  = false;  // the triplet count of this conversion is done
    template <typename KeyT>
    KeyT* key_ptr(int k)
    {
        return reinterpret_cast<KeyT*>(keys_[k].data());
    }
    KeyCodec codec() const { return codec_; }
    int64_t*       pinned_ = nullptr;
};

// Stable LSD radix sort (8-bit digits) of *d_n (device count, at most nmax)
// (key, value) pairs on the low `bits` key bits; input in (k0, v0), (k1, v1)
// is the other buffer.  Returns 0 or 1: the buffer that holds the result.
int radix_sort_pairs(uint64_t*       k0,
                     int32_t*        v0,
                     uint64_t*       k1,
                     int32_t*        v1,
                     const int64_t*  d_n,
                     int64_t         nmax,
                     int             bits,
                     DBuf<uint32_t>& hist,
                     DBuf<uint32_t>& chunk,
                     hipStream_t     s);
// bits needed for keys up to max_key
int radix_key_bits(uint64_t max_key);
// in-place exclusive scan of n uint32 values; *d_total (device, may be NULL) = sum
void exclusive_scan_u32(uint32_t* data, int64_t n, DBuf<uint32_t>& chunk, int64_t* d_total, hipStream_t s);

// ----------------------------------------------------------------- SpMV
// Spmv::rbk_sym_spmv_dot (linear_system/spmv.cu:1023-1159)
struct SpmvScratch
{
    DBuf<double> partials;
    DBuf<float4> x32;  // FP32 full-BSR systems: the packed copy of x when the caller has none
    int          grid     = 0;  // full-BSR kernel
    int          grid_win = 0;  // window kernel
    bool         use32    = false;
};
// the SpMV layout: the window layout unless MI_SPMV_LAYOUT=full selects the full BSR
bool spmv_use_window();
// the window kernel's launch shapes: how many, the default (MI_SPMV_WIN
// overrides it), each one's threads, window rows and blocks per lane and
// stage, and log2 of its window rows (what the layout is built for)
int  spmv_shape_count();
int  spmv_default_shape();
bool spmv_shape_info(int shape, int* threads, int* rows, int* per, int* vec);
int  spmv_shape_shift(int shape);
// the layout codes beside the window shapes (mi_sys_set_spmv_shape)
constexpr int spmv_shape_auto = -1;  // chosen by the system's size
constexpr int spmv_shape_full = -2;  // the full BSR
// the layout a system of B block rows gets when none is set (use32: FP32
// value planes): a window shape index, or spmv_shape_full
int spmv_auto_shape(int64_t B, bool use32);
// workgroup size of the SpMV kernel that streams A's layout
int spmv_threads(const BsrMatrix& A);
// the SpMV's work items (windows or chunks) and the row range [lo, hi) each touches
inline int32_t spmv_items(const BsrMatrix& A)
{
    return A.has_win ? A.wnchunks : A.nchunks;
}
inline const int32_t* spmv_item_lo(const BsrMatrix& A)
{
    return A.has_win ? A.wchunk_lo.data() : A.chunk_row.data();
}
inline const int32_t* spmv_item_hi(const BsrMatrix& A)
{
    return A.has_win ? A.wchunk_hi.data() : A.chunk_row.data() + 1;
}
// allocates the SpMV's reduction scratch and fixes its launch shape for one
// value precision (must run outside graph capture)
void spmv_prepare(SpmvScratch& scr, bool use32);
// y = A x (overwritten); partials[g] = per-workgroup x^T A x; returns the grid g
// x32: the FP32 SpMV's padded float4 copy of x (FP32 systems; packed into the
// scratch when null).  With `active`, only the rows of the listed chunks
// (ascending, *d_nactive of them) are computed; the other rows of y keep
// their values.  `ordered`: x and y are in the window layout's row order
// (A.has_win; no `active` list)
int spmv_full(const BsrMatrix& A, bool use32, const double* x, const float4* x32, double* y, double* partials,
              const int32_t* d_conv, SpmvScratch& scr, hipStream_t s, const int32_t* active = nullptr,
              const int64_t* d_nactive = nullptr, bool ordered = false);
// *out = the sum of partials[0 .. n) (one workgroup, a fixed order)
void sum_partials(const double* partials, int n, double* out, hipStream_t s);
// y += a A x; *d_dot = x^T A x when d_dot != null
void spmv_accumulate_dot(const BsrMatrix& A, bool use32, double a, const double* x, double* y, double* d_dot,
                         SpmvScratch& scr, hipStream_t s);
// y = a A x + b y (Spmv::rbk_sym_spmv)
void spmv_axpby(const BsrMatrix& A, bool use32, double a, const double* x, double b, double* y, SpmvScratch& scr,
                hipStream_t s);
// the fixed scratch of spmv_accumulate_dot_noalloc: its device address on
// `device` (which must be current), looked up once
void spmv_noalloc_init(int device);
// y += A x, *d_dot = x^T A x (when d_dot != null) with the partials in the
// fixed scratch of `device`: no allocation; FP32 systems up to 2^18 block rows
void spmv_accumulate_dot_noalloc(const BsrMatrix& A, bool use32, const double* x, double* y, double* d_dot, int device,
                                 hipStream_t s);

// ----------------------------------------------------------------- vectors
void   fill_f64(double* p, double v, int64_t n, hipStream_t s);
// dst[3 k + c] = src[3 idx[k] + c] for k < rows: vectors between the caller's
// vertex order and the SpMV's row order (idx = wreal in, wpos out)
void   gather_rows3(double* dst, const double* src, const int32_t* idx, int64_t rows, hipStream_t s);
// y += a x
void   axpy_f64(double a, const double* x, double* y, int64_t n, hipStream_t s);
void   fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t s);
// FEM subsystem steps around the solve (subsystem.hip)
void subsystem_fix_vertices(const int32_t* rows,
                            const int32_t* cols,
                            double*        vals,
                            int64_t        T,
                            int32_t        block_offset,
                            const int32_t* is_fixed,
                            int32_t        V,
                            double*        gradient,
                            hipStream_t    s);
void subsystem_distribute(const double* x, double* dxs, int64_t n, hipStream_t s);
// Hessian producers (producers.hip)
void produce_kinetic(int32_t V, const double* mass, int32_t vert_offset, int32_t* rows, int32_t* cols, double* vals,
                     hipStream_t s);
void produce_snh(int64_t        n_tets,
                 const int32_t* tets,
                 const double*  X,
                 const double*  x,
                 double         mu,
                 double         lambda,
                 double         dt,
                 int32_t        vert_offset,
                 int32_t*       rows,
                 int32_t*       cols,
                 double*        vals,
                 hipStream_t    s);
void produce_abd(int64_t        n,
                 const double*  q,
                 const double*  dyadic,
                 const double*  volume,
                 const double*  kappa,
                 const int32_t* is_fixed,
                 double         dt,
                 int32_t        body_offset,
                 int32_t*       rows,
                 int32_t*       cols,
                 double*        vals,
                 hipStream_t    s);
void produce_bending(int64_t        n,
                     const int32_t* hinges,
                     const double*  X,
                     const double*  x,
                     double         kappa,
                     double         dt,
                     int32_t        vert_offset,
                     int32_t*       rows,
                     int32_t*       cols,
                     double*        vals,
                     hipStream_t    s);
void produce_bw_shell(int64_t        n,
                      const int32_t* tris,
                      const double*  X,
                      const double*  x,
                      double         stretch,
                      double         shear,
                      double         rate,
                      double         dt,
                      int32_t        vert_offset,
                      int32_t*       rows,
                      int32_t*       cols,
                      double*        vals,
                      hipStream_t    s);
void energy_kinetic(int32_t V, const double* x, const double* xt, const double* m, double* energy,
                    DBuf<double>& part, hipStream_t s);
void energy_snh(int64_t        n_tets,
                const int32_t* tets,
                const double*  X,
                const double*  x,
                double         mu,
                double         lambda,
                double         dt,
                double*        energy,
                DBuf<double>&  part,
                hipStream_t    s);
void step_forward(const double* x0, const double* dx, double alpha, double* x, int64_t n, hipStream_t s);
void produce_contact(int64_t        n,
                     const int32_t* stencils,
                     const double*  g,
                     const double*  kappa,
                     double         dt,
                     int32_t        vert_offset,
                     int32_t*       rows,
                     int32_t*       cols,
                     double*        vals,
                     hipStream_t    s);
void   copy_f64(double* dst, const double* src, int64_t n, hipStream_t s);
double dot_f64_host(const double* x, const double* y, int64_t n, hipStream_t s);
double norm_f64_host(const double* x, int64_t n, hipStream_t s);
float  dot_f32_host(const float* x, const float* y, int64_t n, hipStream_t s);
float  norm_f32_host(const float* x, int64_t n, hipStream_t s);

// ----------------------------------------------------------------- preconditioners
struct BatchScalars;
struct PcgScalars;
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
    // a preconditioner covering the whole vector may also write the PCG's
    // r.z and r.r partials (pairs, at most max_parts); returns their count,
    // or -1 when it does not (the caller then runs the dot kernel)
    virtual int apply_dot(const double*, double*, const int32_t*, double*, int, hipStream_t) { return -1; }
    // a preconditioner covering the whole vector may take over the PCG's x, r
    // update (x += alpha p, r -= alpha Ap, alpha from the pAp partials) and
    // fold its own first pass over the new r into it; returns false when it
    // does not (the caller then runs pcg_update_xr)
    virtual bool update_xr(PcgScalars*, const double*, int, double*, const double*, double*, const double*,
                           hipStream_t, double* /*zero*/ = nullptr, int /*nzero*/ = 0)
    {
        return false;
    }
    // a preconditioner covering the whole vector whose apply_dot adds its
    // r.z and r.r into dot slots with atomics: how many doubles (pairs x 2)
    // must be zero when it starts (the caller clears them in the x, r
    // update); 0 when apply_dot writes its partials without atomics
    virtual int dot_slots() const { return 0; }
    // MAS only: levels used by apply (MASPreconditionerEngine::set_active_level_num)
    virtual bool set_active_levels(int) { return false; }
    // MAS only, batched solves: a level count per system (block-row ranges)
    virtual bool set_system_levels(int32_t, const int32_t*, const int32_t*) { return false; }
    // MAS only: the rank of every vertex of the range in the partition's part
    // order (parts in id order, then the unpartitioned vertices): the order
    // of locality the SpMV's window layout puts the rows in
    virtual const std::vector<int32_t>* spmv_rank() const { return nullptr; }
    // the vectors apply, apply_dot and update_xr see: the caller's vertex
    // order (false) or the SpMV's row order (true, the PCG's own vectors).
    // Only a preconditioner with an spmv_rank reorders its range; the others
    // see the same order either way
    virtual void set_row_order(bool) {}
    // batched solves: apply may skip the work of systems that have stopped
    // (their r no longer changes, so neither does z); seg = the systems'
    // first block rows (device); seg == nullptr switches the skipping off
    virtual void skip_stopped(const int32_t*, int32_t, const BatchScalars*, hipStream_t) {}
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
// Device scalars of one solve (LinearFusedPCG's d_rz, d_pAp, d_rz_new, d_beta,
// d_converged, d_rz_tol).  Reductions reach them without fences: a kernel
// writes per-workgroup partials, and every workgroup of the next kernel sums
// the partials in the same order; workgroup 0 records the result here.
struct PcgScalars
{
    double  rz;          // r^T z belonging to the current p (beta's denominator)
    double  rz_pending;  // r^T z of the last iteration (alpha's numerator)
    double  rz_new;      // same value, as reported
    double  pAp;
    double  alpha;
    double  beta;
    double  rr;      // r^T r of the last iteration
    double  bb;      // b^T b
    double  rz_tol;  // tol * |rz0|          (criterion 0)
    double  rr_tol;  // tol^2 * ||b||^2      (criterion 1)
    int32_t converged;
    int32_t criterion;
    int32_t iter;
    int32_t conv_iter;
};

// the fixed-order sum of a reduction's partials (stride apart), redone by every
// workgroup of NT threads of the consuming kernel (no fence, no atomics,
// deterministic); `lds` holds NT / 64 values
template <int NT>
__device__ __forceinline__ double pcg_sum_partials(const double* part, int g, int stride, double* lds)
{
    double acc = 0.0;
    for(int k = threadIdx.x; k < g; k += NT)
        acc += part[stride * k];
    return block_sum<NT>(acc, lds);
}

// A thread's first K partials, loaded ahead of their re-sum: a kernel issues
// its own operand loads between load() and sum(), so that the loads' latency
// overlaps the sum's chain of shuffles and barriers.  The sum adds in
// pcg_sum_partials' order (entries past g hold 0.0); partials beyond K * NT
// are loaded by sum() itself.  The loads go through a buffer resource of g
// entries, whose range check returns 0.0 past the end: no branch, so the
// compiler does not wait for them before the caller's own loads.
constexpr uint32_t kBufferRsrcWord3 = 0x00020000;  // raw (untyped) buffer, gfx9 family
__device__ __forceinline__ __amdgpu_buffer_rsrc_t partials_rsrc(const double* part, int g, int per)
{
    return __builtin_amdgcn_make_buffer_rsrc((void*)part, 0, g * per * (int)sizeof(double), kBufferRsrcWord3);
}
template <int NT, int K = 4>
struct PartialsAhead
{
    double q[K];
    __device__ __forceinline__ void load(const double* part, int g)
    {
        const __amdgpu_buffer_rsrc_t rs = partials_rsrc(part, g, 1);
#pragma unroll
        for(int j = 0; j < K; ++j)
            q[j] = __builtin_bit_cast(
                double, __builtin_amdgcn_raw_buffer_load_b64(rs, ((int)threadIdx.x + j * NT) * 8, 0, 0));
    }
    __device__ __forceinline__ double sum(const double* part, int g, double* lds) const
    {
        double acc = 0.0;
#pragma unroll
        for(int j = 0; j < K; ++j)
            acc += q[j];
        for(int k = (int)threadIdx.x + K * NT; k < g; k += NT)
            acc += part[k];
        return block_sum<NT>(acc, lds);
    }
};

// the same for (r.z, r.r) partial pairs: both sums in one pass, each in
// pcg_sum_partials' order; `lds` holds 2 * NT / 64 values
template <int NT, int K = 2>
struct PairsAhead
{
    double2 q[K];
    __device__ __forceinline__ void load(const double* part, int g)
    {
        const __amdgpu_buffer_rsrc_t rs = partials_rsrc(part, g, 2);
#pragma unroll
        for(int j = 0; j < K; ++j)
            q[j] = __builtin_bit_cast(
                double2, __builtin_amdgcn_raw_buffer_load_b128(rs, ((int)threadIdx.x + j * NT) * 16, 0, 0));
    }
    __device__ __forceinline__ void sum(const double* part, int g, double* lds, double& s0, double& s1) const
    {
        double a = 0.0, b = 0.0;
#pragma unroll
        for(int j = 0; j < K; ++j)
        {
            a += q[j].x;
            b += q[j].y;
        }
        for(int k = (int)threadIdx.x + K * NT; k < g; k += NT)
        {
            const double2 w = reinterpret_cast<const double2*>(part)[k];
            a += w.x;
            b += w.y;
        }
        block_sum2<NT>(a, b, lds);
        s0 = a;
        s1 = b;
    }
};

template <int NT>
__device__ __forceinline__ void pcg_sum_partials2(const double* part, int g, double* lds, double& s0, double& s1)
{
    PairsAhead<NT> ahead;
    ahead.load(part, g);
    ahead.sum(part, g, lds, s0, s1);
}

// alpha = rz / pAp from the SpMV's pAp partials; workgroup 0 records it.
// rz is sc->rz_pending, which a kernel may read at its start (one scalar load
// fewer on the chain after the partials' re-sum)
template <int NT>
__device__ __forceinline__ double pcg_alpha_of(PcgScalars* sc, double pAp, double rz)
{
    const double alpha = rz / pAp;
    if(blockIdx.x == 0 && threadIdx.x == 0)
    {
        sc->pAp   = pAp;
        sc->alpha = alpha;
        sc->rz    = rz;
    }
    return alpha;
}
template <int NT>
__device__ __forceinline__ double pcg_alpha_of(PcgScalars* sc, const double* pap_part, int gp, double* lds)
{
    return pcg_alpha_of<NT>(sc, pcg_sum_partials<NT>(pap_part, gp, 1, lds), sc->rz_pending);
}

// ---- batched PCG: independent systems on consecutive block-row ranges
// Tiles of at most kBatchTile DoFs (whole block rows) that never cross a
// system boundary; every per-system sum runs over its tiles in order.
constexpr int64_t kBatchTile = 3 * 1365;
// ----------------------------------------------------------------- learned initial guess
void ws_gemm(int32_t M, int32_t N, int32_t K, const float* A, int64_t sa_m, int64_t sa_k, const float* B,
             int64_t sb_k, int64_t sb_n, float* C, int64_t ldc, bool accumulate, const float* bias, bool relu,
             hipStream_t s);
void ws_aggregate(int32_t n, int32_t F, const int32_t* ptr, const int32_t* idx, const float* w, const float* H,
                  float* Z, hipStream_t s);
void ws_relu_backward(int64_t n, const float* pre, float* d, hipStream_t s);
void ws_colsum(int32_t n, int32_t F, const float* X, float* out, bool accumulate, hipStream_t s);
void ws_adam(int64_t n, float* w, const float* g, float* m, float* v, float lr, float beta1, float beta2, float eps,
             int32_t step, hipStream_t s);
void ws_block3(int32_t n, const double* blocks, double scale, const float* v_f32, double* y_f64, const double* base,
               const double* v_f64, float* y_f32, hipStream_t s);

struct BatchTiles
{
    int32_t        n        = 0;        // tiles
    const int64_t* lo       = nullptr;  // n + 1 DoF boundaries
    const int32_t* sys      = nullptr;  // system of every tile
    const int32_t* sys_tile = nullptr;  // nseg + 1: first tile of every system
    const int64_t* sys_lo   = nullptr;  // nseg + 1: first DoF of every system
};
struct BatchScalars
{
    double  rz, rz_pending, rz_new, pAp, alpha, beta, rr, bb, rz_tol, rr_tol;
    int64_t iter, conv_iter, max_iter;
    int32_t criterion, converged, stopped;
    int32_t bypass;  // the PSD guard took this system's preconditioner out: z = r
};
void batch_tile_dot(const BatchTiles& t, const double* a, const double* b, double* ab, double* aa, hipStream_t s);
// on[k] = 1 while cluster k (slots [bank k, bank (k + 1)) of part_to_real,
// local vertices, -1 = empty) holds a vertex of a running system
void batch_cluster_flags(const int32_t* part_to_real, int64_t clusters, int bank, int32_t row_off, const int32_t* seg,
                         int32_t nseg, const BatchScalars* sc, int32_t* on, hipStream_t s);
// the SpMV chunks that hold rows of a running system, ascending, into list;
// *d_count = their number (seg: the systems' first block rows, device; chunk
// c's rows lie in [lo[c], hi[c]))
void batch_active_chunks(const int32_t* lo, const int32_t* hi, int32_t nchunks, const int32_t* seg, int32_t nseg,
                         const BatchScalars* sc, DBuf<uint32_t>& flags, DBuf<uint32_t>& scan_chunk, int32_t* list,
                         int64_t* d_count, hipStream_t s);
// bypass: per-system flags (device, may be null = none)
void batch_init(const BatchTiles& t, BatchScalars* sc, int32_t nseg, const double* rz, const double* rr,
                int criterion, double tol, double max_iter_ratio, const int32_t* bypass, hipStream_t s);
// z = r on the tiles of the systems flagged in `flags` (device, per system), or
// with flags == null of the running systems whose preconditioner is bypassed
void batch_bypass_z(const BatchTiles& t, const BatchScalars* sc, const int32_t* flags, const double* r, double* z,
                    hipStream_t s);
void batch_alpha_xr(const BatchTiles& t, BatchScalars* sc, int32_t nseg, const double* pap, double* x,
                    const double* p, double* r, const double* Ap, hipStream_t s);
void batch_beta_p(const BatchTiles& t, BatchScalars* sc, int32_t nseg, const double* rz, const double* rr,
                  double* p, const double* z, float4* p32, hipStream_t s);

// allocates the PCG reductions' partials (must run before any pointer into them is taken)
void pcg_prepare(SpmvScratch& scr);
void pcg_init_vectors(double* x, double* r, const double* b, int64_t n, hipStream_t s);
// partials of (r.z, r.r) -> part[2 g]; returns g
int pcg_dot_partials(const double* r, const double* z, int64_t n, double* part, const int32_t* conv,
                     SpmvScratch& scr, hipStream_t s);
// PcgScalars from the initial partials (rz0 = r.z, bb = r.r with r = b)
// b^T b of a solve started from a given x0 (the relative-residual reference)
void pcg_set_bb(PcgScalars* sc, double bb, double tol, hipStream_t s);
void pcg_init_scalars(PcgScalars* sc, const double* part, int g, int criterion, double tol, hipStream_t s);
// x += alpha p, r -= alpha Ap with alpha = rz / sum(pAp partials); zero[0,
// nzero) is cleared on the way (the preconditioner's dot slots, see
// LocalPreconditioner::dot_slots)
void pcg_update_xr(PcgScalars* sc, const double* pap_part, int gp, double* x, const double* p, double* r,
                   const double* Ap, int64_t n, hipStream_t s, double* zero = nullptr, int nzero = 0);
// the same, fused with z = D^-1 r on [blk_off, blk_off + blk_cnt) (z = r elsewhere)
// and the (r.z, r.r) partials -> part; returns the grid
int pcg_update_xr_diag3(PcgScalars* sc, const double* pap_part, int gp, double* x, const double* p, double* r,
                        const double* Ap, double* z, const double* diag_inv, int64_t blk_off, int64_t blk_cnt,
                        int64_t n, double* part, SpmvScratch& scr, hipStream_t s);
// convergence / beta from the (r.z, r.r) partials, then p = z + beta p
// p = z + beta p; with p32 also the FP32 SpMV's padded float4 copy of p
void pcg_update_p(PcgScalars* sc, const double* part, int gd, double* p, const double* z, int64_t n, float4* p32,
                  hipStream_t s);
// x -> one padded float4 per block row (the FP32 SpMV's gather vector)
void pack_f32x4(const double* x, float4* x32, int64_t nb, hipStream_t s);
void pcg_copy_if_active(PcgScalars* sc, double* dst, const double* src, int64_t n, hipStream_t s);
}  // namespace mi
