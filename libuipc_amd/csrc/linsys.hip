// Host orchestration and the C ABI (include/mi_linsys.h).
//
// mi_sys mirrors GlobalLinearSystem::Impl (linear_system/global_linear_system.cu):
// it owns the assembled matrix, the registered local preconditioners and the
// PCG work vectors.  PcgDriver mirrors LinearFusedPCG
// (linear_system/linear_fused_pcg.cu:773-2004): x0 = 0, the rz0 PSD guard,
// check_interval-iteration blocks replayed from a captured hipGraph, device-side
// alpha / beta / convergence scalars, one host read per block.
#include <algorithm>
#include <cmath>
#include <string>

#include "linsys.hpp"
#include "mas.hpp"

namespace mi
{
namespace
{
thread_local std::string g_last_error;
}
void set_last_error(const std::string& m)
{
    g_last_error = m;
}

int device_cu_count()
{
    static int cached = 0;
    if(cached == 0)
    {
        int dev = 0, cu = 0;
        if(hipGetDevice(&dev) == hipSuccess
           && hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cu > 0)
            cached = cu;
        else
            cached = 64;
    }
    return cached;
}
}  // namespace mi

using namespace mi;

struct mi_ctx
{
    int          device = 0;
    hipStream_t  stream = nullptr;
    DBuf<double> energy_part;  // per-workgroup partial energies
};

struct mi_sys
{
    mi_ctx*                                           ctx = nullptr;
    mi_sys_desc                                       desc{};
    BsrMatrix                                         A;
    MatrixConverter                                   conv;
    SpmvScratch                                       spmv_scr, dot_scr;
    std::vector<std::unique_ptr<LocalPreconditioner>> pcs;
    uint64_t                                          pcs_gen       = 0;  // bumped when pcs change
    uint64_t                                          order_gen     = 0;  // bumped when a preconditioner is added or rebuilt
    int32_t                                           spmv_shape    = spmv_shape_auto;  // or full, or a window shape
    bool                                              assembled     = false;
    bool                                              pcs_assembled = false;

    // library-owned triplets and gradient (mi_sys_update_extents)
    DBuf<int32_t> own_rows, own_cols;
    DBuf<double>  own_vals, own_b;
    int64_t       own_T = -1;  // triplets of the last reported extents (-1: none reported)

    // PCG state (LinearFusedPCG's r, z, p, Ap and device scalars)
    DBuf<double> r, z, p, Ap, tmp;
    DBuf<double> xo;   // x in the SpMV's row order (solves that iterate in it)
    DBuf<float4> p32;  // FP32 systems: the SpMV's packed copy of p
    PcgScalars*  d_sc = nullptr;
    PcgScalars*  h_sc = nullptr;  // pinned

    // captured block of iterations
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t     graph      = nullptr;
    struct GraphKey
    {
        const void* x = nullptr;
        const void* b = nullptr;
        int64_t     n = -1, U = -1;
        uint64_t    pcs_gen = ~0ull;
        int         ci = 0, fused = -1, bypass = -1, use32 = -1, ordered = -1;
        const void* bufs[16] = {};
        int64_t     wstride  = -1;
        int32_t     wshape   = -2;  // the SpMV kernel's launch shape (-1: full BSR)
        int32_t     items    = -1;  // the SpMV's work items (windows or chunks)
        hipStream_t stream  = nullptr;
        bool        operator==(const GraphKey& o) const
        {
            if(x != o.x || b != o.b || n != o.n || U != o.U || pcs_gen != o.pcs_gen || ci != o.ci || fused != o.fused
               || bypass != o.bypass || use32 != o.use32 || ordered != o.ordered || stream != o.stream
               || wstride != o.wstride || wshape != o.wshape || items != o.items)
                return false;
            for(int k = 0; k < 16; ++k)
                if(bufs[k] != o.bufs[k])
                    return false;
            return true;
        }
    } graph_key;
    bool graph_failed = false;

    // batched solves (mi_pcg_solve_batched): the tile table of the last
    // segmentation, per-system scalars, tile partials and their own graph
    struct Batch
    {
        std::vector<int32_t> seg_rows;
        DBuf<int64_t>        tile_lo, sys_lo;
        DBuf<int32_t>        tile_sys, sys_tile;
        BatchTiles           tiles;
        DBuf<BatchScalars>   sc;
        BatchScalars*        h_sc = nullptr;  // pinned, nseg
        size_t               h_cap = 0;
        DBuf<double>         part_a, part_b;
        std::vector<double>  h_part;
        // the SpMV chunks still holding rows of a running system
        DBuf<int32_t>        seg_dev, active;
        DBuf<uint32_t>       act_flags, act_scan;
        DBuf<int64_t>        n_active;
        DBuf<int32_t>        bypass;      // per-system PSD-guard flags of the current solve
        bool                 any_bypass = false;
        hipGraphExec_t       graph_exec = nullptr;
        hipGraph_t           graph      = nullptr;
        GraphKey             key;
        uint64_t             seg_gen = 0, key_seg_gen = ~0ull;
    } batch;

    ~mi_sys()
    {
        if(batch.graph_exec)
            (void)hipGraphExecDestroy(batch.graph_exec);
        if(batch.graph)
            (void)hipGraphDestroy(batch.graph);
        if(batch.h_sc)
            (void)hipHostFree(batch.h_sc);
        if(graph_exec)
            (void)hipGraphExecDestroy(graph_exec);
        if(graph)
            (void)hipGraphDestroy(graph);
        if(d_sc)
            (void)hipFree(d_sc);
        if(h_sc)
            (void)hipHostFree(h_sc);
    }

    int64_t n() const { return 3 * (int64_t)desc.block_rows; }

    // GlobalLinearSystem::Impl::apply_preconditioner (global_linear_system.cu:546-590)
    void apply_preconditioner(const double* rr, double* zz, const int32_t* conv_flag, hipStream_t s)
    {
        int64_t covered = 0;
        for(auto& pc : pcs)
            covered += pc->dof_count;
        if(covered < n())
            copy_f64(zz, rr, n(), s);  // identity for DoFs no preconditioner covers
        for(auto& pc : pcs)
            pc->apply(rr + pc->dof_offset, zz + pc->dof_offset, conv_flag, s);
    }
};

#define MI_API_BEGIN try {
#define MI_API_END                                                                  \
    }                                                                               \
    catch(const mi::Error& e)                                                       \
    {                                                                               \
        mi::set_last_error(e.what());                                               \
        return e.code;                                                              \
    }                                                                               \
    catch(const std::exception& e)                                                  \
    {                                                                               \
        mi::set_last_error(std::string("internal error: ") + e.what());             \
        return MI_ERR_HIP;                                                          \
    }                                                                               \
    mi::set_last_error("");                                                         \
    return MI_OK;

static hipStream_t pick(mi_ctx* ctx, void* stream)
{
    return stream ? (hipStream_t)stream : ctx->stream;
}

// the layout the system's SpMV streams: a window shape index, or the full BSR
static int layout_shape(const mi_sys* sys)
{
    if(!spmv_use_window())
        return spmv_shape_full;
    if(sys->spmv_shape == spmv_shape_auto)
        return spmv_auto_shape(sys->desc.block_rows, sys->desc.value_precision == MI_FP32);
    return sys->spmv_shape;
}

// the SpMV's layout of the assembled matrix: the full BSR (built from the
// last conversion if it was not), or the window layout for the current row
// order: the identity, permuted inside every MAS preconditioner's range by its
// part order
static void build_spmv_layout(mi_sys* sys, hipStream_t s)
{
    auto&     A     = sys->A;
    const int shape = layout_shape(sys);
    if(shape < 0)
    {
        if(!A.has_full)
            sys->conv.build_full(A, sys->desc.value_precision == MI_FP32, sys->desc.reserve_ratio, s);
        A.has_win = false;
        return;
    }
    const int32_t B = A.B;
    if(A.win_order != sys->order_gen || A.wpos.size() < (size_t)std::max(B, 1))
    {
        std::vector<int32_t> pos((size_t)std::max(B, 1)), real((size_t)std::max(B, 1));
        for(int32_t v = 0; v < B; ++v)
            pos[v] = v;
        for(auto& pc : sys->pcs)
            if(const std::vector<int32_t>* rk = pc->spmv_rank())
            {
                const int64_t off = pc->dof_offset / 3;
                MI_REQUIRE(off + (int64_t)rk->size() <= B, MI_ERR_STATE, "spmv layout: preconditioner range");
                for(size_t l = 0; l < rk->size(); ++l)
                    pos[off + l] = (int32_t)(off + (*rk)[l]);
            }
        for(int32_t v = 0; v < B; ++v)
            real[pos[v]] = v;
        A.wpos.resize(pos.size());
        A.wreal.resize(real.size());
        MI_HIP(hipMemcpyAsync(A.wpos.data(), pos.data(), sizeof(int32_t) * pos.size(), hipMemcpyHostToDevice, s));
        MI_HIP(hipMemcpyAsync(A.wreal.data(), real.data(), sizeof(int32_t) * real.size(), hipMemcpyHostToDevice, s));
        MI_HIP(hipStreamSynchronize(s));  // the host vectors go out of scope
        A.win_order = sys->order_gen;
    }
    sys->conv.build_window(A, shape, sys->desc.value_precision == MI_FP32, sys->desc.reserve_ratio, s);
}

// the matrix buffers a captured PCG graph launches with
template <typename K>
static void key_matrix(K& key, const BsrMatrix& A)
{
    key.bufs[4]  = A.full_ptr.data();
    key.bufs[5]  = A.fcols.data();
    key.bufs[6]  = A.has32 ? (const void*)A.fvals32.data() : (const void*)A.fvals.data();
    key.bufs[12] = A.has_win ? (const void*)A.wptr.data() : nullptr;
    key.bufs[13] = A.has_win ? (const void*)A.wwords.data() : nullptr;
    key.bufs[14] = A.has_win ? (A.wvals32.data() ? (const void*)A.wvals32.data() : (const void*)A.wvals.data()) : nullptr;
    key.bufs[15] = A.has_win ? (const void*)A.wreal.data() : nullptr;
    key.wstride  = A.has_win ? A.wstride : -1;
    // the captured launch is one kernel instantiation with its window size
    key.wshape   = A.has_win ? A.wshape : -1;
    key.items    = spmv_items(A);
}

extern "C" {

const char* mi_last_error(void)
{
    return mi::g_last_error.c_str();
}

int mi_version(void)
{
    return MI_LINSYS_VERSION;
}

int mi_ctx_create(int device, mi_ctx** out)
{
    MI_API_BEGIN
    MI_REQUIRE(out != nullptr, MI_ERR_ARG, "mi_ctx_create: out is NULL");
    int count = 0;
    MI_HIP(hipGetDeviceCount(&count));
    MI_REQUIRE(device >= 0 && device < count, MI_ERR_ARG, "mi_ctx_create: no such device");
    MI_HIP(hipSetDevice(device));
    spmv_noalloc_init(device);
    auto* c   = new mi_ctx;
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if(e != hipSuccess)
    {
        delete c;
        MI_HIP(e);
    }
    *out = c;
    MI_API_END
}

int mi_ctx_destroy(mi_ctx* ctx)
{
    MI_API_BEGIN
    if(!ctx)
        return MI_OK;
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    MI_API_END
}

int mi_ctx_synchronize(mi_ctx* ctx)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx, MI_ERR_ARG, "null ctx");
    MI_HIP(hipStreamSynchronize(ctx->stream));
    MI_API_END
}

int mi_ctx_stream(mi_ctx* ctx, void** stream_out)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx && stream_out, MI_ERR_ARG, "null argument");
    *stream_out = (void*)ctx->stream;
    MI_API_END
}

int mi_malloc(mi_ctx* ctx, void** dptr, size_t bytes)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx && dptr, MI_ERR_ARG, "null argument");
    hipError_t e = hipMalloc(dptr, bytes ? bytes : 1);
    MI_REQUIRE(e == hipSuccess, MI_ERR_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    MI_API_END
}

int mi_free(mi_ctx* ctx, void* dptr)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx, MI_ERR_ARG, "null ctx");
    if(dptr)
    {
        MI_HIP(hipStreamSynchronize(ctx->stream));
        MI_HIP(hipFree(dptr));
    }
    MI_API_END
}

int mi_memcpy_h2d(mi_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx, MI_ERR_ARG, "null ctx");
    if(bytes)
    {
        MI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        MI_HIP(hipStreamSynchronize(ctx->stream));
    }
    MI_API_END
}

int mi_memcpy_d2h(mi_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx, MI_ERR_ARG, "null ctx");
    if(bytes)
    {
        MI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        MI_HIP(hipStreamSynchronize(ctx->stream));
    }
    MI_API_END
}

int mi_memcpy_d2d(mi_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx, MI_ERR_ARG, "null ctx");
    if(bytes)
    {
        MI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        MI_HIP(hipStreamSynchronize(ctx->stream));
    }
    MI_API_END
}

int mi_memset(mi_ctx* ctx, void* dst, int value, size_t bytes)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx, MI_ERR_ARG, "null ctx");
    if(bytes)
    {
        MI_HIP(hipMemsetAsync(dst, value, bytes, ctx->stream));
        MI_HIP(hipStreamSynchronize(ctx->stream));
    }
    MI_API_END
}

int mi_dot_f64(mi_ctx* ctx, const double* x, const double* y, int64_t n, double* out)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx && out && n >= 0, MI_ERR_ARG, "mi_dot_f64: invalid argument");
    *out = dot_f64_host(x, y, n, ctx->stream);
    MI_API_END
}

int mi_norm_f64(mi_ctx* ctx, const double* x, int64_t n, double* out)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx && out && n >= 0, MI_ERR_ARG, "mi_norm_f64: invalid argument");
    *out = norm_f64_host(x, n, ctx->stream);
    MI_API_END
}

int mi_dot_f32(mi_ctx* ctx, const float* x, const float* y, int64_t n, float* out)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx && out && n >= 0, MI_ERR_ARG, "mi_dot_f32: invalid argument");
    *out = dot_f32_host(x, y, n, ctx->stream);
    MI_API_END
}

int mi_norm_f32(mi_ctx* ctx, const float* x, int64_t n, float* out)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx && out && n >= 0, MI_ERR_ARG, "mi_norm_f32: invalid argument");
    *out = norm_f32_host(x, n, ctx->stream);
    MI_API_END
}

int mi_sys_create(mi_ctx* ctx, const mi_sys_desc* desc, mi_sys** out)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx && desc && out, MI_ERR_ARG, "mi_sys_create: null argument");
    MI_REQUIRE(desc->block_rows >= 0, MI_ERR_ARG, "mi_sys_create: block_rows < 0");
    MI_REQUIRE(desc->value_precision == MI_FP64 || desc->value_precision == MI_FP32, MI_ERR_ARG,
               "mi_sys_create: unknown value_precision");
    MI_REQUIRE((int64_t)desc->block_rows * desc->block_rows < ((int64_t)1 << 62), MI_ERR_ARG,
               "mi_sys_create: block_rows too large");
    auto* s  = new mi_sys;
    s->ctx   = ctx;
    s->desc  = *desc;
    if(!(s->desc.reserve_ratio >= 1.0))
        s->desc.reserve_ratio = 1.1;
    *out = s;
    MI_API_END
}

int mi_sys_destroy(mi_sys* sys)
{
    MI_API_BEGIN
    if(!sys)
        return MI_OK;
    (void)hipStreamSynchronize(sys->ctx->stream);
    delete sys;
    MI_API_END
}

}  // extern "C"

namespace
{
template <typename TV>
void assemble_t(mi_sys* sys, int64_t T, const int32_t* rows, const int32_t* cols, const TV* vals, void* stream)
{
    MI_REQUIRE(sys, MI_ERR_ARG, "mi_sys_assemble: null sys");
    MI_REQUIRE(T >= 0, MI_ERR_ARG, "mi_sys_assemble: negative triplet count");
    hipStream_t s = pick(sys->ctx, stream);
    sys->conv.convert_sym(sys->desc.block_rows,
                          T,
                          rows,
                          cols,
                          vals,
                          sys->A,
                          sys->desc.value_precision == MI_FP32,
                          layout_shape(sys) < 0,
                          sys->desc.reserve_ratio,
                          s);
    build_spmv_layout(sys, s);
    sys->assembled     = true;
    sys->pcs_assembled = false;
}
}  // namespace

extern "C" {

int mi_sys_assemble(mi_sys* sys, int64_t T, const int32_t* rows, const int32_t* cols, const double* vals, void* stream)
{
    MI_API_BEGIN
    assemble_t(sys, T, rows, cols, vals, stream);
    MI_API_END
}

int mi_sys_assemble_f32(mi_sys* sys, int64_t T, const int32_t* rows, const int32_t* cols, const float* vals,
                        void* stream)
{
    MI_API_BEGIN
    MI_REQUIRE(sys, MI_ERR_ARG, "mi_sys_assemble_f32: null sys");
    MI_REQUIRE(sys->desc.value_precision == MI_FP32, MI_ERR_ARG,
               "mi_sys_assemble_f32: FP32 triplets need an FP32 system (value_precision = MI_FP32)");
    assemble_t(sys, T, rows, cols, vals, stream);
    MI_API_END
}

int mi_sys_update_extents(mi_sys*          sys,
                          int32_t          n,
                          const mi_extent* ext,
                          mi_extent_view*  view,
                          int64_t*         total_triplets,
                          int64_t*         triplet_capacity)
{
    MI_API_BEGIN
    MI_REQUIRE(sys && n >= 0 && (n == 0 || (ext && view)), MI_ERR_ARG, "mi_sys_update_extents: bad argument");
    // DoF offsets: the diagonal subsystems in order (diag_dof_offsets_counts.scan)
    std::vector<int64_t> dof_off((size_t)n, -1);
    int64_t              dof = 0;
    for(int32_t k = 0; k < n; ++k)
        if(ext[k].is_diag)
        {
            MI_REQUIRE(ext[k].dof_count >= 0 && ext[k].dof_count % 3 == 0 && ext[k].rl_block_count == 0,
                       MI_ERR_ARG, "mi_sys_update_extents: a diagonal subsystem needs 3k DoFs and no rl blocks");
            dof_off[k] = dof;
            dof += ext[k].dof_count;
        }
    // triplet offsets: every subsystem in order (subsystem_triplet_offsets_counts.scan)
    int64_t T = 0;
    for(int32_t k = 0; k < n; ++k)
    {
        const mi_extent& e = ext[k];
        MI_REQUIRE(e.block_count >= 0 && e.rl_block_count >= 0, MI_ERR_ARG,
                   "mi_sys_update_extents: negative triplet count");
        mi_extent_view& v = view[k];
        v.triplet_offset  = T;
        if(e.is_diag)
        {
            v.dof_offset        = dof_off[k];
            v.rl_triplet_offset = -1;
            v.row_offset = v.col_offset = (int32_t)(dof_off[k] / 3);
            T += e.block_count;
        }
        else
        {
            MI_REQUIRE(e.dof_count == 0 && e.l >= 0 && e.l < n && e.r >= 0 && e.r < n && ext[e.l].is_diag
                           && ext[e.r].is_diag,
                       MI_ERR_ARG, "mi_sys_update_extents: an off-diagonal subsystem couples two diagonal ones");
            v.dof_offset        = -1;
            v.rl_triplet_offset = T + e.block_count;
            v.row_offset        = (int32_t)(dof_off[e.l] / 3);
            v.col_offset        = (int32_t)(dof_off[e.r] / 3);
            T += e.block_count + e.rl_block_count;
        }
    }
    MI_REQUIRE(T < ((int64_t)1 << 31), MI_ERR_ARG, "mi_sys_update_extents: more than 2^31-1 triplets");
    MI_REQUIRE(dof / 3 < ((int64_t)1 << 31), MI_ERR_ARG, "mi_sys_update_extents: too many DoFs");
    const int32_t B = (int32_t)(dof / 3);
    if(B != sys->desc.block_rows)
    {
        MI_HIP(hipStreamSynchronize(sys->ctx->stream));
        sys->desc.block_rows = B;
        sys->assembled       = false;
        sys->pcs_assembled   = false;
        ++sys->order_gen;
    }
    // capacity x reserve_ratio when exceeded, contents discarded
    const double ratio = sys->desc.reserve_ratio;
    const size_t nt    = (size_t)std::max<int64_t>(T, 1);
    if(nt > sys->own_rows.cap)
        MI_HIP(hipStreamSynchronize(sys->ctx->stream));  // no launch may still use the old buffers
    sys->own_rows.resize(nt, ratio);
    sys->own_cols.resize(nt, ratio);
    sys->own_vals.resize(9 * nt, ratio);
    sys->own_b.resize((size_t)std::max<int64_t>(dof, 1), ratio);
    sys->own_T = T;
    if(total_triplets)
        *total_triplets = T;
    if(triplet_capacity)
        *triplet_capacity = (int64_t)sys->own_rows.cap;
    MI_API_END
}

int mi_sys_begin_assembly(mi_sys* sys, int32_t** rows, int32_t** cols, double** vals, double** b, void* stream)
{
    MI_API_BEGIN
    MI_REQUIRE(sys, MI_ERR_ARG, "mi_sys_begin_assembly: null sys");
    MI_REQUIRE(sys->own_T >= 0, MI_ERR_STATE, "mi_sys_begin_assembly: no extents reported");
    hipStream_t s = pick(sys->ctx, stream);
    fill_i32(sys->own_rows.data(), -1, sys->own_T, s);
    fill_i32(sys->own_cols.data(), -1, sys->own_T, s);
    fill_f64(sys->own_b.data(), 0.0, sys->n(), s);
    MI_HIP(hipGetLastError());
    if(rows)
        *rows = sys->own_rows.data();
    if(cols)
        *cols = sys->own_cols.data();
    if(vals)
        *vals = sys->own_vals.data();
    if(b)
        *b = sys->own_b.data();
    MI_API_END
}

int mi_sys_assemble_owned(mi_sys* sys, void* stream)
{
    if(!sys || sys->own_T < 0)
    {
        set_last_error(sys ? "mi_sys_assemble_owned: no extents reported" : "mi_sys_assemble_owned: null sys");
        return sys ? MI_ERR_STATE : MI_ERR_ARG;
    }
    return mi_sys_assemble(sys, sys->own_T, sys->own_rows.data(), sys->own_cols.data(), sys->own_vals.data(), stream);
}

int mi_sys_nnz(mi_sys* sys, int64_t* nnz)
{
    MI_API_BEGIN
    MI_REQUIRE(sys && nnz, MI_ERR_ARG, "mi_sys_nnz: null argument");
    MI_REQUIRE(sys->assembled, MI_ERR_STATE, "mi_sys_nnz: not assembled");
    *nnz = sys->A.U;
    MI_API_END
}

int mi_sys_export_bcoo(mi_sys* sys, int32_t* rows, int32_t* cols, double* vals, int32_t* row_ptr, void* stream)
{
    MI_API_BEGIN
    MI_REQUIRE(sys, MI_ERR_ARG, "null sys");
    MI_REQUIRE(sys->assembled, MI_ERR_STATE, "mi_sys_export_bcoo: not assembled");
    hipStream_t s = pick(sys->ctx, stream);
    auto&       A = sys->A;
    if(rows && A.U)
        MI_HIP(hipMemcpyAsync(rows, A.rows.data(), sizeof(int32_t) * A.U, hipMemcpyDeviceToDevice, s));
    if(cols && A.U)
        MI_HIP(hipMemcpyAsync(cols, A.cols.data(), sizeof(int32_t) * A.U, hipMemcpyDeviceToDevice, s));
    if(row_ptr)
        MI_HIP(hipMemcpyAsync(row_ptr, A.row_ptr.data(), sizeof(int32_t) * (A.B + 1), hipMemcpyDeviceToDevice, s));
    if(vals && A.U)  // the upper values are stored per block, as exported
        MI_HIP(hipMemcpyAsync(vals, A.vals.data(), sizeof(double) * 9 * A.U, hipMemcpyDeviceToDevice, s));
    MI_HIP(hipGetLastError());
    MI_API_END
}

// Matrix Market dumps in the reference's debug format (utils/matrix_market.h:34-110):
// banner, four comment lines, "rows cols entries", then every block's 9 scalars
// row by row within the block, zeros included, %.17g.
int mi_sys_write_mtx(mi_sys* sys, const char* path, int one_based)
{
    MI_API_BEGIN
    MI_REQUIRE(sys && path, MI_ERR_ARG, "mi_sys_write_mtx: null argument");
    MI_REQUIRE(sys->assembled, MI_ERR_STATE, "mi_sys_write_mtx: not assembled");
    hipStream_t          s = sys->ctx->stream;
    auto&                A = sys->A;
    std::vector<int32_t> rows(A.U), cols(A.U);
    std::vector<double>  vals(9 * A.U);
    if(A.U)
    {
        MI_HIP(hipMemcpyAsync(rows.data(), A.rows.data(), sizeof(int32_t) * A.U, hipMemcpyDeviceToHost, s));
        MI_HIP(hipMemcpyAsync(cols.data(), A.cols.data(), sizeof(int32_t) * A.U, hipMemcpyDeviceToHost, s));
        MI_HIP(hipMemcpyAsync(vals.data(), A.vals.data(), sizeof(double) * 9 * A.U, hipMemcpyDeviceToHost, s));
        MI_HIP(hipStreamSynchronize(s));
    }
    FILE* fp = std::fopen(path, "w");
    MI_REQUIRE(fp, MI_ERR_ARG, std::string("mi_sys_write_mtx: cannot open ") + path);
    const long long n = 3LL * A.B, entries = 9LL * A.U;
    const int       o = one_based ? 1 : 0;
    std::fprintf(fp, "%%%%MatrixMarket matrix coordinate real general\n");
    std::fprintf(fp, "%% Exported from block-sparse triplet matrix\n");
    std::fprintf(fp, "%% Block dimension: 3x3\n");
    std::fprintf(fp, "%% Total rows: %lld, Total cols: %lld\n", n, n);
    std::fprintf(fp, "%% Total entries (including zeros): %lld\n", entries);
    std::fprintf(fp, "%lld %lld %lld\n", n, n, entries);
    for(int64_t t = 0; t < A.U; ++t)
        for(int bi = 0; bi < 3; ++bi)
            for(int bj = 0; bj < 3; ++bj)
                std::fprintf(fp, "%lld %lld %.17g\n", 3LL * rows[t] + bi + o, 3LL * cols[t] + bj + o,
                             vals[9 * t + bi + 3 * bj]);
    const bool ok = std::ferror(fp) == 0;
    std::fclose(fp);
    MI_REQUIRE(ok, MI_ERR_ARG, std::string("mi_sys_write_mtx: write failed: ") + path);
    MI_API_END
}

int mi_vec_write_mtx(mi_ctx* ctx, const char* path, const double* x, int64_t n)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx && path && (x || n == 0) && n >= 0, MI_ERR_ARG, "mi_vec_write_mtx: bad argument");
    std::vector<double> h(n);
    if(n)
    {
        MI_HIP(hipMemcpyAsync(h.data(), x, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
        MI_HIP(hipStreamSynchronize(ctx->stream));
    }
    FILE* fp = std::fopen(path, "w");
    MI_REQUIRE(fp, MI_ERR_ARG, std::string("mi_vec_write_mtx: cannot open ") + path);
    std::fprintf(fp, "%%%%MatrixMarket matrix array real general\n");
    std::fprintf(fp, "%% Exported from dense vector\n");
    std::fprintf(fp, "%% Vector size: %lld\n", (long long)n);
    std::fprintf(fp, "%lld 1\n", (long long)n);
    for(int64_t i = 0; i < n; ++i)
        std::fprintf(fp, "%.17g\n", h[i]);
    const bool ok = std::ferror(fp) == 0;
    std::fclose(fp);
    MI_REQUIRE(ok, MI_ERR_ARG, std::string("mi_vec_write_mtx: write failed: ") + path);
    MI_API_END
}

int mi_spmv_dot(mi_sys* sys, const double* x, double* y, double* d_dot, void* stream)
{
    MI_API_BEGIN
    MI_REQUIRE(sys && x && y, MI_ERR_ARG, "mi_spmv_dot: null argument");
    MI_REQUIRE(sys->assembled, MI_ERR_STATE, "mi_spmv_dot: not assembled");
    hipStream_t s = pick(sys->ctx, stream);
    spmv_accumulate_dot(sys->A, sys->desc.value_precision == MI_FP32, 1.0, x, y, d_dot, sys->spmv_scr, s);
    MI_API_END
}

int mi_spmv_shapes(int32_t* count, int32_t* dflt, int32_t* threads, int32_t* rows, int32_t* per, int32_t* vec)
{
    MI_API_BEGIN
    const int n = spmv_shape_count();
    if(count)
        *count = n;
    if(dflt)
        *dflt = spmv_default_shape();
    for(int k = 0; k < n; ++k)
        spmv_shape_info(k, threads ? threads + k : nullptr, rows ? rows + k : nullptr, per ? per + k : nullptr,
                        vec ? vec + k : nullptr);
    MI_API_END
}

int mi_sys_set_spmv_shape(mi_sys* sys, int32_t shape, void* stream)
{
    MI_API_BEGIN
    MI_REQUIRE(sys, MI_ERR_ARG, "mi_sys_set_spmv_shape: null sys");
    MI_REQUIRE(shape == spmv_shape_auto || shape == spmv_shape_full || (shape >= 0 && shape < spmv_shape_count()),
               MI_ERR_ARG, "mi_sys_set_spmv_shape: unknown shape");
    sys->spmv_shape = shape;
    // an assembled system gets the new layout at once
    const int want = layout_shape(sys);
    if(sys->assembled && (want < 0 ? sys->A.has_win : (!sys->A.has_win || sys->A.wshape != want)))
        build_spmv_layout(sys, pick(sys->ctx, stream));
    MI_API_END
}

int mi_sys_spmv_layout(mi_sys* sys, int32_t* layout, int32_t* window_rows, int32_t* threads, int64_t* blocks,
                       int32_t* shape)
{
    MI_API_BEGIN
    MI_REQUIRE(sys, MI_ERR_ARG, "mi_sys_spmv_layout: null sys");
    MI_REQUIRE(sys->assembled, MI_ERR_STATE, "mi_sys_spmv_layout: not assembled");
    const auto& A = sys->A;
    if(layout)
        *layout = A.has_win ? 1 : 0;
    if(window_rows)
        *window_rows = A.has_win ? (1 << A.wshift) : 0;
    if(threads)
        *threads = spmv_threads(A);
    if(shape)
        *shape = A.has_win ? A.wshape : -1;
    if(blocks)
    {
        int32_t n = 0;
        const int32_t* ptr = A.has_win ? A.wptr.data() : A.full_ptr.data();
        if(A.B > 0 && ptr)
        {
            MI_HIP(hipStreamSynchronize(sys->ctx->stream));
            MI_HIP(hipMemcpy(&n, ptr + A.B, sizeof(int32_t), hipMemcpyDeviceToHost));
        }
        *blocks = n;
    }
    MI_API_END
}

int mi_spmv_dot_noalloc(mi_sys* sys, const double* x, double* y, double* d_dot, void* stream)
{
    MI_API_BEGIN
    MI_REQUIRE(sys && x && y, MI_ERR_ARG, "mi_spmv_dot_noalloc: null argument");
    MI_REQUIRE(sys->assembled, MI_ERR_STATE, "mi_spmv_dot_noalloc: not assembled");
    spmv_accumulate_dot_noalloc(sys->A, sys->desc.value_precision == MI_FP32, x, y, d_dot, sys->ctx->device,
                                pick(sys->ctx, stream));
    MI_API_END
}

int mi_spmv(mi_sys* sys, double a, const double* x, double b, double* y, void* stream)
{
    MI_API_BEGIN
    MI_REQUIRE(sys && x && y, MI_ERR_ARG, "mi_spmv: null argument");
    MI_REQUIRE(sys->assembled, MI_ERR_STATE, "mi_spmv: not assembled");
    spmv_axpby(sys->A, sys->desc.value_precision == MI_FP32, a, x, b, y, sys->spmv_scr, pick(sys->ctx, stream));
    MI_API_END
}

int mi_sys_spmv_order(mi_sys* sys, int32_t* pos)
{
    MI_API_BEGIN
    MI_REQUIRE(sys && pos, MI_ERR_ARG, "mi_sys_spmv_order: null argument");
    MI_REQUIRE(sys->assembled, MI_ERR_STATE, "mi_sys_spmv_order: not assembled");
    const auto& A = sys->A;
    if(A.has_win && A.B > 0)
    {
        MI_HIP(hipStreamSynchronize(sys->ctx->stream));
        MI_HIP(hipMemcpy(pos, A.wpos.data(), sizeof(int32_t) * A.B, hipMemcpyDeviceToHost));
    }
    else
        for(int32_t v = 0; v < A.B; ++v)
            pos[v] = v;
    MI_API_END
}

int mi_spmv_ordered(mi_sys* sys, const double* x, double* y, double* d_dot, void* stream)
{
    MI_API_BEGIN
    MI_REQUIRE(sys && x && y, MI_ERR_ARG, "mi_spmv_ordered: null argument");
    MI_REQUIRE(sys->assembled, MI_ERR_STATE, "mi_spmv_ordered: not assembled");
    MI_REQUIRE(sys->A.has_win, MI_ERR_STATE, "mi_spmv_ordered: the matrix has no window layout");
    hipStream_t s     = pick(sys->ctx, stream);
    const bool  use32 = sys->desc.value_precision == MI_FP32;
    spmv_prepare(sys->spmv_scr, use32);
    const int np = spmv_full(sys->A, use32, x, nullptr, y, sys->spmv_scr.partials.data(), nullptr, sys->spmv_scr, s,
                             nullptr, nullptr, true);
    if(d_dot)
        sum_partials(sys->spmv_scr.partials.data(), np, d_dot, s);
    MI_API_END
}
}  // extern "C"

namespace
{
// a preconditioner for `d`, checked against the system and the other ranges
// (the one at `skip` excepted)
std::unique_ptr<LocalPreconditioner> build_precond(mi_sys* sys, const mi_precond_desc* d, int skip, const char* who)
{
    MI_REQUIRE(sys && d, MI_ERR_ARG, std::string(who) + ": null argument");
    MI_REQUIRE(d->dof_offset >= 0 && d->dof_count >= 0 && d->dof_offset + d->dof_count <= sys->n(), MI_ERR_ARG,
               std::string(who) + ": DoF range outside the system");
    for(int k = 0; k < (int)sys->pcs.size(); ++k)
    {
        const auto& pc = sys->pcs[k];
        MI_REQUIRE(k == skip || d->dof_offset >= pc->dof_offset + pc->dof_count
                       || d->dof_offset + d->dof_count <= pc->dof_offset,
                   MI_ERR_ARG, std::string(who) + ": overlapping preconditioner ranges");
    }
    switch(d->kind)
    {
        case MI_PRECOND_DIAG3: return make_diag3(d->dof_offset, d->dof_count);
        case MI_PRECOND_DIAG12: return make_diag12(d->dof_offset, d->dof_count);
        case MI_PRECOND_MAS: return make_mas(*d);
        default: throw Error(MI_ERR_ARG, std::string(who) + ": unknown kind");
    }
}
}  // namespace

extern "C" {
int mi_precond_add(mi_sys* sys, const mi_precond_desc* d, int* handle)
{
    MI_API_BEGIN
    auto pc = build_precond(sys, d, -1, "mi_precond_add");
    sys->pcs.push_back(std::move(pc));
    sys->pcs_assembled = false;
    ++sys->pcs_gen;
    ++sys->order_gen;
    if(handle)
        *handle = (int)sys->pcs.size() - 1;
    MI_API_END
}

int mi_precond_reinit(mi_sys* sys, int handle, const mi_precond_desc* d)
{
    MI_API_BEGIN
    MI_REQUIRE(sys && handle >= 0 && handle < (int)sys->pcs.size(), MI_ERR_ARG, "mi_precond_reinit: bad handle");
    auto pc = build_precond(sys, d, handle, "mi_precond_reinit");
    MI_HIP(hipDeviceSynchronize());  // no launch may still read the preconditioner being replaced
    sys->pcs[handle]   = std::move(pc);
    sys->pcs_assembled = false;
    ++sys->pcs_gen;
    ++sys->order_gen;
    MI_API_END
}

int mi_precond_mas_set_active_levels(mi_sys* sys, int handle, int32_t levels)
{
    MI_API_BEGIN
    MI_REQUIRE(sys && handle >= 0 && handle < (int)sys->pcs.size(), MI_ERR_ARG,
               "mi_precond_mas_set_active_levels: bad handle");
    MI_REQUIRE(sys->pcs[handle]->set_active_levels(levels), MI_ERR_ARG,
               "mi_precond_mas_set_active_levels: not a MAS preconditioner");
    ++sys->pcs_gen;  // a captured solve launches the old level count's kernels
    MI_API_END
}

int mi_precond_mas_set_system_levels(mi_sys* sys, int handle, int32_t nseg, const int32_t* seg_rows,
                                     const int32_t* levels)
{
    MI_API_BEGIN
    MI_REQUIRE(sys && handle >= 0 && handle < (int)sys->pcs.size(), MI_ERR_ARG,
               "mi_precond_mas_set_system_levels: bad handle");
    MI_REQUIRE(nseg >= 0 && (nseg == 0 || (seg_rows && levels)), MI_ERR_ARG,
               "mi_precond_mas_set_system_levels: null arrays");
    for(int32_t g = 0; g < nseg; ++g)
        MI_REQUIRE(seg_rows[g + 1] > seg_rows[g], MI_ERR_ARG,
                   "mi_precond_mas_set_system_levels: empty or decreasing segment");
    MI_REQUIRE(sys->pcs[handle]->set_system_levels(nseg, seg_rows, levels), MI_ERR_ARG,
               "mi_precond_mas_set_system_levels: not a MAS preconditioner");
    ++sys->pcs_gen;  // a captured solve launches the old counts' kernels
    MI_API_END
}

int mi_precond_assemble(mi_sys* sys, void* stream)
{
    MI_API_BEGIN
    MI_REQUIRE(sys, MI_ERR_ARG, "null sys");
    MI_REQUIRE(sys->assembled, MI_ERR_STATE, "mi_precond_assemble: assemble the system first");
    hipStream_t s = pick(sys->ctx, stream);
    // a preconditioner added since the assembly may change the SpMV's row order
    if(sys->A.has_win && sys->A.win_order != sys->order_gen)
        build_spmv_layout(sys, s);
    for(auto& pc : sys->pcs)
        pc->assemble(sys->A, s);
    sys->pcs_assembled = true;
    MI_API_END
}

int mi_precond_apply(mi_sys* sys, const double* r, double* z, const int32_t* d_conv, void* stream)
{
    MI_API_BEGIN
    MI_REQUIRE(sys && r && z, MI_ERR_ARG, "mi_precond_apply: null argument");
    MI_REQUIRE(sys->pcs.empty() || sys->pcs_assembled, MI_ERR_STATE, "mi_precond_apply: preconditioners not assembled");
    sys->apply_preconditioner(r, z, d_conv, pick(sys->ctx, stream));
    MI_HIP(hipGetLastError());
    MI_API_END
}

int mi_precond_mas_export(mi_sys*  sys,
                          int      handle,
                          int32_t* info,
                          int32_t* p2r,
                          int32_t* r2p,
                          int32_t* gn,
                          int32_t* ls,
                          int32_t* ct,
                          double*  hess,
                          double*  inv)
{
    MI_API_BEGIN
    MI_REQUIRE(sys && handle >= 0 && handle < (int)sys->pcs.size(), MI_ERR_ARG, "mi_precond_mas_export: bad handle");
    MI_REQUIRE(sys->pcs[handle]->mas_export(info, p2r, r2p, gn, ls, ct, hess, inv), MI_ERR_ARG,
               "mi_precond_mas_export: not a MAS preconditioner");
    MI_API_END
}

int mi_subsystem_fix_vertices(mi_ctx*        ctx,
                              int64_t        n_triplets,
                              const int32_t* rows,
                              const int32_t* cols,
                              double*        vals,
                              int32_t        block_offset,
                              const int32_t* is_fixed,
                              int32_t        vert_num,
                              double*        gradient,
                              void*          stream)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx && n_triplets >= 0 && vert_num >= 0 && block_offset >= 0, MI_ERR_ARG,
               "mi_subsystem_fix_vertices: bad argument");
    MI_REQUIRE(n_triplets == 0 || (rows && cols && vals), MI_ERR_ARG, "mi_subsystem_fix_vertices: null triplets");
    MI_REQUIRE(vert_num == 0 || is_fixed, MI_ERR_ARG, "mi_subsystem_fix_vertices: null is_fixed");
    subsystem_fix_vertices(rows, cols, vals, n_triplets, block_offset, is_fixed, vert_num, gradient,
                           pick(ctx, stream));
    MI_API_END
}

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
                      void*          stream)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx, MI_ERR_ARG, "mi_mesh_partition: null context");
    MI_REQUIRE(n_tets >= 0 && n_tris >= 0 && n_edges >= 0, MI_ERR_ARG, "mi_mesh_partition: negative count");
    const int32_t p = mesh_partition_device(vert_num, tets, n_tets, tris, n_tris, edges, n_edges, part_max_size,
                                            part_ids, pick(ctx, stream));
    if(n_parts_out)
        *n_parts_out = p;
    MI_API_END
}

int mi_subsystem_distribute(mi_ctx* ctx, const double* x, int64_t dof_offset, int32_t vert_num, double* dxs,
                            void* stream)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx && dof_offset >= 0 && vert_num >= 0, MI_ERR_ARG, "mi_subsystem_distribute: bad argument");
    MI_REQUIRE(vert_num == 0 || (x && dxs), MI_ERR_ARG, "mi_subsystem_distribute: null vector");
    subsystem_distribute(x + dof_offset, dxs, 3 * (int64_t)vert_num, pick(ctx, stream));
    MI_API_END
}

int mi_produce_kinetic(mi_ctx*       ctx,
                       int32_t       vert_num,
                       const double* mass,
                       int32_t       vert_offset,
                       int32_t*      rows,
                       int32_t*      cols,
                       double*       vals,
                       void*         stream)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx && vert_num >= 0 && vert_offset >= 0, MI_ERR_ARG, "mi_produce_kinetic: bad argument");
    MI_REQUIRE(vert_num == 0 || (mass && rows && cols && vals), MI_ERR_ARG, "mi_produce_kinetic: null buffer");
    produce_kinetic(vert_num, mass, vert_offset, rows, cols, vals, pick(ctx, stream));
    MI_API_END
}

int mi_produce_snh(mi_ctx*        ctx,
                   int64_t        n_tets,
                   const int32_t* tets,
                   const double*  rest_positions,
                   const double*  positions,
                   double         mu,
                   double         lambda,
                   double         dt,
                   int32_t        vert_offset,
                   int32_t*       rows,
                   int32_t*       cols,
                   double*        vals,
                   void*          stream)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx && n_tets >= 0 && vert_offset >= 0, MI_ERR_ARG, "mi_produce_snh: bad argument");
    MI_REQUIRE(mu > 0.0 && lambda > 0.0 && dt > 0.0, MI_ERR_ARG, "mi_produce_snh: mu, lambda and dt must be > 0");
    MI_REQUIRE(n_tets == 0 || (tets && rest_positions && positions && rows && cols && vals), MI_ERR_ARG,
               "mi_produce_snh: null buffer");
    produce_snh(n_tets, tets, rest_positions, positions, mu, lambda, dt, vert_offset, rows, cols, vals,
                pick(ctx, stream));
    MI_API_END
}

int mi_produce_abd(mi_ctx*        ctx,
                   int64_t        n_bodies,
                   const double*  q,
                   const double*  dyadic_mass,
                   const double*  volume,
                   const double*  kappa,
                   const int32_t* is_fixed,
                   double         dt,
                   int32_t        body_offset,
                   int32_t*       rows,
                   int32_t*       cols,
                   double*        vals,
                   void*          stream)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx && n_bodies >= 0 && body_offset >= 0, MI_ERR_ARG, "mi_produce_abd: bad argument");
    MI_REQUIRE(dt > 0.0, MI_ERR_ARG, "mi_produce_abd: dt must be > 0");
    MI_REQUIRE(n_bodies == 0 || (q && dyadic_mass && volume && kappa && rows && cols && vals), MI_ERR_ARG,
               "mi_produce_abd: null buffer");
    produce_abd(n_bodies, q, dyadic_mass, volume, kappa, is_fixed, dt, body_offset, rows, cols, vals,
                pick(ctx, stream));
    MI_API_END
}

int mi_produce_bw_shell(mi_ctx*        ctx,
                        int64_t        n_tris,
                        const int32_t* tris,
                        const double*  rest_positions,
                        const double*  positions,
                        double         stretch_stiffness,
                        double         shear_stiffness,
                        double         strain_rate,
                        double         dt,
                        int32_t        vert_offset,
                        int32_t*       rows,
                        int32_t*       cols,
                        double*        vals,
                        void*          stream)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx && n_tris >= 0 && vert_offset >= 0, MI_ERR_ARG, "mi_produce_bw_shell: bad argument");
    MI_REQUIRE(dt > 0.0, MI_ERR_ARG, "mi_produce_bw_shell: dt must be > 0");
    MI_REQUIRE(n_tris == 0 || (tris && rest_positions && positions && rows && cols && vals), MI_ERR_ARG,
               "mi_produce_bw_shell: null buffer");
    produce_bw_shell(n_tris, tris, rest_positions, positions, stretch_stiffness, shear_stiffness, strain_rate, dt,
                     vert_offset, rows, cols, vals, pick(ctx, stream));
    MI_API_END
}

// ---- learned initial guess (include/mi_warmstart.h)
int mi_ws_gemm(mi_ctx* ctx, int32_t M, int32_t N, int32_t K, const float* A, int64_t sa_m, int64_t sa_k,
               const float* B, int64_t sb_k, int64_t sb_n, float* C, int64_t ldc, int32_t accumulate,
               const float* bias, int32_t relu, void* stream)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx && M >= 0 && N >= 0 && K >= 0 && (M == 0 || N == 0 || (A && B && C)), MI_ERR_ARG,
               "mi_ws_gemm: bad argument");
    ws_gemm(M, N, K, A, sa_m, sa_k, B, sb_k, sb_n, C, ldc, accumulate != 0, bias, relu != 0, pick(ctx, stream));
    MI_API_END
}

int mi_ws_aggregate(mi_ctx* ctx, int32_t n, int32_t F, const int32_t* ptr, const int32_t* idx, const float* w,
                    const float* H, float* Z, void* stream)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx && n >= 0 && F > 0 && F <= 64, MI_ERR_ARG, "mi_ws_aggregate: 1 <= F <= 64");
    MI_REQUIRE(n == 0 || (ptr && idx && w && H && Z), MI_ERR_ARG, "mi_ws_aggregate: null buffer");
    ws_aggregate(n, F, ptr, idx, w, H, Z, pick(ctx, stream));
    MI_API_END
}

int mi_ws_relu_backward(mi_ctx* ctx, int64_t n, const float* pre, float* d, void* stream)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx && n >= 0 && (n == 0 || (pre && d)), MI_ERR_ARG, "mi_ws_relu_backward: bad argument");
    ws_relu_backward(n, pre, d, pick(ctx, stream));
    MI_API_END
}

int mi_ws_colsum(mi_ctx* ctx, int32_t n, int32_t F, const float* X, float* out, int32_t accumulate, void* stream)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx && n >= 0 && F > 0 && X && out, MI_ERR_ARG, "mi_ws_colsum: bad argument");
    ws_colsum(n, F, X, out, accumulate != 0, pick(ctx, stream));
    MI_API_END
}

int mi_ws_adam(mi_ctx* ctx, int64_t n, float* w, const float* g, float* m, float* v, float lr, float beta1,
               float beta2, float eps, int32_t step, void* stream)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx && n >= 0 && step >= 1 && (n == 0 || (w && g && m && v)), MI_ERR_ARG, "mi_ws_adam: bad argument");
    ws_adam(n, w, g, m, v, lr, beta1, beta2, eps, step, pick(ctx, stream));
    MI_API_END
}

int mi_ws_block3(mi_ctx* ctx, int32_t n, const double* blocks, double scale, const float* v_f32, double* y_f64,
                 const double* base, const double* v_f64, float* y_f32, void* stream)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx && n >= 0 && blocks && ((v_f32 && y_f64 && !v_f64 && !y_f32) || (v_f64 && y_f32 && !v_f32 && !y_f64)),
               MI_ERR_ARG, "mi_ws_block3: give (v_f32, y_f64) or (v_f64, y_f32)");
    MI_REQUIRE(!base || v_f32, MI_ERR_ARG, "mi_ws_block3: base goes with (v_f32, y_f64)");
    ws_block3(n, blocks, scale, v_f32, y_f64, base, v_f64, y_f32, pick(ctx, stream));
    MI_API_END
}

int mi_produce_bending(mi_ctx*        ctx,
                       int64_t        n_hinges,
                       const int32_t* hinges,
                       const double*  rest_positions,
                       const double*  positions,
                       double         bending_stiffness,
                       double         dt,
                       int32_t        vert_offset,
                       int32_t*       rows,
                       int32_t*       cols,
                       double*        vals,
                       void*          stream)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx && n_hinges >= 0 && vert_offset >= 0, MI_ERR_ARG, "mi_produce_bending: bad argument");
    MI_REQUIRE(dt > 0.0 && bending_stiffness >= 0.0, MI_ERR_ARG, "mi_produce_bending: dt must be > 0, kappa >= 0");
    MI_REQUIRE(n_hinges == 0 || (hinges && rest_positions && positions && rows && cols && vals), MI_ERR_ARG,
               "mi_produce_bending: null buffer");
    produce_bending(n_hinges, hinges, rest_positions, positions, bending_stiffness, dt, vert_offset, rows, cols, vals,
                    pick(ctx, stream));
    MI_API_END
}

int mi_energy_kinetic(mi_ctx*       ctx,
                      int32_t       vert_num,
                      const double* x,
                      const double* x_tilde,
                      const double* mass,
                      double*       energy,
                      void*         stream)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx && vert_num >= 0 && energy, MI_ERR_ARG, "mi_energy_kinetic: bad argument");
    MI_REQUIRE(vert_num == 0 || (x && x_tilde && mass), MI_ERR_ARG, "mi_energy_kinetic: null buffer");
    energy_kinetic(vert_num, x, x_tilde, mass, energy, ctx->energy_part, pick(ctx, stream));
    MI_API_END
}

int mi_energy_snh(mi_ctx*        ctx,
                  int64_t        n_tets,
                  const int32_t* tets,
                  const double*  rest_positions,
                  const double*  positions,
                  double         mu,
                  double         lambda,
                  double         dt,
                  double*        energy,
                  void*          stream)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx && n_tets >= 0 && energy, MI_ERR_ARG, "mi_energy_snh: bad argument");
    MI_REQUIRE(mu > 0.0 && lambda > 0.0 && dt > 0.0, MI_ERR_ARG, "mi_energy_snh: mu, lambda and dt must be > 0");
    MI_REQUIRE(n_tets == 0 || (tets && rest_positions && positions), MI_ERR_ARG, "mi_energy_snh: null buffer");
    energy_snh(n_tets, tets, rest_positions, positions, mu, lambda, dt, energy, ctx->energy_part, pick(ctx, stream));
    MI_API_END
}

int mi_step_forward(mi_ctx* ctx, const double* x0, const double* dx, double alpha, double* x, int64_t n, void* stream)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx && n >= 0, MI_ERR_ARG, "mi_step_forward: bad argument");
    MI_REQUIRE(n == 0 || (x0 && dx && x), MI_ERR_ARG, "mi_step_forward: null buffer");
    step_forward(x0, dx, alpha, x, n, pick(ctx, stream));
    MI_API_END
}

int mi_produce_contact(mi_ctx*        ctx,
                       int64_t        n_stencils,
                       const int32_t* stencils,
                       const double*  g,
                       const double*  kappa,
                       double         dt,
                       int32_t        vert_offset,
                       int32_t*       rows,
                       int32_t*       cols,
                       double*        vals,
                       void*          stream)
{
    MI_API_BEGIN
    MI_REQUIRE(ctx && n_stencils >= 0 && vert_offset >= 0 && dt > 0.0, MI_ERR_ARG,
               "mi_produce_contact: bad argument");
    MI_REQUIRE(n_stencils == 0 || (stencils && g && kappa && rows && cols && vals), MI_ERR_ARG,
               "mi_produce_contact: null buffer");
    produce_contact(n_stencils, stencils, g, kappa, dt, vert_offset, rows, cols, vals, pick(ctx, stream));
    MI_API_END
}
}  // extern "C"

namespace
{
void write_int_array(FILE* fp, const char* key, const int32_t* a, int64_t n, bool last)
{
    std::fprintf(fp, "    \"%s\": [", key);
    for(int64_t i = 0; i < n; ++i)
        std::fprintf(fp, i ? ", %d" : "%d", a[i]);
    std::fprintf(fp, "]%s\n", last ? "" : ",");
}
}  // namespace

extern "C" {

// The files of MASPreconditionerEngine::dump_cluster_matrices_debug
// (finite_element/mas_preconditioner_engine.cu:3154-3301):
//   mas_cluster_hess.f<frame>.n<iter>.mtx, mas_cluster_inv.f<frame>.n<iter>.mtx
//     coordinate real general, (clusters * 48)^2 block diagonal, per cluster the
//     136 upper 3x3 blocks row by row within the block, %.17g, 1-based;
//   mas_cluster_meta.f<frame>.n<iter>.json  sizes, levels, part_to_real,
//     going_next, real_to_part.
int mi_precond_mas_dump(mi_sys* sys, int handle, const char* output_dir, int64_t frame, int64_t newton_iter)
{
    MI_API_BEGIN
    MI_REQUIRE(sys && output_dir && handle >= 0 && handle < (int)sys->pcs.size(), MI_ERR_ARG,
               "mi_precond_mas_dump: bad argument");
    auto&   pc = *sys->pcs[handle];
    int32_t info[6];
    MI_REQUIRE(pc.mas_export(info, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr), MI_ERR_ARG,
               "mi_precond_mas_dump: not a MAS preconditioner");
    const int32_t V = info[0], N = info[1], L = info[2], total = info[3];
    const int64_t C = info[5];
    constexpr int kSymBlocks = 136, kDim = 48;
    std::vector<int32_t> p2r(std::max(N, 1)), r2p(std::max(V, 1)), ls(2 * (L + 1));
    std::vector<int32_t> gn((size_t)std::max(1, ((std::max(N, V) + 15) / 16 * 16) * L));
    std::vector<double>  hess((size_t)std::max<int64_t>(1, C * kSymBlocks * 9)), inv(hess.size());
    pc.mas_export(info, p2r.data(), r2p.data(), gn.data(), ls.data(), nullptr, hess.data(), inv.data());
    const std::string dir = output_dir;
    const std::string tag = ".f" + std::to_string(frame) + ".n" + std::to_string(newton_iter);
    for(int which = 0; which < 2; ++which)
    {
        const char*       kind = which == 0 ? "hess" : "inv";
        const double*     M    = which == 0 ? hess.data() : inv.data();
        const std::string path = dir + "/mas_cluster_" + kind + tag + ".mtx";
        FILE*             fp   = std::fopen(path.c_str(), "w");
        MI_REQUIRE(fp, MI_ERR_ARG, "mi_precond_mas_dump: cannot open " + path);
        std::fprintf(fp, "%%%%MatrixMarket matrix coordinate real general\n");
        std::fprintf(fp, "%% MAS cluster %s block-upper-triangle (%lld clusters, banksize=16)\n", kind, (long long)C);
        std::fprintf(fp, "%lld %lld %lld\n", (long long)C * kDim, (long long)C * kDim, (long long)C * kSymBlocks * 9);
        for(int64_t c = 0; c < C; ++c)
        {
            const long long base = c * kDim;
            for(int br = 0; br < 16; ++br)
                for(int bc = br; bc < 16; ++bc)
                {
                    const int     k   = 16 * br - br * (br + 1) / 2 + bc;
                    const double* blk = M + (c * kSymBlocks + k) * 9;
                    for(int i = 0; i < 3; ++i)
                        for(int j = 0; j < 3; ++j)
                            std::fprintf(fp, "%lld %lld %.17g\n", base + br * 3 + i + 1, base + bc * 3 + j + 1,
                                         blk[i + 3 * j]);
                }
        }
        const bool ok = std::ferror(fp) == 0;
        std::fclose(fp);
        MI_REQUIRE(ok, MI_ERR_ARG, "mi_precond_mas_dump: write failed: " + path);
    }
    const std::string path = dir + "/mas_cluster_meta" + tag + ".json";
    FILE*             fp   = std::fopen(path.c_str(), "w");
    MI_REQUIRE(fp, MI_ERR_ARG, "mi_precond_mas_dump: cannot open " + path);
    std::fprintf(fp, "{\n    \"banksize\": 16,\n    \"frame\": %lld,\n", (long long)frame);
    write_int_array(fp, "going_next", gn.data(), total, false);
    std::fprintf(fp, "    \"level_num\": %d,\n    \"levels\": [", L);
    for(int l = 0; l <= L; ++l)
        std::fprintf(fp, "%s{\"count\": %d, \"level\": %d, \"offset\": %d}", l ? ", " : "", ls[2 * l], l,
                     ls[2 * l + 1]);
    std::fprintf(fp, "],\n    \"newton_iter\": %lld,\n    \"num_cluster_blocks\": %lld,\n", (long long)newton_iter,
                 (long long)C);
    write_int_array(fp, "part_to_real", p2r.data(), N, false);
    write_int_array(fp, "real_to_part", r2p.data(), V, false);
    std::fprintf(fp, "    \"sym_block_count\": 136,\n    \"total_clusters\": %d,\n    \"total_map_nodes\": %d,\n"
                     "    \"total_nodes\": %d\n}\n",
                 total, N, V);
    const bool ok = std::ferror(fp) == 0;
    std::fclose(fp);
    MI_REQUIRE(ok, MI_ERR_ARG, "mi_precond_mas_dump: write failed: " + path);
    MI_API_END
}

int mi_pcg_default_opts(mi_pcg_opts* o)
{
    MI_API_BEGIN
    MI_REQUIRE(o, MI_ERR_ARG, "null opts");
    o->criterion      = MI_CRIT_RZ_RATIO;
    o->tol            = 1e-3;
    o->max_iter_ratio = 2.0;
    o->check_interval = 5;
    o->use_graph      = 1;
    o->psd_fallback   = 1;
    o->force_bypass   = 0;
    o->true_residual  = 0;
    o->graph_checks   = 4;
    o->x_init         = 0;
    o->keep_order     = 0;
    MI_API_END
}
}  // extern "C"

namespace
{
// one PCG iteration (LinearFusedPCG::run_iteration, linear_fused_pcg.cu:1374-1594):
// SpMV + pAp partials -> x, r update (alpha) [+ block Jacobi + r.z partials]
// -> preconditioner -> r.z partials -> convergence, beta, p update
void run_iteration(mi_sys* sys, double* x, bool bypass, const LocalPreconditioner* fused_diag, bool ordered,
                   hipStream_t s)
{
    const int64_t n     = sys->n();
    PcgScalars*   sc    = sys->d_sc;
    const bool    use32 = sys->desc.value_precision == MI_FP32;
    double*       partA = sys->spmv_scr.partials.data();
    double*       partB = sys->dot_scr.partials.data();
    float4*       p32   = use32 && !sys->A.has_win ? sys->p32.data() : nullptr;
    const int     gA    = spmv_full(sys->A, use32, sys->p.data(), p32, sys->Ap.data(), partA, &sc->converged,
                                    sys->spmv_scr, s, nullptr, nullptr, ordered);
    int gB;
    if(fused_diag)
    {
        gB = pcg_update_xr_diag3(sc, partA, gA, x, sys->p.data(), sys->r.data(), sys->Ap.data(), sys->z.data(),
                                 fused_diag->diag3_inverse(), fused_diag->dof_offset / 3,
                                 fused_diag->dof_count / 3, n, partB, sys->dot_scr, s);
    }
    else
    {
        const bool whole = sys->pcs.size() == 1 && sys->pcs[0]->dof_offset == 0 && sys->pcs[0]->dof_count == n;
        // a preconditioner whose apply adds the r.z partials into dot slots
        // gets them cleared by the x, r update (its apply_dot refuses a
        // buffer too small for them)
        const int nz   = (whole && !bypass) ? sys->pcs[0]->dot_slots() : 0;
        double*   zero = (nz > 0 && nz <= 2 * sys->dot_scr.grid) ? partB : nullptr;
        // one preconditioner over the whole vector may fold its first pass into the x, r update
        if(bypass || !whole
           || !sys->pcs[0]->update_xr(sc, partA, gA, x, sys->p.data(), sys->r.data(), sys->Ap.data(), s, zero, nz))
            pcg_update_xr(sc, partA, gA, x, sys->p.data(), sys->r.data(), sys->Ap.data(), n, s, zero, nz);
        gB = -1;
        if(bypass)
            pcg_copy_if_active(sc, sys->z.data(), sys->r.data(), n, s);
        else if(whole)
        {
            // one preconditioner over the whole vector: its last kernel writes the r.z partials
            pcg_prepare(sys->dot_scr);
            gB = sys->pcs[0]->apply_dot(sys->r.data(), sys->z.data(), &sc->converged, partB, sys->dot_scr.grid, s);
            if(gB < 0)
                sys->pcs[0]->apply(sys->r.data(), sys->z.data(), &sc->converged, s);
        }
        else
            sys->apply_preconditioner(sys->r.data(), sys->z.data(), &sc->converged, s);
        if(gB < 0)
            gB = pcg_dot_partials(sys->r.data(), sys->z.data(), n, partB, &sc->converged, sys->dot_scr, s);
    }
    pcg_update_p(sc, partB, gB, sys->p.data(), sys->z.data(), n, p32, s);
}

void read_scalars(mi_sys* sys, hipStream_t s)
{
    MI_HIP(hipMemcpyAsync(sys->h_sc, sys->d_sc, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
}

// is the SpMV's row order another than the caller's? (a window layout built
// for the current MAS part orders)
bool spmv_reorders(const mi_sys* sys)
{
    if(!sys->A.has_win || sys->A.win_order != sys->order_gen)
        return false;
    for(auto& pc : sys->pcs)
        if(pc->spmv_rank())
            return true;
    return false;
}

// the preconditioners see the solve's vector order while it runs
struct RowOrderScope
{
    mi_sys* sys;
    RowOrderScope(mi_sys* s_, bool on) : sys(s_)
    {
        for(auto& pc : sys->pcs)
            pc->set_row_order(on);
    }
    ~RowOrderScope()
    {
        for(auto& pc : sys->pcs)
            pc->set_row_order(false);
    }
};
}  // namespace

extern "C" int mi_pcg_solve(mi_sys* sys, const double* b, double* x, const mi_pcg_opts* opts, mi_pcg_stats* st, void* stream)
{
    MI_API_BEGIN
    MI_REQUIRE(sys && b && x && st, MI_ERR_ARG, "mi_pcg_solve: null argument");
    MI_REQUIRE(sys->assembled, MI_ERR_STATE, "mi_pcg_solve: system not assembled");
    MI_REQUIRE(sys->pcs.empty() || sys->pcs_assembled, MI_ERR_STATE, "mi_pcg_solve: preconditioners not assembled");
    mi_pcg_opts o;
    mi_pcg_default_opts(&o);
    if(opts)
        o = *opts;
    MI_REQUIRE(o.tol > 0.0 && o.max_iter_ratio > 0.0, MI_ERR_ARG, "mi_pcg_solve: tol and max_iter_ratio must be > 0");
    MI_REQUIRE(o.criterion == MI_CRIT_RZ_RATIO || o.criterion == MI_CRIT_REL_RESIDUAL, MI_ERR_ARG,
               "mi_pcg_solve: unknown criterion");
    hipStream_t   s = pick(sys->ctx, stream);
    const int64_t n = sys->n();
    *st             = mi_pcg_stats{};
    st->rel_res_true = NAN;
    if(n == 0)
    {
        st->converged = 1;
        return MI_OK;
    }
    double ratio = sys->desc.reserve_ratio;
    sys->r.resize((size_t)n, ratio);
    sys->z.resize((size_t)n, ratio);
    sys->p.resize((size_t)n, ratio);
    sys->Ap.resize((size_t)n, ratio);
    if(!sys->d_sc)
    {
        MI_HIP(hipMalloc((void**)&sys->d_sc, sizeof(PcgScalars)));
        MI_HIP(hipHostMalloc((void**)&sys->h_sc, sizeof(PcgScalars)));
    }
    // scratch first: the partial buffers' addresses are taken below and baked
    // into the captured graph; no allocation may happen inside a captured block
    spmv_prepare(sys->spmv_scr, sys->desc.value_precision == MI_FP32);
    pcg_prepare(sys->dot_scr);
    PcgScalars& h = *sys->h_sc;

    // The iteration runs on vectors in the SpMV's row order when that is a
    // MAS part order: the window layout's rows and the MAS slots are then
    // read contiguously.  b is gathered into it, x scattered back at the end.
    const bool    ordered = !o.keep_order && spmv_reorders(sys);
    RowOrderScope order_scope(sys, ordered);
    const int64_t B       = sys->A.B;
    double*       xs      = x;  // the iterate, in the order the iteration runs on
    if(ordered)
    {
        sys->xo.resize((size_t)n, ratio);
        xs = sys->xo.data();
    }
    auto x_out = [&]
    {
        if(ordered)
            gather_rows3(x, xs, sys->A.wpos.data(), B, s);
    };

    // r = b, x = 0, z = P r, p = z, rz0 = r.z (linear_fused_pcg.cu:1811-1824);
    // with a given x0, r = b - A x0
    const bool use32_sys = sys->desc.value_precision == MI_FP32;
    double     bb_given  = -1.0;
    if(o.x_init)
    {
        if(ordered)
        {
            gather_rows3(xs, x, sys->A.wreal.data(), B, s);
            gather_rows3(sys->r.data(), b, sys->A.wreal.data(), B, s);
            spmv_full(sys->A, use32_sys, xs, nullptr, sys->Ap.data(), sys->spmv_scr.partials.data(), nullptr,
                      sys->spmv_scr, s, nullptr, nullptr, true);
            axpy_f64(-1.0, sys->Ap.data(), sys->r.data(), n, s);
        }
        else
        {
            copy_f64(sys->r.data(), b, n, s);
            spmv_axpby(sys->A, use32_sys, -1.0, x, 1.0, sys->r.data(), sys->spmv_scr, s);
        }
        const double bn = norm_f64_host(b, n, s);
        bb_given        = bn * bn;
    }
    else if(ordered)
    {
        gather_rows3(sys->r.data(), b, sys->A.wreal.data(), B, s);
        fill_f64(xs, 0.0, n, s);
    }
    else
        pcg_init_vectors(x, sys->r.data(), b, n, s);
    // the relative-residual reference stays ||b|| when x0 is given
    auto init_scalars = [&](int g)
    {
        pcg_init_scalars(sys->d_sc, sys->dot_scr.partials.data(), g, o.criterion, o.tol, s);
        if(bb_given >= 0.0)
            pcg_set_bb(sys->d_sc, bb_given, o.tol, s);
    };
    bool bypass = o.force_bypass != 0;
    if(bypass)
        copy_f64(sys->z.data(), sys->r.data(), n, s);
    else
        sys->apply_preconditioner(sys->r.data(), sys->z.data(), nullptr, s);
    int g0 = pcg_dot_partials(sys->r.data(), sys->z.data(), n, sys->dot_scr.partials.data(), nullptr, sys->dot_scr, s);
    init_scalars(g0);
    read_scalars(sys, s);

    // PSD guard (linear_fused_pcg.cu:1843-1875): rz0 < 0, or non-finite z with a
    // finite r, means the preconditioner is not usable -> plain CG for this solve.
    // h.rr is r0.r0 (h.bb holds ||b||^2 instead when x0 is given).
    if(!bypass && o.psd_fallback && (h.rz < 0.0 || !std::isfinite(h.rz)))
    {
        if(h.rz < 0.0 || std::isfinite(h.rr))
        {
            bypass = true;
            copy_f64(sys->z.data(), sys->r.data(), n, s);
            g0 = pcg_dot_partials(sys->r.data(), sys->z.data(), n, sys->dot_scr.partials.data(), nullptr,
                                  sys->dot_scr, s);
            init_scalars(g0);
            read_scalars(sys, s);
        }
    }
    const double rz0 = h.rz, bb = h.bb;
    st->bypassed     = bypass ? 1 : 0;
    st->rz0          = rz0;
    MI_REQUIRE(std::isfinite(rz0) && std::isfinite(bb) && std::isfinite(h.rr), MI_ERR_NONFINITE,
               "mi_pcg_solve: non-finite r^T z at start");
    copy_f64(sys->p.data(), sys->z.data(), n, s);
    if(sys->desc.value_precision == MI_FP32 && !sys->A.has_win)
    {
        sys->p32.resize((size_t)sys->desc.block_rows, ratio);
        pack_f32x4(sys->p.data(), sys->p32.data(), sys->desc.block_rows, s);
    }
    if(std::fabs(rz0) == 0.0)
    {
        st->converged = 1;
        st->rel_res   = 0.0;
        x_out();
        MI_HIP(hipStreamSynchronize(s));
        return MI_OK;
    }

    const LocalPreconditioner* fused = nullptr;
    if(!bypass && sys->pcs.size() == 1 && sys->pcs[0]->diag3_inverse())
        fused = sys->pcs[0].get();

    const int64_t max_iter = (int64_t)(o.max_iter_ratio * (double)n);
    const int64_t total    = max_iter > 0 ? max_iter - 1 : 0;
    const int     ci       = o.check_interval > 0 ? o.check_interval : 1;
    const int     gb       = ci * (o.graph_checks > 0 ? o.graph_checks : 4);  // iterations per host round trip
    int64_t       done     = 0;
    bool          converged = false;

    mi_sys::GraphKey key;
    key.x       = xs;
    key.b       = b;
    key.n       = n;
    key.U       = sys->A.U;
    key.pcs_gen = sys->pcs_gen;
    key.ci      = gb;
    key.fused   = fused ? 1 : 0;
    key.bypass  = bypass ? 1 : 0;
    key.use32   = sys->desc.value_precision;
    key.ordered = ordered ? 1 : 0;
    key.bufs[0] = sys->r.data();
    key.bufs[1] = sys->z.data();
    key.bufs[2] = sys->p.data();
    key.bufs[3] = sys->Ap.data();
    key_matrix(key, sys->A);
    key.bufs[7] = sys->pcs.empty() ? nullptr : (const void*)sys->pcs[0].get();
    key.bufs[8] = fused ? (const void*)fused->diag3_inverse() : nullptr;
    key.bufs[9] = sys->p32.data();
    key.stream  = s;

    while(done < total)
    {
        int64_t blk       = std::min<int64_t>(gb, total - done);
        bool    use_graph = o.use_graph && !sys->graph_failed && blk == gb && s != nullptr;
        if(use_graph && !(sys->graph_exec && sys->graph_key == key))
        {
            if(sys->graph_exec)
                (void)hipGraphExecDestroy(sys->graph_exec);
            if(sys->graph)
                (void)hipGraphDestroy(sys->graph);
            sys->graph_exec = nullptr;
            sys->graph      = nullptr;
            hipError_t e    = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
            if(e == hipSuccess)
            {
                for(int k = 0; k < gb; ++k)
                    run_iteration(sys, xs, bypass, fused, ordered, s);
                e = hipStreamEndCapture(s, &sys->graph);
                if(e == hipSuccess)
                    e = hipGraphInstantiate(&sys->graph_exec, sys->graph, nullptr, nullptr, 0);
            }
            if(e != hipSuccess)
            {
                (void)hipGetLastError();
                sys->graph_failed = true;  // permanent plain-launch fallback (linear_fused_pcg.cu:1748-1757)
                sys->graph_exec   = nullptr;
                use_graph         = false;
            }
            else
                sys->graph_key = key;
        }
        if(use_graph)
            MI_HIP(hipGraphLaunch(sys->graph_exec, s));
        else
            for(int64_t k = 0; k < blk; ++k)
                run_iteration(sys, xs, bypass, fused, ordered, s);
        MI_HIP(hipGetLastError());
        done += blk;
        read_scalars(sys, s);
        MI_REQUIRE(std::isfinite(h.rz_new), MI_ERR_NONFINITE,
                   "mi_pcg_solve: non-finite r^T z at iteration " + std::to_string(h.iter));
        if(h.converged)
        {
            converged = true;
            break;
        }
    }
    // the reference stops at the first check (every ci iterations) that passes
    st->iters       = converged ? std::min<int64_t>(((int64_t)h.conv_iter + ci - 1) / ci * ci, done) : max_iter;
    st->iters_exact = converged ? h.conv_iter : -1;
    st->converged   = converged ? 1 : 0;
    st->rz_final    = h.rz_new;
    st->rel_res     = bb > 0 ? std::sqrt(h.rr / bb) : 0.0;
    x_out();

    if(o.true_residual)
    {
        sys->tmp.resize((size_t)n, ratio);
        copy_f64(sys->tmp.data(), b, n, s);
        spmv_axpby(sys->A, sys->desc.value_precision == MI_FP32, -1.0, x, 1.0, sys->tmp.data(), sys->spmv_scr, s);
        double rn = norm_f64_host(sys->tmp.data(), n, s);
        double bn = std::sqrt(bb);
        st->rel_res_true = bn > 0 ? rn / bn : 0.0;
    }
    MI_API_END
}

namespace
{
// one batched iteration: the union SpMV, then per-system alpha, x / r,
// preconditioner, per-system convergence, beta and p
void run_batch_iteration(mi_sys* sys, double* x, hipStream_t s)
{
    auto&         bt    = sys->batch;
    const int32_t nseg  = (int32_t)bt.seg_rows.size() - 1;
    const bool    use32 = sys->desc.value_precision == MI_FP32;
    float4*       p32   = use32 && !sys->A.has_win ? sys->p32.data() : nullptr;
    // rows of stopped systems are skipped: their Ap is never read again
    spmv_full(sys->A, use32, sys->p.data(), p32, sys->Ap.data(), sys->spmv_scr.partials.data(), nullptr,
              sys->spmv_scr, s, bt.active.data(), bt.n_active.data());
    batch_tile_dot(bt.tiles, sys->p.data(), sys->Ap.data(), bt.part_a.data(), nullptr, s);
    batch_alpha_xr(bt.tiles, bt.sc.data(), nseg, bt.part_a.data(), x, sys->p.data(), sys->r.data(), sys->Ap.data(),
                   s);
    sys->apply_preconditioner(sys->r.data(), sys->z.data(), nullptr, s);
    if(bt.any_bypass)
        batch_bypass_z(bt.tiles, bt.sc.data(), nullptr, sys->r.data(), sys->z.data(), s);
    batch_tile_dot(bt.tiles, sys->r.data(), sys->z.data(), bt.part_a.data(), bt.part_b.data(), s);
    batch_beta_p(bt.tiles, bt.sc.data(), nseg, bt.part_a.data(), bt.part_b.data(), sys->p.data(), sys->z.data(), p32,
                 s);
}

void read_batch(mi_sys* sys, int32_t nseg, hipStream_t s)
{
    MI_HIP(hipMemcpyAsync(sys->batch.h_sc, sys->batch.sc.data(), sizeof(BatchScalars) * nseg, hipMemcpyDeviceToHost,
                          s));
    MI_HIP(hipStreamSynchronize(s));
}

// per-system sums of tile partials on the host, in tile order
std::vector<double> host_seg_sums(mi_sys* sys, const double* part, int32_t nseg, hipStream_t s)
{
    auto& bt = sys->batch;
    bt.h_part.resize((size_t)bt.tiles.n);
    MI_HIP(hipMemcpyAsync(bt.h_part.data(), part, sizeof(double) * bt.tiles.n, hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    std::vector<double> out((size_t)nseg, 0.0);
    int32_t             t = 0;
    for(int32_t g = 0; g < nseg; ++g)
    {
        const int64_t hi = 3 * (int64_t)bt.seg_rows[g + 1];
        double        acc = 0.0;
        for(int64_t lo = 3 * (int64_t)bt.seg_rows[g]; lo < hi; lo += kBatchTile)
            acc += bt.h_part[t++];
        out[g] = acc;
    }
    return out;
}
}  // namespace

extern "C" int mi_pcg_solve_batched(mi_sys*            sys,
                                    int32_t            nseg,
                                    const int32_t*     seg_rows,
                                    const double*      b,
                                    double*            x,
                                    const mi_pcg_opts* opts,
                                    mi_pcg_stats*      st,
                                    void*              stream)
{
    MI_API_BEGIN
    MI_REQUIRE(sys && seg_rows && b && x && st && nseg > 0, MI_ERR_ARG, "mi_pcg_solve_batched: null argument");
    MI_REQUIRE(sys->assembled, MI_ERR_STATE, "mi_pcg_solve_batched: system not assembled");
    MI_REQUIRE(sys->pcs.empty() || sys->pcs_assembled, MI_ERR_STATE,
               "mi_pcg_solve_batched: preconditioners not assembled");
    MI_REQUIRE(seg_rows[0] == 0 && seg_rows[nseg] == sys->desc.block_rows, MI_ERR_ARG,
               "mi_pcg_solve_batched: the segments must cover the block rows from 0");
    for(int32_t g = 0; g < nseg; ++g)
        MI_REQUIRE(seg_rows[g + 1] > seg_rows[g], MI_ERR_ARG, "mi_pcg_solve_batched: empty or decreasing segment");
    mi_pcg_opts o;
    mi_pcg_default_opts(&o);
    if(opts)
        o = *opts;
    MI_REQUIRE(o.tol > 0.0 && o.max_iter_ratio > 0.0, MI_ERR_ARG,
               "mi_pcg_solve_batched: tol and max_iter_ratio must be > 0");
    MI_REQUIRE(o.criterion == MI_CRIT_RZ_RATIO || o.criterion == MI_CRIT_REL_RESIDUAL, MI_ERR_ARG,
               "mi_pcg_solve_batched: unknown criterion");
    hipStream_t   s     = pick(sys->ctx, stream);
    const int64_t n     = sys->n();
    const double  ratio = sys->desc.reserve_ratio;
    auto&         bt    = sys->batch;

    // the tile table (rebuilt when the segmentation changes)
    if(bt.seg_rows.size() != (size_t)nseg + 1 || !std::equal(bt.seg_rows.begin(), bt.seg_rows.end(), seg_rows))
    {
        bt.seg_rows.assign(seg_rows, seg_rows + nseg + 1);
        std::vector<int64_t> lo, slo((size_t)nseg + 1);
        std::vector<int32_t> tsys, stile((size_t)nseg + 1);
        for(int32_t g = 0; g < nseg; ++g)
        {
            stile[g]        = (int32_t)lo.size();
            slo[g]          = 3 * (int64_t)seg_rows[g];
            const int64_t e = 3 * (int64_t)seg_rows[g + 1];
            for(int64_t t = slo[g]; t < e; t += kBatchTile)
            {
                lo.push_back(t);
                tsys.push_back(g);
            }
        }
        stile[nseg] = (int32_t)lo.size();
        slo[nseg]   = n;
        lo.push_back(n);
        MI_REQUIRE(lo.size() < ((size_t)1 << 31), MI_ERR_ARG, "mi_pcg_solve_batched: too many tiles");
        auto up = [&](auto& d, const auto& h)
        {
            d.resize(h.size());
            MI_HIP(hipMemcpyAsync(d.data(), h.data(), sizeof(h[0]) * h.size(), hipMemcpyHostToDevice, s));
        };
        up(bt.tile_lo, lo);
        up(bt.seg_dev, bt.seg_rows);
        up(bt.tile_sys, tsys);
        up(bt.sys_tile, stile);
        up(bt.sys_lo, slo);
        MI_HIP(hipStreamSynchronize(s));
        bt.tiles          = BatchTiles{(int32_t)tsys.size(), bt.tile_lo.data(), bt.tile_sys.data(), bt.sys_tile.data(),
                                       bt.sys_lo.data()};
        bt.part_a.resize(tsys.size());
        bt.part_b.resize(tsys.size());
        bt.sc.resize((size_t)nseg);
        if(bt.h_cap < (size_t)nseg)
        {
            if(bt.h_sc)
                (void)hipHostFree(bt.h_sc);
            bt.h_sc = nullptr;
            MI_HIP(hipHostMalloc((void**)&bt.h_sc, sizeof(BatchScalars) * nseg));
            bt.h_cap = (size_t)nseg;
        }
        ++bt.seg_gen;
    }
    for(int32_t g = 0; g < nseg; ++g)
    {
        st[g]              = mi_pcg_stats{};
        st[g].rel_res_true = NAN;
    }
    sys->r.resize((size_t)n, ratio);
    sys->z.resize((size_t)n, ratio);
    sys->p.resize((size_t)n, ratio);
    sys->Ap.resize((size_t)n, ratio);
    const bool use32 = sys->desc.value_precision == MI_FP32;
    if(use32 && !sys->A.has_win)
        sys->p32.resize((size_t)sys->desc.block_rows, ratio);
    spmv_prepare(sys->spmv_scr, use32);

    // r = b, x = 0, z = P r, p = z; per-system rz0, b^T b
    pcg_init_vectors(x, sys->r.data(), b, n, s);
    sys->apply_preconditioner(sys->r.data(), sys->z.data(), nullptr, s);
    batch_tile_dot(bt.tiles, sys->r.data(), sys->z.data(), bt.part_a.data(), bt.part_b.data(), s);
    batch_init(bt.tiles, bt.sc.data(), nseg, bt.part_a.data(), bt.part_b.data(), o.criterion, o.tol, o.max_iter_ratio,
               nullptr, s);
    read_batch(sys, nseg, s);
    // PSD guard per system (linear_fused_pcg.cu:1843-1875, run by the reference
    // once per system): rz0 < 0, or a non-finite z with a finite r, solves that
    // system as plain CG; the other systems keep their preconditioner.  A
    // non-finite r (a NaN in b) fails the call, as check_init_rz_nan_inf does.
    std::vector<int32_t> byp((size_t)nseg, 0);
    bt.any_bypass = false;
    for(int32_t g = 0; g < nseg; ++g)
    {
        const double z0 = bt.h_sc[g].rz, r0 = bt.h_sc[g].rr;
        MI_REQUIRE(std::isfinite(r0), MI_ERR_NONFINITE,
                   "mi_pcg_solve_batched: non-finite r at start of system " + std::to_string(g));
        if(o.force_bypass || (o.psd_fallback && (z0 < 0.0 || !std::isfinite(z0))))
            byp[g] = 1;
        MI_REQUIRE(byp[g] || std::isfinite(z0), MI_ERR_NONFINITE,
                   "mi_pcg_solve_batched: non-finite r^T z at start of system " + std::to_string(g));
        bt.any_bypass = bt.any_bypass || byp[g];
    }
    if(bt.any_bypass)
    {
        bt.bypass.resize((size_t)nseg);
        MI_HIP(hipMemcpyAsync(bt.bypass.data(), byp.data(), sizeof(int32_t) * nseg, hipMemcpyHostToDevice, s));
        batch_bypass_z(bt.tiles, bt.sc.data(), bt.bypass.data(), sys->r.data(), sys->z.data(), s);
        batch_tile_dot(bt.tiles, sys->r.data(), sys->z.data(), bt.part_a.data(), bt.part_b.data(), s);
        batch_init(bt.tiles, bt.sc.data(), nseg, bt.part_a.data(), bt.part_b.data(), o.criterion, o.tol,
                   o.max_iter_ratio, bt.bypass.data(), s);
        read_batch(sys, nseg, s);
    }
    copy_f64(sys->p.data(), sys->z.data(), n, s);
    if(use32 && !sys->A.has_win)
        pack_f32x4(sys->p.data(), sys->p32.data(), sys->desc.block_rows, s);
    std::vector<double> rz0((size_t)nseg), bb((size_t)nseg);
    int64_t             longest = 0;
    bool                all     = true;
    for(int32_t g = 0; g < nseg; ++g)
    {
        rz0[g] = bt.h_sc[g].rz;
        bb[g]  = bt.h_sc[g].bb;
        longest = std::max<int64_t>(longest, bt.h_sc[g].max_iter);
        all     = all && bt.h_sc[g].stopped;
        MI_REQUIRE(std::isfinite(rz0[g]) && std::isfinite(bb[g]), MI_ERR_NONFINITE,
                   "mi_pcg_solve_batched: non-finite r^T z at start of system " + std::to_string(g));
    }

    const int     ci    = o.check_interval > 0 ? o.check_interval : 1;
    const int     gb    = ci * (o.graph_checks > 0 ? o.graph_checks : 4);
    const int64_t total = longest > 0 ? longest - 1 : 0;
    mi_sys::GraphKey key;
    key.x       = x;
    key.b       = b;
    key.n       = n;
    key.U       = sys->A.U;
    key.pcs_gen = sys->pcs_gen;
    key.ci      = gb;
    key.use32   = sys->desc.value_precision;
    key.bypass  = bt.any_bypass ? 1 : 0;
    key.bufs[0] = sys->r.data();
    key.bufs[1] = sys->z.data();
    key.bufs[2] = sys->p.data();
    key.bufs[3] = sys->Ap.data();
    key_matrix(key, sys->A);
    key.bufs[7] = bt.sc.data();
    key.bufs[8] = bt.part_a.data();
    key.bufs[9] = sys->p32.data();
    // the SpMV's chunk list, rebuilt (outside the graph) whenever a system stops
    bt.active.resize((size_t)std::max(1, spmv_items(sys->A)));
    bt.n_active.resize(1);
    int32_t n_stopped = -1;
    auto    refresh   = [&](int32_t stopped_now)
    {
        if(stopped_now == n_stopped)
            return;
        batch_active_chunks(spmv_item_lo(sys->A), spmv_item_hi(sys->A), spmv_items(sys->A), bt.seg_dev.data(), nseg,
                            bt.sc.data(),
                            bt.act_flags, bt.act_scan, bt.active.data(), bt.n_active.data(), s);
        for(auto& pc : sys->pcs)
            pc->skip_stopped(bt.seg_dev.data(), nseg, bt.sc.data(), s);
        n_stopped = stopped_now;
    };
    // the preconditioners skip stopped systems only inside this solve
    struct SkipOff
    {
        mi_sys*     sys;
        hipStream_t s;
        ~SkipOff()
        {
            for(auto& pc : sys->pcs)
                pc->skip_stopped(nullptr, 0, nullptr, s);
        }
    } skip_off{sys, s};
    auto count_stopped = [&]
    {
        int32_t c = 0;
        for(int32_t g = 0; g < nseg; ++g)
            c += bt.h_sc[g].stopped ? 1 : 0;
        return c;
    };
    refresh(count_stopped());
    key.bufs[10] = bt.active.data();
    key.bufs[11] = bt.n_active.data();
    key.stream   = s;
    int64_t done = 0;
    while(!all && done < total)
    {
        const int64_t blk       = std::min<int64_t>(gb, total - done);
        bool          use_graph = o.use_graph && !sys->graph_failed && blk == gb && s != nullptr;
        if(use_graph && !(bt.graph_exec && bt.key == key && bt.key_seg_gen == bt.seg_gen))
        {
            if(bt.graph_exec)
                (void)hipGraphExecDestroy(bt.graph_exec);
            if(bt.graph)
                (void)hipGraphDestroy(bt.graph);
            bt.graph_exec = nullptr;
            bt.graph      = nullptr;
            hipError_t e  = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
            if(e == hipSuccess)
            {
                for(int k = 0; k < gb; ++k)
                    run_batch_iteration(sys, x, s);
                e = hipStreamEndCapture(s, &bt.graph);
                if(e == hipSuccess)
                    e = hipGraphInstantiate(&bt.graph_exec, bt.graph, nullptr, nullptr, 0);
            }
            if(e != hipSuccess)
            {
                (void)hipGetLastError();
                sys->graph_failed = true;
                bt.graph_exec     = nullptr;
                use_graph         = false;
            }
            else
            {
                bt.key         = key;
                bt.key_seg_gen = bt.seg_gen;
            }
        }
        if(use_graph)
            MI_HIP(hipGraphLaunch(bt.graph_exec, s));
        else
            for(int64_t k = 0; k < blk; ++k)
                run_batch_iteration(sys, x, s);
        MI_HIP(hipGetLastError());
        done += blk;
        read_batch(sys, nseg, s);
        const int32_t stopped_now = count_stopped();
        all                       = stopped_now == nseg;
        if(!all)
            refresh(stopped_now);
    }
    read_batch(sys, nseg, s);
    for(int32_t g = 0; g < nseg; ++g)
    {
        const BatchScalars& q = bt.h_sc[g];
        const bool          c = q.converged != 0;
        st[g].iters       = c ? std::min<int64_t>((q.conv_iter + ci - 1) / ci * ci, done) : q.max_iter;
        st[g].iters_exact = c ? q.conv_iter : -1;
        st[g].converged   = c ? 1 : 0;
        st[g].rz0         = rz0[g];
        st[g].bypassed    = byp[g];
        st[g].rz_final    = q.rz_new;
        st[g].rel_res     = bb[g] > 0 ? std::sqrt(q.rr / bb[g]) : 0.0;
    }
    if(o.true_residual)
    {
        sys->tmp.resize((size_t)n, ratio);
        copy_f64(sys->tmp.data(), b, n, s);
        spmv_axpby(sys->A, use32, -1.0, x, 1.0, sys->tmp.data(), sys->spmv_scr, s);
        batch_tile_dot(bt.tiles, sys->tmp.data(), sys->tmp.data(), bt.part_a.data(), nullptr, s);
        const auto rr = host_seg_sums(sys, bt.part_a.data(), nseg, s);
        for(int32_t g = 0; g < nseg; ++g)
            st[g].rel_res_true = bb[g] > 0 ? std::sqrt(rr[g] / bb[g]) : 0.0;
    }
    MI_API_END
}