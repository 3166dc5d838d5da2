// Block-symmetric SpMV on the upper 3x3 BCOO, fused with p^T A p.
//
// Replaces Spmv::rbk_sym_spmv_dot (linear_system/spmv.cu:1023-1159; kernel
// Spmv_rbk_sym_spmv_dot_chunked_kernel<1,0>, 531-756) with the semantics of
// Spmv::cpu_sym_spmv (spmv.cu:1210-1296): y = (U + U^T - diag U) x.
//
// gfx950 design:
//   * one lane per upper block; the 9 values come from 9 SoA planes, so every
//     plane load of a wave is one contiguous 512-byte (FP64) / 256-byte (FP32)
//     stream;
//   * the row sums A x_j are combined across the wave with a segmented
//     shuffle reduction keyed by the (sorted) row, one FP64 atomic triple per
//     row segment per wave;
//   * the transposed contribution A^T x_i of an off-diagonal block is one FP64
//     atomic triple into y_j (the reference does the same, spmv.cu:683-723);
//   * p^T A p = sum_blocks x_i . (A x_j) (x2 off the diagonal) needs no read of
//     y; per-workgroup partials are summed in a fixed order by the last
//     workgroup to finish (deterministic), which also forms alpha = rz / pAp;
//   * a persistent grid; workgroups are renumbered so that each XCD (blocks are
//     dispatched round-robin over the 8 XCDs) walks one contiguous eighth of
//     the matrix and keeps its part of x in its own L2.
#include "linsys.hpp"

namespace mi
{
namespace
{
constexpr int kThreads = 256;
constexpr int kXcds    = 8;

template <typename VT>
__global__ __launch_bounds__(kThreads) void spmv_sym_dot_kernel(int64_t        U,
                                                               const int32_t* __restrict__ rows,
                                                               const int32_t* __restrict__ cols,
                                                               const VT* __restrict__ vals,
                                                               int64_t       stride,
                                                               const double* __restrict__ x,
                                                               double*       y,
                                                               double        a,
                                                               double*       partials,
                                                               uint32_t*     ticket,
                                                               double*       d_dot,
                                                               const double* d_num,
                                                               double*       d_alpha,
                                                               const int32_t* d_conv)
{
    if(d_conv && *d_conv != 0)
        return;
    __shared__ double lds[kThreads / kWave];
    __shared__ int    s_last;

    // XCD-blocked renumbering of the workgroup
    const int G   = gridDim.x;
    const int b   = blockIdx.x;
    const int per = G / kXcds;
    const int lb  = (per > 0 && b < per * kXcds) ? (b % kXcds) * per + b / kXcds : b;

    const int64_t nchunks = (U + kThreads - 1) / kThreads;
    const int64_t c0      = nchunks * lb / G;
    const int64_t c1      = nchunks * (lb + 1) / G;
    const int     lane    = threadIdx.x % kWave;

    double dot = 0.0;
    for(int64_t ch = c0; ch < c1; ++ch)
    {
        const int64_t u    = ch * kThreads + threadIdx.x;
        const bool    live = u < U;
        int32_t       i = -1, j = 0;
        double        A[9];
        if(live)
        {
            i = rows[u];
            j = cols[u];
#pragma unroll
            for(int q = 0; q < 9; ++q)
                A[q] = (double)vals[q * stride + u];
        }
        else
        {
#pragma unroll
            for(int q = 0; q < 9; ++q)
                A[q] = 0.0;
        }
        double xj0 = 0, xj1 = 0, xj2 = 0, xi0 = 0, xi1 = 0, xi2 = 0;
        if(live)
        {
            xj0 = x[3 * (int64_t)j + 0];
            xj1 = x[3 * (int64_t)j + 1];
            xj2 = x[3 * (int64_t)j + 2];
            xi0 = x[3 * (int64_t)i + 0];
            xi1 = x[3 * (int64_t)i + 1];
            xi2 = x[3 * (int64_t)i + 2];
        }
        // column-major: A(r, c) = A[r + 3 c]
        double r0 = A[0] * xj0 + A[3] * xj1 + A[6] * xj2;
        double r1 = A[1] * xj0 + A[4] * xj1 + A[7] * xj2;
        double r2 = A[2] * xj0 + A[5] * xj1 + A[8] * xj2;
        double d  = xi0 * r0 + xi1 * r1 + xi2 * r2;
        dot += (live && i != j) ? 2.0 * d : d;
        if(live && i != j)
        {
            double w0 = A[0] * xi0 + A[1] * xi1 + A[2] * xi2;
            double w1 = A[3] * xi0 + A[4] * xi1 + A[5] * xi2;
            double w2 = A[6] * xi0 + A[7] * xi1 + A[8] * xi2;
            double* yj = y + 3 * (int64_t)j;
            atomic_add_f64(yj + 0, a * w0);
            atomic_add_f64(yj + 1, a * w1);
            atomic_add_f64(yj + 2, a * w2);
        }
        // segmented suffix reduction of the row sums over the wave
#pragma unroll
        for(int o = 1; o < kWave; o <<= 1)
        {
            double  s0 = __shfl_down(r0, o, kWave);
            double  s1 = __shfl_down(r1, o, kWave);
            double  s2 = __shfl_down(r2, o, kWave);
            int32_t io = __shfl_down(i, o, kWave);
            if(lane + o < kWave && io == i)
            {
                r0 += s0;
                r1 += s1;
                r2 += s2;
            }
        }
        int32_t iprev = __shfl_up(i, 1, kWave);
        if(live && (lane == 0 || iprev != i))
        {
            double* yi = y + 3 * (int64_t)i;
            atomic_add_f64(yi + 0, a * r0);
            atomic_add_f64(yi + 1, a * r1);
            atomic_add_f64(yi + 2, a * r2);
        }
    }

    if(!d_dot)
        return;
    double bs = block_sum<kThreads>(dot, lds);
    if(threadIdx.x == 0)
    {
        partials[blockIdx.x] = bs;
        __threadfence();
        uint32_t t = atomicAdd(ticket, 1u);
        s_last     = (t == (uint32_t)G - 1u);
    }
    __syncthreads();
    if(s_last)
    {
        __threadfence();
        double acc = 0.0;
        for(int k = threadIdx.x; k < G; k += kThreads)
            acc += __hip_atomic_load(&partials[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        acc = block_sum<kThreads>(acc, lds);
        if(threadIdx.x == 0)
        {
            acc *= a;
            *d_dot = acc;
            if(d_num)
                *d_alpha = *d_num / acc;
            *ticket = 0u;
        }
    }
}

__global__ void scale_kernel(double* y, double b, int64_t n)
{
    for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = (b == 0.0) ? 0.0 : b * y[i];
}
}  // namespace

void spmv_prepare(SpmvScratch& scr)
{
    if(scr.grid == 0)
    {
        scr.grid = device_cu_count() * 4;
        scr.grid = (scr.grid / kXcds) * kXcds;
        if(scr.grid < kXcds)
            scr.grid = kXcds;
        scr.partials.resize((size_t)scr.grid);
        scr.ticket.resize(4);
        MI_HIP(hipMemset(scr.ticket.data(), 0, 4 * sizeof(uint32_t)));
    }
}

static int spmv_grid(const SpmvScratch& scr, int64_t U)
{
    int64_t nchunks = (U + kThreads - 1) / kThreads;
    int     g       = scr.grid;
    if(nchunks < g)
        g = (int)std::max<int64_t>(1, nchunks);
    return g;
}

void spmv_sym_dot(const BsrMatrix& A,
                  bool             use32,
                  const double*    x,
                  double*          y,
                  double*          d_dot,
                  double*          d_num,
                  double*          d_alpha,
                  const int32_t*   d_conv,
                  SpmvScratch&     scr,
                  hipStream_t      s)
{
    spmv_prepare(scr);
    if(A.U == 0)
    {
        if(d_dot)
            MI_HIP(hipMemsetAsync(d_dot, 0, sizeof(double), s));
        return;
    }
    int g = spmv_grid(scr, A.U);
    if(use32)
    {
        MI_REQUIRE(A.has32, MI_ERR_STATE, "spmv: FP32 values were not assembled");
        spmv_sym_dot_kernel<float><<<g, kThreads, 0, s>>>(A.U,
                                                          A.rows.data(),
                                                          A.cols.data(),
                                                          A.vals32.data(),
                                                          A.stride,
                                                          x,
                                                          y,
                                                          1.0,
                                                          scr.partials.data(),
                                                          scr.ticket.data(),
                                                          d_dot,
                                                          d_num,
                                                          d_alpha,
                                                          d_conv);
    }
    else
    {
        spmv_sym_dot_kernel<double><<<g, kThreads, 0, s>>>(A.U,
                                                           A.rows.data(),
                                                           A.cols.data(),
                                                           A.vals.data(),
                                                           A.stride,
                                                           x,
                                                           y,
                                                           1.0,
                                                           scr.partials.data(),
                                                           scr.ticket.data(),
                                                           d_dot,
                                                           d_num,
                                                           d_alpha,
                                                           d_conv);
    }
    MI_HIP(hipGetLastError());
}

void spmv_axpby(const BsrMatrix& A, double a, const double* x, double b, double* y, SpmvScratch& scr, hipStream_t s)
{
    spmv_prepare(scr);
    int64_t n = 3 * (int64_t)A.B;
    if(n > 0)
        scale_kernel<<<std::max(1, std::min(ceil_div(n, 256), device_cu_count() * 8)), 256, 0, s>>>(y, b, n);
    if(A.U == 0 || a == 0.0)
        return;
    int g = spmv_grid(scr, A.U);
    spmv_sym_dot_kernel<double><<<g, kThreads, 0, s>>>(A.U,
                                                       A.rows.data(),
                                                       A.cols.data(),
                                                       A.vals.data(),
                                                       A.stride,
                                                       x,
                                                       y,
                                                       a,
                                                       scr.partials.data(),
                                                       scr.ticket.data(),
                                                       nullptr,
                                                       nullptr,
                                                       nullptr,
                                                       nullptr);
    MI_HIP(hipGetLastError());
}
}  // namespace mi
