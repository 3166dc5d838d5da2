// Block-symmetric SpMV fused with x^T A x, on gfx950.
//
// Replaces Spmv::rbk_sym_spmv_dot (linear_system/spmv.cu:1023-1159; kernel
// Spmv_rbk_sym_spmv_dot_chunked_kernel<1,0>, 531-756) with the semantics of
// Spmv::cpu_sym_spmv (spmv.cu:1210-1296): y = (U + U^T - diag U) x.
//
// gfx950 design: the reference accumulates A^T x_i of every upper block into
// y_j with FP64 atomics (spmv.cu:683-723).  On MI355X the device-scope atomics
// of eight non-coherent XCD L2s serialise badly, so the assembly also builds
// the full BSR (both triangles, BsrMatrix::full_*) and the SpMV is a pure row
// gather:
//   * 16 lanes per block row stride over the row's blocks; each of the 9 SoA
//     value planes is read as 128 contiguous bytes per row group;
//   * the row sum is reduced across the 16 lanes with xor shuffles and stored
//     once (no atomics, deterministic);
//   * x^T A x = sum_i x_i . (A x)_i per workgroup into partials[blockIdx.x];
//     the consumer kernel (PCG update) sums the partials in a fixed order;
//   * a persistent grid; workgroups are renumbered so that each XCD (blocks are
//     dispatched round-robin over the 8 XCDs) walks one contiguous eighth of
//     the rows and keeps its part of x in its own L2.
#include <cstdlib>

#include "linsys.hpp"

namespace mi
{
namespace
{
constexpr int kThreads = 256;
constexpr int kLanes   = 16;  // lanes per block row
constexpr int kGroups  = kThreads / kLanes;
constexpr int kXcds    = 8;

typedef int   v2i __attribute__((ext_vector_type(2)));
typedef float v2f __attribute__((ext_vector_type(2)));

template <typename VT, bool ACCUM, int UNROLL>
__global__ __launch_bounds__(kThreads) void spmv_full_kernel(int32_t        B,
                                                            const int32_t* __restrict__ fptr,
                                                            const int32_t* __restrict__ fcols,
                                                            const VT* __restrict__ fvals,
                                                            int64_t       fstride,
                                                            const double* __restrict__ x,
                                                            double*       y,
                                                            double        a,
                                                            double*       partials,
                                                            const int32_t* d_conv)
{
    if(d_conv && *d_conv != 0)
        return;
    __shared__ double lds[kThreads / kWave];

    const int G   = gridDim.x;
    const int b   = blockIdx.x;
    const int per = G / kXcds;
    const int lb  = (per > 0 && b < per * kXcds) ? (b % kXcds) * per + b / kXcds : b;
    const int64_t r0 = (int64_t)B * lb / G;
    const int64_t r1 = (int64_t)B * (lb + 1) / G;

    const int g    = threadIdx.x / kLanes;
    const int lane = threadIdx.x % kLanes;
    double    dot  = 0.0;
    for(int64_t row = r0 + g; row < r1; row += kGroups)
    {
        const int32_t s = fptr[row], e = fptr[row + 1];
        double        acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
        // UNROLL blocks per lane per trip: all column indices, value sets and
        // x gathers are in flight before the first FMA.  Plane bases are
        // uniform and offsets 32-bit (scalar base + vector offset addressing).
        if constexpr(sizeof(VT) == 8)
        {
            for(int32_t k = s + lane; k < e; k += UNROLL * kLanes)
            {
                uint32_t jj[UNROLL];
                bool     ok[UNROLL];
#pragma unroll
                for(int u = 0; u < UNROLL; ++u)
                {
                    const int32_t ku = k + u * kLanes;
                    ok[u]            = ku < e;
                    jj[u]            = ok[u] ? (uint32_t)fcols[ku] : 0u;
                }
                VT A[UNROLL][9];
#pragma unroll
                for(int u = 0; u < UNROLL; ++u)
#pragma unroll
                    for(int q = 0; q < 9; ++q)
                        A[u][q] = ok[u] ? fvals[q * fstride + k + u * kLanes] : VT(0);
                double xv[UNROLL][3];
#pragma unroll
                for(int u = 0; u < UNROLL; ++u)
#pragma unroll
                    for(int c = 0; c < 3; ++c)
                        xv[u][c] = x[3u * jj[u] + c];
#pragma unroll
                for(int u = 0; u < UNROLL; ++u)
                {
                    acc0 += (double)A[u][0] * xv[u][0] + (double)A[u][3] * xv[u][1] + (double)A[u][6] * xv[u][2];
                    acc1 += (double)A[u][1] * xv[u][0] + (double)A[u][4] * xv[u][1] + (double)A[u][7] * xv[u][2];
                    acc2 += (double)A[u][2] * xv[u][0] + (double)A[u][5] * xv[u][1] + (double)A[u][8] * xv[u][2];
                }
            }
        }
        else
        {
            // FP32 planes: each lane takes two neighbouring blocks with 8-byte
            // loads, so 16 lanes read 128 contiguous bytes per plane.  The pair
            // grid starts at the even block at or before the row start; the
            // halves outside [s, e) are masked.  The planes are 8-byte aligned
            // (the plane stride is even).
            const int32_t s2 = s & ~1;
            for(int32_t k = s2 + 2 * lane; k < e; k += UNROLL * 2 * kLanes)
            {
                uint32_t jj[UNROLL][2];
                bool     ok[UNROLL][2];
#pragma unroll
                for(int u = 0; u < UNROLL; ++u)
                {
                    const int32_t ku = k + u * 2 * kLanes;
                    ok[u][0]         = ku >= s && ku < e;
                    ok[u][1]         = ku + 1 < e;
                    const bool any   = ku < e;
                    const v2i  cc    = any ? *reinterpret_cast<const v2i*>(fcols + ku) : v2i{0, 0};
                    jj[u][0]         = ok[u][0] ? (uint32_t)cc.x : 0u;
                    jj[u][1]         = ok[u][1] ? (uint32_t)cc.y : 0u;
                }
                v2f A[UNROLL][9];
#pragma unroll
                for(int u = 0; u < UNROLL; ++u)
                {
                    const int32_t ku  = k + u * 2 * kLanes;
                    const bool    any = ku < e;
#pragma unroll
                    for(int q = 0; q < 9; ++q)
                        A[u][q] = any ? *reinterpret_cast<const v2f*>(fvals + q * fstride + ku) : v2f{0.f, 0.f};
                }
#pragma unroll
                for(int u = 0; u < UNROLL; ++u)
#pragma unroll
                    for(int h = 0; h < 2; ++h)
                    {
                        if(!ok[u][h])
                            continue;
                        const double x0 = x[3u * jj[u][h]], x1 = x[3u * jj[u][h] + 1], x2 = x[3u * jj[u][h] + 2];
                        float        a[9];
#pragma unroll
                        for(int q = 0; q < 9; ++q)
                            a[q] = h ? A[u][q].y : A[u][q].x;
                        acc0 += (double)a[0] * x0 + (double)a[3] * x1 + (double)a[6] * x2;
                        acc1 += (double)a[1] * x0 + (double)a[4] * x1 + (double)a[7] * x2;
                        acc2 += (double)a[2] * x0 + (double)a[5] * x1 + (double)a[8] * x2;
                    }
            }
        }
#pragma unroll
        for(int o = kLanes / 2; o > 0; o >>= 1)
        {
            acc0 += __shfl_xor(acc0, o, kWave);
            acc1 += __shfl_xor(acc1, o, kWave);
            acc2 += __shfl_xor(acc2, o, kWave);
        }
        if(lane == 0)
        {
            double* yr = y + 3 * row;
            if constexpr(ACCUM)
            {
                yr[0] += a * acc0;
                yr[1] += a * acc1;
                yr[2] += a * acc2;
            }
            else
            {
                yr[0] = acc0;
                yr[1] = acc1;
                yr[2] = acc2;
            }
            dot += x[3 * row] * acc0 + x[3 * row + 1] * acc1 + x[3 * row + 2] * acc2;
        }
    }
    if(partials)
    {
        double bs = block_sum<kThreads>(dot, lds);
        if(threadIdx.x == 0)
            partials[blockIdx.x] = bs;
    }
}

__global__ void scale_kernel(double* y, double b, int64_t n)
{
    for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = (b == 0.0) ? 0.0 : b * y[i];
}

// sum of partials -> *out (one workgroup), for the ABI's spmv_dot
__global__ __launch_bounds__(kThreads) void sum_partials_kernel(const double* partials, int n, double* out)
{
    __shared__ double lds[kThreads / kWave];
    double            acc = 0.0;
    for(int k = threadIdx.x; k < n; k += kThreads)
        acc += partials[k];
    acc = block_sum<kThreads>(acc, lds);
    if(threadIdx.x == 0)
        *out = acc;
}

// blocks per lane per trip of the row loop: 4 for FP64 planes, 1 (pairs) for
// FP32; MI_SPMV_UNROLL = 1, 2 or 4 overrides for launch sweeps
int spmv_unroll(bool use32)
{
    int v = use32 ? 1 : 4;
    if(const char* e = std::getenv("MI_SPMV_UNROLL"))
    {
        const int w = std::atoi(e);
        if(w == 1 || w == 2 || w == 4)
            v = w;
    }
    return v;
}

template <typename VT, bool ACCUM>
const void* spmv_kernel_of(int unroll)
{
    return unroll == 1   ? (const void*)spmv_full_kernel<VT, ACCUM, 1>
           : unroll == 4 ? (const void*)spmv_full_kernel<VT, ACCUM, 4>
                         : (const void*)spmv_full_kernel<VT, ACCUM, 2>;
}

template <typename VT, bool ACCUM>
void launch_spmv(const SpmvScratch& scr, int g, hipStream_t s, const BsrMatrix& A, const VT* vals, const double* x,
                 double* y, double a, double* partials, const int32_t* d_conv)
{
    switch(scr.unroll)
    {
        case 1:
            spmv_full_kernel<VT, ACCUM, 1><<<g, kThreads, 0, s>>>(A.B, A.full_ptr.data(), A.fcols.data(), vals,
                                                                  A.fstride, x, y, a, partials, d_conv);
            break;
        case 4:
            spmv_full_kernel<VT, ACCUM, 4><<<g, kThreads, 0, s>>>(A.B, A.full_ptr.data(), A.fcols.data(), vals,
                                                                  A.fstride, x, y, a, partials, d_conv);
            break;
        default:
            spmv_full_kernel<VT, ACCUM, 2><<<g, kThreads, 0, s>>>(A.B, A.full_ptr.data(), A.fcols.data(), vals,
                                                                  A.fstride, x, y, a, partials, d_conv);
    }
}
}  // namespace

void spmv_prepare(SpmvScratch& scr, bool use32)
{
    if(scr.grid == 0 || scr.use32 != use32)
    {
        scr.use32  = use32;
        scr.unroll = spmv_unroll(use32);
        // a persistent grid of as many 4-wave workgroups per CU as fit at once
        // (the kernel's occupancy); MI_SPMV_BLOCKS_PER_CU overrides for sweeps
        int per_cu = 0;
        if(const char* e = std::getenv("MI_SPMV_BLOCKS_PER_CU"))
            per_cu = std::max(1, std::min(32, std::atoi(e)));
        if(per_cu == 0)
        {
            const void* fn = use32 ? spmv_kernel_of<float, false>(scr.unroll) : spmv_kernel_of<double, false>(scr.unroll);
            MI_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kThreads, 0));
            per_cu = std::max(1, per_cu);
        }
        int grid = device_cu_count() * per_cu;
        grid     = std::max(kXcds, (grid / kXcds) * kXcds);
        scr.partials.resize((size_t)2 * grid);
        scr.grid = grid;
    }
}

int spmv_grid(const SpmvScratch& scr, int32_t B)
{
    int64_t need = ((int64_t)B + kGroups - 1) / kGroups;
    int     g    = scr.grid;
    if(need < g)
        g = (int)std::max<int64_t>(1, need);
    return g;
}

int spmv_full(const BsrMatrix& A, bool use32, const double* x, double* y, double* partials, const int32_t* d_conv,
              SpmvScratch& scr, hipStream_t s)
{
    spmv_prepare(scr, use32);
    const int g = spmv_grid(scr, A.B);
    if(A.B == 0)
        return 0;
    MI_REQUIRE(x && y && partials && A.full_ptr.data() && A.fcols.data(), MI_ERR_STATE, "spmv_full: null buffer");
    MI_REQUIRE(partials >= scr.partials.data() && partials + g <= scr.partials.data() + scr.partials.size(),
               MI_ERR_STATE, "spmv_full: partials outside the scratch");
    if(use32)
    {
        MI_REQUIRE(A.has32, MI_ERR_STATE, "spmv: FP32 values were not assembled");
        launch_spmv<float, false>(scr, g, s, A, A.fvals32.data(), x, y, 1.0, partials, d_conv);
    }
    else
        launch_spmv<double, false>(scr, g, s, A, A.fvals.data(), x, y, 1.0, partials, d_conv);
    MI_HIP(hipGetLastError());
    return g;
}

void spmv_accumulate_dot(const BsrMatrix& A, bool use32, double a, const double* x, double* y, double* d_dot,
                         SpmvScratch& scr, hipStream_t s)
{
    spmv_prepare(scr, use32);
    if(A.B == 0)
    {
        if(d_dot)
            fill_f64(d_dot, 0.0, 1, s);
        return;
    }
    const int g = spmv_grid(scr, A.B);
    double*   p = d_dot ? scr.partials.data() : nullptr;
    if(use32)
    {
        MI_REQUIRE(A.has32, MI_ERR_STATE, "spmv: FP32 values were not assembled");
        launch_spmv<float, true>(scr, g, s, A, A.fvals32.data(), x, y, a, p, nullptr);
    }
    else
        launch_spmv<double, true>(scr, g, s, A, A.fvals.data(), x, y, a, p, nullptr);
    if(d_dot)
        sum_partials_kernel<<<1, kThreads, 0, s>>>(scr.partials.data(), g, d_dot);
    MI_HIP(hipGetLastError());
}

void spmv_axpby(const BsrMatrix& A, bool use32, double a, const double* x, double b, double* y, SpmvScratch& scr,
                hipStream_t s)
{
    int64_t n = 3 * (int64_t)A.B;
    if(n > 0 && b != 1.0)
        scale_kernel<<<std::max(1, std::min(ceil_div(n, 256), device_cu_count() * 8)), 256, 0, s>>>(y, b, n);
    if(a != 0.0)
        spmv_accumulate_dot(A, use32, a, x, y, nullptr, scr, s);
}
}  // namespace mi
