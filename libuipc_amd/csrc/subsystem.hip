// The two FEM-subsystem steps on either side of the solve (gfx950):
//   before assembly, the fixed-vertex treatment of the subsystem's gradient and
//   Hessian triplets (FEMLinearSubsystem_assemble_k1/k2_kernel,
//   finite_element/fem_linear_subsystem.cu:23-53);
//   after the solve, dx_i = -x_i for every vertex of the subsystem
//   (FEMLinearSubsystem_retrieve_solution_kernel, fem_linear_subsystem.cu:119-126,
//   called from GlobalLinearSystem::Impl::distribute_solution,
//   global_linear_system.cu:531-544).
#include "linsys.hpp"

namespace mi
{
namespace
{
constexpr int kThreads = 256;

int grid_of(int64_t n)
{
    return std::max(1, std::min(ceil_div(n, kThreads), device_cu_count() * 8));
}

__global__ void fix_gradient_kernel(const int32_t* __restrict__ is_fixed, double* g, int32_t V)
{
    for(int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += gridDim.x * blockDim.x)
        if(is_fixed[i])
        {
            g[3 * (int64_t)i]     = 0.0;
            g[3 * (int64_t)i + 1] = 0.0;
            g[3 * (int64_t)i + 2] = 0.0;
        }
}

// a triplet touching a fixed vertex becomes 0 off the diagonal and I on it;
// slots with a negative row (unwritten) are left alone
__global__ void fix_hessian_kernel(const int32_t* __restrict__ rows,
                                   const int32_t* __restrict__ cols,
                                   double*        vals,
                                   int64_t        T,
                                   int32_t        block_offset,
                                   const int32_t* __restrict__ is_fixed,
                                   int32_t        V)
{
    for(int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < T; t += (int64_t)gridDim.x * blockDim.x)
    {
        const int32_t i = rows[t] - block_offset, j = cols[t] - block_offset;
        if(rows[t] < 0 || i < 0 || j < 0 || i >= V || j >= V)
            continue;
        if(is_fixed[i] || is_fixed[j])
        {
            double* v = vals + 9 * t;
#pragma unroll
            for(int q = 0; q < 9; ++q)
                v[q] = (i == j && q % 4 == 0) ? 1.0 : 0.0;
        }
    }
}

__global__ void distribute_kernel(const double* __restrict__ x, double* dxs, int64_t n)
{
    for(int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
        dxs[k] = -x[k];
}
}  // namespace

void subsystem_fix_vertices(const int32_t* rows,
                            const int32_t* cols,
                            double*        vals,
                            int64_t        T,
                            int32_t        block_offset,
                            const int32_t* is_fixed,
                            int32_t        V,
                            double*        gradient,
                            hipStream_t    s)
{
    if(V > 0 && gradient)
        fix_gradient_kernel<<<grid_of(V), kThreads, 0, s>>>(is_fixed, gradient, V);
    if(T > 0)
        fix_hessian_kernel<<<grid_of(T), kThreads, 0, s>>>(rows, cols, vals, T, block_offset, is_fixed, V);
    MI_HIP(hipGetLastError());
}

void subsystem_distribute(const double* x, double* dxs, int64_t n, hipStream_t s)
{
    if(n > 0)
        distribute_kernel<<<grid_of(n), kThreads, 0, s>>>(x, dxs, n);
    MI_HIP(hipGetLastError());
}
}  // namespace mi
