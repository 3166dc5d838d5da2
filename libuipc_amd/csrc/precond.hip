// Block-Jacobi local preconditioners on gfx950.
//
//   Diag3Preconditioner  replaces FEMDiagPreconditioner
//                        (finite_element/fem_diag_preconditioner.cu:18-117):
//                        inverse of each 3x3 diagonal block.
//   Diag12Preconditioner replaces ABDDiagPreconditioner
//                        (affine_body/abd_diag_preconditioner.cu:15-122):
//                        inverse of each affine body's 12x12 diagonal Hessian
//                        (its 4x4 upper 3x3 blocks, mirrored).
//
// Both find their blocks through the BSR row offsets: a row's diagonal block
// is its first block (cols are ascending and >= row in the upper storage).
#include "linsys.hpp"

namespace mi
{
namespace
{
constexpr int kThreads = 256;

__device__ __forceinline__ void inv3(const double m[9], double o[9])
{
    double c00 = m[4] * m[8] - m[7] * m[5];
    double c01 = m[7] * m[2] - m[1] * m[8];
    double c02 = m[1] * m[5] - m[4] * m[2];
    double det = m[0] * c00 + m[3] * c01 + m[6] * c02;
    double id  = 1.0 / det;
    o[0]       = c00 * id;
    o[1]       = c01 * id;
    o[2]       = c02 * id;
    o[3]       = (m[6] * m[5] - m[3] * m[8]) * id;
    o[4]       = (m[0] * m[8] - m[6] * m[2]) * id;
    o[5]       = (m[3] * m[2] - m[0] * m[5]) * id;
    o[6]       = (m[3] * m[7] - m[6] * m[4]) * id;
    o[7]       = (m[6] * m[1] - m[0] * m[7]) * id;
    o[8]       = (m[0] * m[4] - m[3] * m[1]) * id;
}

__global__ void diag3_assemble_kernel(const int32_t* row_ptr,
                                      const int32_t* cols,
                                      const double*  vals,
                                      int64_t        blk_off,
                                      int64_t        blk_cnt,
                                      double*        D)
{
    for(int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < blk_cnt; k += (int64_t)gridDim.x * blockDim.x)
    {
        int64_t row = blk_off + k;
        int32_t u   = row_ptr[row];
        double  o[9];
        if(u < row_ptr[row + 1] && cols[u] == row)
        {
            double m[9];
#pragma unroll
            for(int q = 0; q < 9; ++q)
                m[q] = vals[upper_at(u, q)];
            inv3(m, o);
        }
        else
        {
#pragma unroll
            for(int q = 0; q < 9; ++q)
                o[q] = 0.0;
        }
#pragma unroll
        for(int q = 0; q < 9; ++q)
            D[9 * k + q] = o[q];
    }
}

__global__ void diag3_apply_kernel(const double* D, const double* r, double* z, const int32_t* conv, int64_t blk_cnt)
{
    if(conv && *conv != 0)
        return;
    for(int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < blk_cnt; k += (int64_t)gridDim.x * blockDim.x)
    {
        const double* Di = D + 9 * k;
        double        r0 = r[3 * k], r1 = r[3 * k + 1], r2 = r[3 * k + 2];
#pragma unroll
        for(int a = 0; a < 3; ++a)
            z[3 * k + a] = Di[a] * r0 + Di[a + 3] * r1 + Di[a + 6] * r2;
    }
}

// one wave per body (four bodies per workgroup): gather the 12x12 symmetric
// body Hessian into [H | I] in LDS and invert it by Gauss-Jordan with
// partial pivoting (Eigen's inverse of a fixed 12x12 goes through a
// partial-pivot LU; same operator).  Every entry sees the operations and the
// order of a serial sweep: pivot = first row of the largest |A[r][c]|, row c
// scaled by 1 / A[c][c], then A[r][k] -= A[r][c] A[c][k] for r != c.  (A
// thread per body kept the 12 x 24 array in scratch memory: 1.9 ms for
// 10,000 bodies.)
__global__ __launch_bounds__(256) void diag12_assemble_kernel(const int32_t* row_ptr,
                                                              const int32_t* cols,
                                                              const double*  vals,
                                                              int64_t        blk_off,
                                                              int64_t        bodies,
                                                              double*        Hinv)
{
    constexpr int     kBodies = 256 / kWave;
    __shared__ double S[kBodies][12][25];
    const int         w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int64_t     bi     = (int64_t)blockIdx.x * kBodies + w;
    const bool        active = bi < bodies;  // every wave reaches every barrier
    double (*A)[25]          = S[w];
    const int64_t r0         = blk_off + 4 * bi;
    if(active)
        for(int e = lane; e < 12 * 24; e += kWave)
        {
            const int i = e / 24, j = e % 24;
            A[i][j]     = (j - 12 == i) ? 1.0 : 0.0;
        }
    __syncthreads();
    // lane 4 p + q (q >= p) gathers upper block (p, q) and its transpose
    if(active && lane < 16 && lane % 4 >= lane / 4)
    {
        const int     p = lane / 4, q = lane % 4;
        const int64_t row = r0 + p;
        for(int32_t u = row_ptr[row]; u < row_ptr[row + 1]; ++u)
        {
            const int64_t c = cols[u];
            if(c > r0 + q)
                break;
            if(c < r0 + q)
                continue;
            for(int a = 0; a < 3; ++a)
                for(int b = 0; b < 3; ++b)
                {
                    const double v          = vals[upper_at(u, a + 3 * b)];
                    A[3 * p + a][3 * q + b] = v;
                    A[3 * q + b][3 * p + a] = v;
                }
            break;
        }
    }
    __syncthreads();
    for(int c = 0; c < 12; ++c)
    {
        int    piv = c;
        double pv  = active ? fabs(A[c][c]) : 0.0;
        for(int r = c + 1; active && r < 12; ++r)
            if(fabs(A[r][c]) > pv)
            {
                pv  = fabs(A[r][c]);
                piv = r;
            }
        __syncthreads();
        if(active && piv != c && lane < 24)
        {
            const double t = A[c][lane];
            A[c][lane]     = A[piv][lane];
            A[piv][lane]   = t;
        }
        __syncthreads();
        const double d = active ? 1.0 / A[c][c] : 0.0;
        __syncthreads();
        if(active && lane < 24)
            A[c][lane] *= d;
        __syncthreads();
        // the multipliers A[r][c] are read before any of them changes
        double f[(12 * 24 + kWave - 1) / kWave];
#pragma unroll
        for(int t = 0; t < (12 * 24 + kWave - 1) / kWave; ++t)
        {
            const int e = lane + t * kWave;
            f[t]        = (active && e < 12 * 24) ? A[e / 24][c] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for(int t = 0; t < (12 * 24 + kWave - 1) / kWave; ++t)
        {
            const int e = lane + t * kWave;
            if(active && e < 12 * 24)
            {
                const int r = e / 24, k = e % 24;
                if(r != c)
                    A[r][k] -= f[t] * A[c][k];
            }
        }
        __syncthreads();
    }
    if(active)
    {
        double* H = Hinv + 144 * bi;
        for(int e = lane; e < 144; e += kWave)
        {
            const int i = e % 12, j = e / 12;
            H[i + 12 * j] = A[i][12 + j];
        }
    }
}

// 12 lanes per body (abd_diag_preconditioner's _lanes variant)
__global__ void diag12_apply_kernel(const double* Hinv, const double* r, double* z, const int32_t* conv, int64_t bodies)
{
    if(conv && *conv != 0)
        return;
    for(int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < 12 * bodies; t += (int64_t)gridDim.x * blockDim.x)
    {
        int64_t       bi = t / 12;
        int           a  = (int)(t % 12);
        const double* H  = Hinv + 144 * bi;
        const double* rb = r + 12 * bi;
        double        acc = 0.0;
#pragma unroll
        for(int c = 0; c < 12; ++c)
            acc += H[a + 12 * c] * rb[c];
        z[12 * bi + a] = acc;
    }
}

int grid_for(int64_t n)
{
    return std::max(1, std::min(ceil_div(n, kThreads), device_cu_count() * 8));
}

class Diag3Preconditioner final : public LocalPreconditioner
{
  public:
    DBuf<double> D;
    void         assemble(const BsrMatrix& A, hipStream_t s) override
    {
        int64_t off = dof_offset / 3, cnt = dof_count / 3;
        MI_REQUIRE(off + cnt <= A.B, MI_ERR_ARG, "diag3: subsystem range exceeds the matrix");
        D.resize((size_t)std::max<int64_t>(1, 9 * cnt));
        if(cnt > 0)
            diag3_assemble_kernel<<<grid_for(cnt), kThreads, 0, s>>>(
                A.row_ptr.data(), A.cols.data(), A.vals.data(), off, cnt, D.data());
        MI_HIP(hipGetLastError());
    }
    void apply(const double* r, double* z, const int32_t* conv, hipStream_t s) override
    {
        int64_t cnt = dof_count / 3;
        if(cnt > 0)
            diag3_apply_kernel<<<grid_for(cnt), kThreads, 0, s>>>(D.data(), r, z, conv, cnt);
    }
    const double* diag3_inverse() const override { return D.data(); }
};

class Diag12Preconditioner final : public LocalPreconditioner
{
  public:
    DBuf<double> H;
    void         assemble(const BsrMatrix& A, hipStream_t s) override
    {
        int64_t off = dof_offset / 3, bodies = dof_count / 12;
        MI_REQUIRE(off + 4 * bodies <= A.B, MI_ERR_ARG, "diag12: subsystem range exceeds the matrix");
        H.resize((size_t)std::max<int64_t>(1, 144 * bodies));
        if(bodies > 0)
            diag12_assemble_kernel<<<(unsigned)ceil_div(bodies, (int64_t)4), 256, 0, s>>>(
                A.row_ptr.data(), A.cols.data(), A.vals.data(), off, bodies, H.data());
        MI_HIP(hipGetLastError());
    }
    void apply(const double* r, double* z, const int32_t* conv, hipStream_t s) override
    {
        int64_t bodies = dof_count / 12;
        if(bodies > 0)
            diag12_apply_kernel<<<grid_for(12 * bodies), kThreads, 0, s>>>(H.data(), r, z, conv, bodies);
    }
};
}  // namespace

std::unique_ptr<LocalPreconditioner> make_diag3(int64_t off, int64_t cnt)
{
    MI_REQUIRE(off % 3 == 0 && cnt % 3 == 0 && off >= 0 && cnt >= 0, MI_ERR_ARG, "diag3: DoF range must be 3-aligned");
    auto p        = std::make_unique<Diag3Preconditioner>();
    p->dof_offset = off;
    p->dof_count  = cnt;
    return p;
}

std::unique_ptr<LocalPreconditioner> make_diag12(int64_t off, int64_t cnt)
{
    MI_REQUIRE(off % 3 == 0 && cnt % 12 == 0 && off >= 0 && cnt >= 0, MI_ERR_ARG,
               "diag12: DoF offset must be 3-aligned and the count a multiple of 12");
    auto p        = std::make_unique<Diag12Preconditioner>();
    p->dof_offset = off;
    p->dof_count  = cnt;
    return p;
}
}  // namespace mi
