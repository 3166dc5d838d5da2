// Learned initial guess for the PCG (no counterpart in the reference, whose
// solve starts from x0 = 0): the device kernels of a small message-passing
// network over the block graph of the assembled system, for its inference
// and its training.
//
//   gemm    C = op(A) op(B) (+ C) (+ bias) (relu), BF16 operands on the matrix
//           cores (v_mfma_f32_16x16x16_bf16), FP32 accumulation and storage;
//           one wave per 16x16 tile of C, strides give the transposes
//   agg     Z_i = sum_j w_ij H_j over a CSR graph (one wave per row, one lane
//           per feature); the weights are symmetric, so the same kernel is
//           its own adjoint
//   relu'   dA = dH * (A > 0)
//   colsum  the bias gradients
//   adam    the optimiser step on FP32 master weights
//   blk3    y_i = s * M_i v_i for 3x3 blocks M_i (the output transform and
//           its adjoint)
#include "linsys.hpp"

namespace mi
{
namespace
{
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// round-to-nearest-even FP32 -> BF16 bits
__device__ __forceinline__ short bf16_bits(float f)
{
    unsigned int u = __float_as_uint(f);
    if((u & 0x7F800000u) != 0x7F800000u)
        u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

// v_mfma_f32_16x16x16_bf16 lane layout (lane l, g = l / 16, c = l % 16):
// A operand A[c][4 g + e], B operand B[4 g + e][c], accumulator v holds
// D[4 g + v][c] (checked on gfx950 with distinct operands)
__global__ __launch_bounds__(64) void ws_gemm_kernel(int32_t      M,
                                                     int32_t      N,
                                                     int32_t      K,
                                                     const float* A,
                                                     int64_t      sa_m,
                                                     int64_t      sa_k,
                                                     const float* B,
                                                     int64_t      sb_k,
                                                     int64_t      sb_n,
                                                     float*       C,
                                                     int64_t      ldc,
                                                     int32_t      accumulate,
                                                     const float* bias,
                                                     int32_t      relu)
{
    const int     tiles_n = (N + 15) / 16;
    const int64_t tile    = blockIdx.x;
    const int     m0 = (int)(tile / tiles_n) * 16, n0 = (int)(tile % tiles_n) * 16;
    const int     l = threadIdx.x, g = l / 16, c = l % 16;
    f32x4         acc = {0.f, 0.f, 0.f, 0.f};
    for(int k0 = 0; k0 < K; k0 += 16)
    {
        bf16x4 av, bv;
#pragma unroll
        for(int e = 0; e < 4; ++e)
        {
            const int k  = k0 + 4 * g + e;
            const int am = m0 + c, bn = n0 + c;
            av[e] = bf16_bits(am < M && k < K ? A[am * sa_m + k * sa_k] : 0.f);
            bv[e] = bf16_bits(bn < N && k < K ? B[k * sb_k + bn * sb_n] : 0.f);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for(int v = 0; v < 4; ++v)
    {
        const int m = m0 + 4 * g + v, n = n0 + c;
        if(m < M && n < N)
        {
            float out = acc[v];
            if(accumulate)
                out += C[(int64_t)m * ldc + n];
            if(bias)
                out += bias[n];
            if(relu)
                out = out > 0.f ? out : 0.f;
            C[(int64_t)m * ldc + n] = out;
        }
    }
}

// Z[i][f] = sum over the CSR row i of w * H[j][f] (F <= 64: one lane per feature)
__global__ __launch_bounds__(256) void ws_agg_kernel(int32_t        n,
                                                     int32_t        F,
                                                     const int32_t* ptr,
                                                     const int32_t* idx,
                                                     const float*   w,
                                                     const float*   H,
                                                     float*         Z)
{
    const int row = (int)((blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / 64);
    const int f   = threadIdx.x % 64;
    if(row >= n || f >= F)
        return;
    float acc = 0.f;
    for(int32_t k = ptr[row]; k < ptr[row + 1]; ++k)
        acc += w[k] * H[(int64_t)idx[k] * F + f];
    Z[(int64_t)row * F + f] = acc;
}

__global__ void ws_relu_bwd_kernel(int64_t n, const float* pre, float* d)
{
    for(int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if(!(pre[i] > 0.f))
            d[i] = 0.f;
}

// out[f] (+)= sum_i X[i][f], one workgroup per feature, fixed order
__global__ __launch_bounds__(256) void ws_colsum_kernel(int32_t n, int32_t F, const float* X, float* out,
                                                        int32_t accumulate)
{
    __shared__ float lds[256 / kWave];
    const int        f   = blockIdx.x;
    float            acc = 0.f;
    for(int32_t i = threadIdx.x; i < n; i += 256)
        acc += X[(int64_t)i * F + f];
    acc = block_sum<256>(acc, lds);
    if(threadIdx.x == 0)
        out[f] = accumulate ? out[f] + acc : acc;
}

__global__ void ws_adam_kernel(int64_t n, float* w, const float* gr, float* m, float* v, float lr, float b1, float b2,
                               float eps, float c1, float c2)
{
    for(int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    {
        const float g = gr[i];
        m[i]          = b1 * m[i] + (1.f - b1) * g;
        v[i]          = b2 * v[i] + (1.f - b2) * g * g;
        w[i] -= lr * (m[i] / c1) / (sqrtf(v[i] / c2) + eps);
    }
}

// y_i = (base_i +) s M_i v_i, 3x3 M_i column-major (double): FP32 v into
// FP64 y, or FP64 v into FP32 y
__global__ void ws_blk3_kernel(int32_t n, const double* Mb, double s, const float* vf, double* yd, const double* base,
                               const double* vd, float* yf)
{
    for(int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        const double* M = Mb + 9 * (int64_t)i;
        double        v0, v1, v2;
        if(vf)
        {
            v0 = vf[3 * (int64_t)i];
            v1 = vf[3 * (int64_t)i + 1];
            v2 = vf[3 * (int64_t)i + 2];
        }
        else
        {
            v0 = vd[3 * (int64_t)i];
            v1 = vd[3 * (int64_t)i + 1];
            v2 = vd[3 * (int64_t)i + 2];
        }
        for(int a = 0; a < 3; ++a)
        {
            const double y = s * (M[a] * v0 + M[a + 3] * v1 + M[a + 6] * v2);
            if(yd)
                yd[3 * (int64_t)i + a] = base ? base[3 * (int64_t)i + a] + y : y;
            else
                yf[3 * (int64_t)i + a] = (float)y;
        }
    }
}

int grid_of(int64_t n)
{
    return std::max(1, std::min(ceil_div(n, 256), device_cu_count() * 8));
}
}  // namespace

void ws_gemm(int32_t M, int32_t N, int32_t K, const float* A, int64_t sa_m, int64_t sa_k, const float* B,
             int64_t sb_k, int64_t sb_n, float* C, int64_t ldc, bool accumulate, const float* bias, bool relu,
             hipStream_t s)
{
    if(M == 0 || N == 0)
        return;
    const int64_t tiles = (int64_t)((M + 15) / 16) * ((N + 15) / 16);
    ws_gemm_kernel<<<(unsigned)tiles, 64, 0, s>>>(M, N, K, A, sa_m, sa_k, B, sb_k, sb_n, C, ldc, accumulate ? 1 : 0,
                                                  bias, relu ? 1 : 0);
    MI_HIP(hipGetLastError());
}

void ws_aggregate(int32_t n, int32_t F, const int32_t* ptr, const int32_t* idx, const float* w, const float* H,
                  float* Z, hipStream_t s)
{
    if(n > 0)
        ws_agg_kernel<<<ceil_div((int64_t)n * 64, 256), 256, 0, s>>>(n, F, ptr, idx, w, H, Z);
    MI_HIP(hipGetLastError());
}

void ws_relu_backward(int64_t n, const float* pre, float* d, hipStream_t s)
{
    if(n > 0)
        ws_relu_bwd_kernel<<<grid_of(n), 256, 0, s>>>(n, pre, d);
    MI_HIP(hipGetLastError());
}

void ws_colsum(int32_t n, int32_t F, const float* X, float* out, bool accumulate, hipStream_t s)
{
    ws_colsum_kernel<<<F, 256, 0, s>>>(n, F, X, out, accumulate ? 1 : 0);
    MI_HIP(hipGetLastError());
}

void ws_adam(int64_t n, float* w, const float* g, float* m, float* v, float lr, float beta1, float beta2, float eps,
             int32_t step, hipStream_t s)
{
    const float c1 = 1.f - std::pow(beta1, (float)step), c2 = 1.f - std::pow(beta2, (float)step);
    if(n > 0)
        ws_adam_kernel<<<grid_of(n), 256, 0, s>>>(n, w, g, m, v, lr, beta1, beta2, eps, c1, c2);
    MI_HIP(hipGetLastError());
}

void ws_block3(int32_t n, const double* blocks, double scale, const float* v_f32, double* y_f64, const double* base,
               const double* v_f64, float* y_f32, hipStream_t s)
{
    if(n > 0)
        ws_blk3_kernel<<<grid_of(n), 256, 0, s>>>(n, blocks, scale, v_f32, y_f64, base, v_f64, y_f32);
    MI_HIP(hipGetLastError());
}
}  // namespace mi
