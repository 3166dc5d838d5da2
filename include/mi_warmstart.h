/* Learned initial guess for the PCG: device kernels of a message-passing
 * network over the block graph of an assembled system (inference and
 * training).  No counterpart in the reference, whose solve starts from x0 = 0;
 * a host feeds the prediction to mi_pcg_solve with mi_pcg_opts.x_init = 1.
 * All pointers are device pointers; matrices are row-major FP32 unless noted. */
#ifndef MI_WARMSTART_H
#define MI_WARMSTART_H

#include <stdint.h>

#include "mi_linsys.h"

#ifdef __cplusplus
extern "C" {
#endif

/* C[m][n] = sum_k A(m, k) B(k, n) (+ C[m][n] if accumulate) (+ bias[n]) (relu),
 * A(m, k) = A[m sa_m + k sa_k], B(k, n) = B[k sb_k + n sb_n]: the operands are
 * rounded to BF16 and multiplied on the matrix cores with FP32 accumulation. */
int mi_ws_gemm(mi_ctx* ctx, int32_t M, int32_t N, int32_t K, const float* A, int64_t sa_m, int64_t sa_k,
               const float* B, int64_t sb_k, int64_t sb_n, float* C, int64_t ldc, int32_t accumulate,
               const float* bias, int32_t relu, void* stream);

/* Z[i][f] = sum over k in [ptr[i], ptr[i+1]) of w[k] H[idx[k]][f], 1 <= F <= 64.
 * With symmetric weights this is also its own adjoint. */
int mi_ws_aggregate(mi_ctx* ctx, int32_t n, int32_t F, const int32_t* ptr, const int32_t* idx, const float* w,
                    const float* H, float* Z, void* stream);

/* d[i] = 0 where pre[i] <= 0 (the ReLU's backward, in place) */
int mi_ws_relu_backward(mi_ctx* ctx, int64_t n, const float* pre, float* d, void* stream);

/* out[f] (+)= sum_i X[i][f] (F columns, fixed summation order) */
int mi_ws_colsum(mi_ctx* ctx, int32_t n, int32_t F, const float* X, float* out, int32_t accumulate, void* stream);

/* one Adam step (bias-corrected, step >= 1) on FP32 weights w with gradient g */
int mi_ws_adam(mi_ctx* ctx, int64_t n, float* w, const float* g, float* m, float* v, float lr, float beta1,
               float beta2, float eps, int32_t step, void* stream);

/* y_i = scale * M_i v_i for n column-major 3x3 FP64 blocks: FP32 v into FP64 y
 * (plus base_i when base is given), or FP64 v into FP32 y (base NULL); exactly
 * one (v, y) pair given */
int mi_ws_block3(mi_ctx* ctx, int32_t n, const double* blocks, double scale, const float* v_f32, double* y_f64,
                 const double* base, const double* v_f64, float* y_f32, void* stream);

#ifdef __cplusplus
}
#endif
#endif
