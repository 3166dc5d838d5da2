// Vector reductions and the PCG vector kernels on gfx950.
//
// Replaces LinearSystemContext::dot / norm (cuda_tool/linear_system.h:436-502,
// cuda_tool/linear_reduction.h:14-60) and LinearFusedPCG's per-iteration
// vector kernels (linear_system/linear_fused_pcg.cu): fused_dot (245-279) with
// the scalar node folded into its last workgroup (pcg_scalar_body, 309-327),
// fused_update_xr (504-549) and fused_update_p_beta (593-611).
//
// Reductions are deterministic: every workgroup writes a partial, the last
// workgroup to arrive (atomic ticket) sums the partials in index order.
#include <cmath>

#include "linsys.hpp"

namespace mi
{
namespace
{
constexpr int kThreads = 256;

int vec_grid(int64_t n)
{
    return std::max(1, std::min(ceil_div(n, kThreads), device_cu_count() * 8));
}

__global__ void fill_kernel(double* p, double v, int64_t n)
{
    for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void copy_kernel(double* dst, const double* src, int64_t n)
{
    for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// ---------------------------------------------------------------- KAT dot / norm
template <typename T>
__global__ __launch_bounds__(kThreads) void dot_partial_kernel(const T* x, const T* y, int64_t n, double* partials)
{
    __shared__ double lds[kThreads / kWave];
    double            s = 0.0;
    for(int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
        s += (double)x[i] * (double)y[i];
    s = block_sum<kThreads>(s, lds);
    if(threadIdx.x == 0)
        partials[blockIdx.x] = s;
}

// scaled square sum (scale, ssq) as in linear_reduction.h:14-40
struct Sss
{
    double scale, ssq;
};
__device__ __forceinline__ Sss sss_combine(Sss a, Sss b)
{
    if(a.scale < b.scale)
    {
        double r = a.scale / b.scale;
        return {b.scale, b.ssq + a.ssq * r * r};
    }
    if(b.scale < a.scale)
    {
        double r = b.scale / a.scale;
        return {a.scale, a.ssq + b.ssq * r * r};
    }
    if(a.scale == b.scale)
        return {a.scale, a.ssq + b.ssq};
    return {a.scale + b.scale, a.ssq + b.ssq};  // NaN propagation
}

template <typename T>
__global__ __launch_bounds__(kThreads) void norm_partial_kernel(const T* x, int64_t n, double* partials)
{
    __shared__ double ls[kThreads], lq[kThreads];
    Sss               acc{0.0, 0.0};
    for(int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
    {
        double v = fabs((double)x[i]);
        if(v != 0.0 || isnan(v))
            acc = sss_combine(acc, Sss{v, 1.0});
    }
    ls[threadIdx.x] = acc.scale;
    lq[threadIdx.x] = acc.ssq;
    __syncthreads();
    for(int o = kThreads / 2; o > 0; o >>= 1)
    {
        if((int)threadIdx.x < o)
        {
            Sss c           = sss_combine(Sss{ls[threadIdx.x], lq[threadIdx.x]},
                                Sss{ls[threadIdx.x + o], lq[threadIdx.x + o]});
            ls[threadIdx.x] = c.scale;
            lq[threadIdx.x] = c.ssq;
        }
        __syncthreads();
    }
    if(threadIdx.x == 0)
    {
        partials[2 * blockIdx.x]     = ls[0];
        partials[2 * blockIdx.x + 1] = lq[0];
    }
}

// ---------------------------------------------------------------- PCG
__global__ void pcg_init_kernel(double* x, double* r, const double* b, int64_t n)
{
    for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    {
        x[i] = 0.0;
        r[i] = b[i];
    }
}

// last-workgroup finish of a pair of sums (s0, s1) -> device
__device__ __forceinline__ bool finish_pair(double s0, double s1, double* partials, uint32_t* ticket, double* lds, double out[2])
{
    __shared__ int s_last;
    double         b0 = block_sum<kThreads>(s0, lds);
    double         b1 = block_sum<kThreads>(s1, lds);
    if(threadIdx.x == 0)
    {
        partials[2 * blockIdx.x]     = b0;
        partials[2 * blockIdx.x + 1] = b1;
        __threadfence();
        uint32_t t = atomicAdd(ticket, 1u);
        s_last     = (t == gridDim.x - 1u);
    }
    __syncthreads();
    if(!s_last)
        return false;
    __threadfence();
    double a0 = 0.0, a1 = 0.0;
    for(int k = threadIdx.x; k < (int)gridDim.x; k += kThreads)
    {
        a0 += __hip_atomic_load(&partials[2 * k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        a1 += __hip_atomic_load(&partials[2 * k + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    out[0] = block_sum<kThreads>(a0, lds);
    out[1] = block_sum<kThreads>(a1, lds);
    if(threadIdx.x == 0)
        *ticket = 0u;
    return true;
}

// the scalar node of one iteration (pcg_scalar_body, linear_fused_pcg.cu:309-327):
// converged = |rz_new| <= rz_tol (or ||r||^2 <= rr_tol), beta = rz_new / rz,
// rz <- rz_new unless converged
__device__ __forceinline__ void pcg_scalar_update(PcgScalars* sc, double rz_new, double rr)
{
    sc->rz_new = rz_new;
    sc->rr     = rr;
    sc->iter += 1;
    int conv = sc->criterion == 0 ? (fabs(rz_new) <= sc->rz_tol) : (rr <= sc->rr_tol);
    if(!isfinite(rz_new) || !isfinite(rr))
        conv = 0;
    sc->beta = rz_new / sc->rz;
    if(conv)
    {
        sc->converged = 1;
        sc->conv_iter = sc->iter;
    }
    else
        sc->rz = rz_new;
}

__global__ __launch_bounds__(kThreads) void pcg_dot_rz_rr_kernel(const double* r,
                                                                const double* z,
                                                                int64_t       n,
                                                                double*       partials,
                                                                uint32_t*     ticket,
                                                                double*       out3)
{
    __shared__ double lds[kThreads / kWave];
    double            s0 = 0.0, s1 = 0.0;
    for(int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
    {
        double ri = r[i];
        s0 += ri * z[i];
        s1 += ri * ri;
    }
    double out[2];
    if(finish_pair(s0, s1, partials, ticket, lds, out) && threadIdx.x == 0)
    {
        out3[0] = out[0];
        out3[1] = out[1];
    }
}

__global__ __launch_bounds__(kThreads) void pcg_dot_scalar_kernel(PcgScalars*   sc,
                                                                 const double* r,
                                                                 const double* z,
                                                                 int64_t       n,
                                                                 double*       partials,
                                                                 uint32_t*     ticket)
{
    if(sc->converged)
        return;
    __shared__ double lds[kThreads / kWave];
    double            s0 = 0.0, s1 = 0.0;
    for(int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
    {
        double ri = r[i];
        s0 += ri * z[i];
        s1 += ri * ri;
    }
    double out[2];
    if(finish_pair(s0, s1, partials, ticket, lds, out) && threadIdx.x == 0)
        pcg_scalar_update(sc, out[0], out[1]);
}

// fused_update_xr (linear_fused_pcg.cu:504-549): x += alpha p, r -= alpha Ap,
// and Ap is zeroed for the next SpMV's atomics
__global__ void pcg_update_xr_kernel(PcgScalars* sc, double* x, const double* p, double* r, double* Ap, int64_t n)
{
    if(sc->converged)
        return;
    const double alpha = sc->alpha;
    for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    {
        x[i] += alpha * p[i];
        r[i] -= alpha * Ap[i];
        Ap[i] = 0.0;
    }
}

// fused_update_xr + FEMDiagPreconditioner apply + fused_dot + scalar node, one
// thread per 3x3 block row; DoFs outside [blk_off, blk_off + blk_cnt) get z = r
__global__ __launch_bounds__(kThreads) void pcg_update_xr_diag3_kernel(PcgScalars*   sc,
                                                                      double*       x,
                                                                      const double* p,
                                                                      double*       r,
                                                                      double*       Ap,
                                                                      double*       z,
                                                                      const double* __restrict__ D,
                                                                      int64_t       blk_off,
                                                                      int64_t       blk_cnt,
                                                                      int64_t       nb,
                                                                      double*       partials,
                                                                      uint32_t*     ticket)
{
    if(sc->converged)
        return;
    __shared__ double lds[kThreads / kWave];
    const double      alpha = sc->alpha;
    double            s0 = 0.0, s1 = 0.0;
    for(int64_t k = (int64_t)blockIdx.x * kThreads + threadIdx.x; k < nb; k += (int64_t)gridDim.x * kThreads)
    {
        double rv[3];
#pragma unroll
        for(int c = 0; c < 3; ++c)
        {
            int64_t i = 3 * k + c;
            x[i] += alpha * p[i];
            rv[c] = r[i] - alpha * Ap[i];
            r[i]  = rv[c];
            Ap[i] = 0.0;
        }
        double  zv[3];
        int64_t lb = k - blk_off;
        if(lb >= 0 && lb < blk_cnt)
        {
            const double* Di = D + 9 * lb;
#pragma unroll
            for(int a = 0; a < 3; ++a)
                zv[a] = Di[a] * rv[0] + Di[a + 3] * rv[1] + Di[a + 6] * rv[2];
        }
        else
        {
            zv[0] = rv[0];
            zv[1] = rv[1];
            zv[2] = rv[2];
        }
#pragma unroll
        for(int c = 0; c < 3; ++c)
        {
            z[3 * k + c] = zv[c];
            s0 += rv[c] * zv[c];
            s1 += rv[c] * rv[c];
        }
    }
    double out[2];
    if(finish_pair(s0, s1, partials, ticket, lds, out) && threadIdx.x == 0)
        pcg_scalar_update(sc, out[0], out[1]);
}

__global__ void pcg_update_p_kernel(PcgScalars* sc, double* p, const double* z, int64_t n)
{
    if(sc->converged)
        return;
    const double beta = sc->beta;
    for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = z[i] + beta * p[i];
}

__global__ void pcg_copy_if_active_kernel(PcgScalars* sc, double* dst, const double* src, int64_t n)
{
    if(sc->converged)
        return;
    for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

template <typename T>
double host_dot(const T* x, const T* y, int64_t n, hipStream_t s)
{
    if(n <= 0)
        return 0.0;
    int           g = vec_grid(n);
    DBuf<double>  part;
    part.resize((size_t)g);
    dot_partial_kernel<T><<<g, kThreads, 0, s>>>(x, y, n, part.data());
    MI_HIP(hipGetLastError());
    std::vector<double> h((size_t)g);
    MI_HIP(hipMemcpyAsync(h.data(), part.data(), sizeof(double) * g, hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    double acc = 0.0;
    for(double v : h)
        acc += v;
    return acc;
}

template <typename T>
double host_norm(const T* x, int64_t n, hipStream_t s)
{
    if(n <= 0)
        return 0.0;
    int          g = vec_grid(n);
    DBuf<double> part;
    part.resize((size_t)2 * g);
    norm_partial_kernel<T><<<g, kThreads, 0, s>>>(x, n, part.data());
    MI_HIP(hipGetLastError());
    std::vector<double> h((size_t)2 * g);
    MI_HIP(hipMemcpyAsync(h.data(), part.data(), sizeof(double) * 2 * g, hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    double scale = 0.0, ssq = 0.0;
    for(int k = 0; k < g; ++k)
    {
        double bs = h[2 * k], bq = h[2 * k + 1];
        if(scale < bs)
        {
            double r = scale / bs;
            ssq      = bq + ssq * r * r;
            scale    = bs;
        }
        else if(bs < scale)
        {
            double r = bs / scale;
            ssq      = ssq + bq * r * r;
        }
        else if(scale == bs)
            ssq += bq;
        else
        {
            scale += bs;
            ssq += bq;
        }
    }
    if(scale == 0.0)
        return 0.0;
    return scale * std::sqrt(ssq);
}
}  // namespace

void fill_f64(double* p, double v, int64_t n, hipStream_t s)
{
    if(n > 0)
        fill_kernel<<<vec_grid(n), kThreads, 0, s>>>(p, v, n);
}
void copy_f64(double* dst, const double* src, int64_t n, hipStream_t s)
{
    if(n > 0)
        copy_kernel<<<vec_grid(n), kThreads, 0, s>>>(dst, src, n);
}

double dot_f64_host(const double* x, const double* y, int64_t n, hipStream_t s)
{
    return host_dot<double>(x, y, n, s);
}
double norm_f64_host(const double* x, int64_t n, hipStream_t s)
{
    return host_norm<double>(x, n, s);
}
float dot_f32_host(const float* x, const float* y, int64_t n, hipStream_t s)
{
    return (float)host_dot<float>(x, y, n, s);
}
float norm_f32_host(const float* x, int64_t n, hipStream_t s)
{
    return (float)host_norm<float>(x, n, s);
}

static void ensure(SpmvScratch& scr)
{
    if(scr.grid == 0)
    {
        scr.grid = device_cu_count() * 4;
        scr.partials.resize((size_t)2 * scr.grid);
        scr.ticket.resize(4);
        MI_HIP(hipMemset(scr.ticket.data(), 0, 4 * sizeof(uint32_t)));
    }
}

void pcg_init_vectors(double* x, double* r, const double* b, int64_t n, hipStream_t s)
{
    if(n > 0)
        pcg_init_kernel<<<vec_grid(n), kThreads, 0, s>>>(x, r, b, n);
}

void pcg_dot_rz_rr(const double* r, const double* z, int64_t n, double* out3, SpmvScratch& scr, hipStream_t s)
{
    ensure(scr);
    int g = std::min(scr.grid, std::max(1, ceil_div(n, kThreads)));
    pcg_dot_rz_rr_kernel<<<g, kThreads, 0, s>>>(r, z, n, scr.partials.data(), scr.ticket.data(), out3);
}

void pcg_update_xr(PcgScalars* sc, double* x, const double* p, double* r, double* Ap, int64_t n, hipStream_t s)
{
    pcg_update_xr_kernel<<<vec_grid(n), kThreads, 0, s>>>(sc, x, p, r, Ap, n);
}

void pcg_update_xr_diag3(PcgScalars*   sc,
                         double*       x,
                         const double* p,
                         double*       r,
                         double*       Ap,
                         double*       z,
                         const double* D,
                         int64_t       blk_off,
                         int64_t       blk_cnt,
                         int64_t       n,
                         SpmvScratch&  scr,
                         hipStream_t   s)
{
    ensure(scr);
    int64_t nb = n / 3;
    int     g  = std::min(scr.grid, std::max(1, ceil_div(nb, kThreads)));
    pcg_update_xr_diag3_kernel<<<g, kThreads, 0, s>>>(
        sc, x, p, r, Ap, z, D, blk_off, blk_cnt, nb, scr.partials.data(), scr.ticket.data());
}

void pcg_dot_scalar(PcgScalars* sc, const double* r, const double* z, int64_t n, SpmvScratch& scr, hipStream_t s)
{
    ensure(scr);
    int g = std::min(scr.grid, std::max(1, ceil_div(n, kThreads)));
    pcg_dot_scalar_kernel<<<g, kThreads, 0, s>>>(sc, r, z, n, scr.partials.data(), scr.ticket.data());
}

void pcg_update_p(PcgScalars* sc, double* p, const double* z, int64_t n, hipStream_t s)
{
    pcg_update_p_kernel<<<vec_grid(n), kThreads, 0, s>>>(sc, p, z, n);
}

void pcg_copy_if_active(PcgScalars* sc, double* dst, const double* src, int64_t n, hipStream_t s)
{
    pcg_copy_if_active_kernel<<<vec_grid(n), kThreads, 0, s>>>(sc, dst, src, n);
}
}  // namespace mi
