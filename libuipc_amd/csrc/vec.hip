// Vector reductions and the PCG vector kernels on gfx950.
//
// Replaces LinearSystemContext::dot / norm (cuda_tool/linear_system.h:436-502,
// cuda_tool/linear_reduction.h:14-60) and LinearFusedPCG's per-iteration
// vector kernels (linear_system/linear_fused_pcg.cu): fused_dot (245-279),
// fused_update_xr (504-549), the scalar node (pcg_scalar_body, 309-327) and
// fused_update_p_beta (593-611).
//
// Reductions are deterministic and need neither fences nor atomics: every
// workgroup writes a partial, and every workgroup of the consuming kernel sums
// all partials in index order (an 8 KB read that stays in L2).  An agent-scope
// fence per workgroup, which the last-workgroup-finishes pattern needs, costs
// an L2 write-back on each of the 8 XCDs.
#include <cmath>

#include <cstdlib>

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

int env_per_cu(const char* name, int dflt)
{
    if(const char* e = std::getenv(name))
        return std::max(1, std::min(32, std::atoi(e)));
    return dflt;
}

// the PCG's vector updates: every workgroup first re-sums the previous
// kernel's partials, so fewer, fatter workgroups (MI_PCG_WG_PER_CU per CU;
// the p update on 786,432 values: 6.1 us at 8 per CU, 5.5 at 4 or 2)
int pcg_grid(int64_t n)
{
    static const int per = env_per_cu("MI_PCG_WG_PER_CU", 4);
    return std::max(1, std::min(ceil_div(n, kThreads), device_cu_count() * per));
}

__global__ void fill_kernel(double* p, double v, int64_t n)
{
    for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void fill_i32_kernel(int32_t* p, int32_t v, int64_t n)
{
    for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void copy_kernel(double* dst, const double* src, int64_t n)
{
    for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}
__global__ void axpy_kernel(double a, const double* __restrict__ x, double* y, int64_t n)
{
    for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] += a * x[i];
}
// one thread per scalar: consecutive threads write consecutive entries
__global__ void gather_rows3_kernel(double* dst, const double* __restrict__ src, const int32_t* __restrict__ idx,
                                    int64_t rows)
{
    const int64_t n = 3 * rows;
    for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    {
        const int64_t k = i / 3;
        dst[i]          = src[3 * (int64_t)idx[k] + (i - 3 * k)];
    }
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

__global__ __launch_bounds__(kThreads) void pcg_dot_partials_kernel(const double* r,
                                                                   const double* z,
                                                                   int64_t       n,
                                                                   double*       part,
                                                                   const int32_t* conv)
{
    if(conv && *conv != 0)
        return;
    __shared__ double lds[kThreads / kWave];
    double            s0 = 0.0, s1 = 0.0;
    for(int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
    {
        double ri = r[i];
        s0 += ri * z[i];
        s1 += ri * ri;
    }
    s0 = block_sum<kThreads>(s0, lds);
    s1 = block_sum<kThreads>(s1, lds);
    if(threadIdx.x == 0)
    {
        part[2 * blockIdx.x]     = s0;
        part[2 * blockIdx.x + 1] = s1;
    }
}

__global__ __launch_bounds__(kThreads) void pcg_init_scalars_kernel(PcgScalars* sc, const double* part, int g,
                                                                   int criterion, double tol)
{
    __shared__ double lds[2 * kThreads / kWave];
    double            rz, bb;
    pcg_sum_partials2<kThreads>(part, g, lds, rz, bb);
    if(threadIdx.x == 0)
    {
        sc->rz         = rz;
        sc->rz_pending = rz;
        sc->rz_new     = rz;
        sc->bb         = bb;
        sc->rr         = bb;
        sc->pAp = sc->alpha = sc->beta = 0.0;
        sc->rz_tol     = tol * fabs(rz);
        sc->rr_tol     = tol * tol * bb;
        sc->converged  = 0;
        sc->criterion  = criterion;
        sc->iter       = 0;
        sc->conv_iter  = 0;
    }
}

// fused_update_xr (linear_fused_pcg.cu:504-549): alpha = rz / pAp, x += alpha p,
// r -= alpha Ap.  Workgroup 0 records pAp / alpha and moves rz_pending to rz
// (nothing else reads rz in this kernel).
__device__ __forceinline__ double pcg_alpha(PcgScalars* sc, const double* pap_part, int gp, double* lds)
{
    return pcg_alpha_of<kThreads>(sc, pap_part, gp, lds);
}

__global__ __launch_bounds__(kThreads) void pcg_update_xr_kernel(PcgScalars*   sc,
                                                                const double* pap_part,
                                                                int           gp,
                                                                double*       x,
                                                                const double* p,
                                                                double*       r,
                                                                const double* Ap,
                                                                int64_t       n,
                                                                double*       zero,
                                                                int           nzero)
{
    if(sc->converged)
        return;
    const double rz = sc->rz_pending;
    if(blockIdx.x == 0)
        for(int k = threadIdx.x; k < nzero; k += kThreads)
            zero[k] = 0.0;
    __shared__ double lds[kThreads / kWave];
    // the partials, then the first entry's operands are loaded before
    // alpha's re-sum (their latency overlaps it)
    PartialsAhead<kThreads> ahead;
    ahead.load(pap_part, gp);
    // (unconditional loads from a clamped index: no branch, no early wait)
    const int64_t i0 = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const int64_t ic = i0 < n ? i0 : n - 1;
    const double  xa = x[ic], pa = p[ic], ra = r[ic], aa = Ap[ic];
    const double alpha = pcg_alpha_of<kThreads>(sc, ahead.sum(pap_part, gp, lds), rz);
    if(i0 < n)
    {
        x[i0] = xa + alpha * pa;
        r[i0] = ra - alpha * aa;
    }
    for(int64_t i = i0 + (int64_t)gridDim.x * kThreads; i < n; i += (int64_t)gridDim.x * kThreads)
    {
        x[i] += alpha * p[i];
        r[i] -= alpha * Ap[i];
    }
}

// fused_update_xr + FEMDiagPreconditioner apply + fused_dot partials, one
// thread per 3x3 block row; DoFs outside [blk_off, blk_off + blk_cnt) get z = r
__global__ __launch_bounds__(kThreads) void pcg_update_xr_diag3_kernel(PcgScalars*   sc,
                                                                      const double* pap_part,
                                                                      int           gp,
                                                                      double*       x,
                                                                      const double* p,
                                                                      double*       r,
                                                                      const double* Ap,
                                                                      double*       z,
                                                                      const double* __restrict__ D,
                                                                      int64_t       blk_off,
                                                                      int64_t       blk_cnt,
                                                                      int64_t       nb,
                                                                      double*       part)
{
    if(sc->converged)
        return;
    __shared__ double lds[kThreads / kWave];
    const double      alpha = pcg_alpha(sc, pap_part, gp, lds);
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
    s0 = block_sum<kThreads>(s0, lds);
    s1 = block_sum<kThreads>(s1, lds);
    if(threadIdx.x == 0)
    {
        part[2 * blockIdx.x]     = s0;
        part[2 * blockIdx.x + 1] = s1;
    }
}

// the scalar node (pcg_scalar_body, linear_fused_pcg.cu:309-327) folded into
// fused_update_p_beta (593-611): converged = |rz_new| <= rz_tol (or
// ||r||^2 <= rr_tol), beta = rz_new / rz, p = z + beta p unless converged.
// Every workgroup derives the same values from the partials; workgroup 0
// records them.  Returns whether the solve has converged.
// The scalars it reads are loaded at the kernel's start (PScalarsIn), off the
// chain that follows the partials' re-sum.
struct PScalarsIn
{
    double rz, rz_tol, rr_tol;
    int    criterion;
    __device__ __forceinline__ explicit PScalarsIn(const PcgScalars* sc)
        : rz(sc->rz), rz_tol(sc->rz_tol), rr_tol(sc->rr_tol), criterion(sc->criterion)
    {
    }
};
__device__ __forceinline__ bool pcg_p_scalars(PcgScalars*                 sc,
                                              const PScalarsIn&           in,
                                              const PairsAhead<kThreads>& ahead,
                                              const double*               part,
                                              int                         gd,
                                              double*                     lds,
                                              double&                     beta)
{
    double rz_new, rr;
    ahead.sum(part, gd, lds, rz_new, rr);
    bool conv = in.criterion == 0 ? (fabs(rz_new) <= in.rz_tol) : (rr <= in.rr_tol);
    if(!isfinite(rz_new) || !isfinite(rr))
        conv = false;
    beta = rz_new / in.rz;
    if(blockIdx.x == 0 && threadIdx.x == 0)
    {
        const int it   = sc->iter + 1;
        sc->iter       = it;
        sc->rz_new     = rz_new;
        sc->rr         = rr;
        sc->beta       = beta;
        sc->rz_pending = rz_new;
        if(conv)
        {
            sc->converged = 1;
            sc->conv_iter = it;
        }
    }
    return conv;
}

// A thread's partials, then its first operands are loaded before the
// partials' re-sum, so that their latency overlaps the sum's chain of
// shuffles and barriers.
__global__ __launch_bounds__(kThreads) void pcg_update_p_kernel(PcgScalars*   sc,
                                                               const double* part,
                                                               int           gd,
                                                               double*       p,
                                                               const double* z,
                                                               int64_t       n,
                                                               float4*       p32)
{
    if(sc->converged)
        return;
    const PScalarsIn  in(sc);
    __shared__ double lds[2 * kThreads / kWave];
    const int64_t     i0     = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const int64_t     stride = (int64_t)gridDim.x * kThreads;
    double            beta;
    PairsAhead<kThreads> ahead;
    ahead.load(part, gd);
    if(p32)
    {
        // the FP32 SpMV's copy of p: one padded float4 per block row
        // (unconditional loads from a clamped row: no branch, no early wait)
        const int64_t kc = 3 * i0 < n ? i0 : n / 3 - 1;
        double        zq[3], pq[3];
#pragma unroll
        for(int c = 0; c < 3; ++c)
        {
            zq[c] = z[3 * kc + c];
            pq[c] = p[3 * kc + c];
        }
        if(pcg_p_scalars(sc, in, ahead, part, gd, lds, beta))
            return;
        for(int64_t k = i0; 3 * k < n; k += stride)
        {
            if(k != i0)
            {
#pragma unroll
                for(int c = 0; c < 3; ++c)
                {
                    zq[c] = z[3 * k + c];
                    pq[c] = p[3 * k + c];
                }
            }
            const double q0 = zq[0] + beta * pq[0], q1 = zq[1] + beta * pq[1], q2 = zq[2] + beta * pq[2];
            p[3 * k]     = q0;
            p[3 * k + 1] = q1;
            p[3 * k + 2] = q2;
            p32[k]       = make_float4((float)q0, (float)q1, (float)q2, 0.f);
        }
        return;
    }
    const int64_t i1 = i0 + stride;
    const int64_t c0 = i0 < n ? i0 : n - 1, c1 = i1 < n ? i1 : n - 1;
    const double  za = z[c0], pa = p[c0], zb = z[c1], pb = p[c1];
    if(pcg_p_scalars(sc, in, ahead, part, gd, lds, beta))
        return;
    if(i0 < n)
        p[i0] = za + beta * pa;
    if(i1 < n)
        p[i1] = zb + beta * pb;
    for(int64_t i = i1 + stride; i < n; i += stride)
        p[i] = z[i] + beta * p[i];
}

__global__ void pack_f32x4_kernel(const double* __restrict__ x, float4* x32, int64_t nb)
{
    for(int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nb; k += (int64_t)gridDim.x * blockDim.x)
        x32[k] = make_float4((float)x[3 * k], (float)x[3 * k + 1], (float)x[3 * k + 2], 0.f);
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
void fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t s)
{
    if(n > 0)
        fill_i32_kernel<<<vec_grid(n), kThreads, 0, s>>>(p, v, n);
}
void gather_rows3(double* dst, const double* src, const int32_t* idx, int64_t rows, hipStream_t s)
{
    if(rows > 0)
        gather_rows3_kernel<<<vec_grid(3 * rows), kThreads, 0, s>>>(dst, src, idx, rows);
}
void axpy_f64(double a, const double* x, double* y, int64_t n, hipStream_t s)
{
    if(n > 0)
        axpy_kernel<<<vec_grid(n), kThreads, 0, s>>>(a, x, y, n);
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

void pcg_prepare(SpmvScratch& scr)
{
    if(scr.grid == 0)
    {
        // the r.z / r.r partials: MI_DOT_WG_PER_CU workgroups per CU
        static const int per = env_per_cu("MI_DOT_WG_PER_CU", 4);
        scr.grid = device_cu_count() * per;
        scr.partials.resize((size_t)2 * scr.grid);
    }
}
static void ensure(SpmvScratch& scr)
{
    pcg_prepare(scr);
}

void pcg_init_vectors(double* x, double* r, const double* b, int64_t n, hipStream_t s)
{
    if(n > 0)
        pcg_init_kernel<<<vec_grid(n), kThreads, 0, s>>>(x, r, b, n);
}

int pcg_dot_partials(const double* r, const double* z, int64_t n, double* part, const int32_t* conv,
                     SpmvScratch& scr, hipStream_t s)
{
    ensure(scr);
    MI_REQUIRE(r && z && part, MI_ERR_STATE, "pcg_dot_partials: null buffer");
    int g = std::min(scr.grid, std::max(1, ceil_div(n, kThreads)));
    MI_REQUIRE(part != scr.partials.data() || (size_t)(2 * g) <= scr.partials.size(), MI_ERR_STATE,
               "pcg_dot_partials: partials too small");
    pcg_dot_partials_kernel<<<g, kThreads, 0, s>>>(r, z, n, part, conv);
    return g;
}

void pcg_init_scalars(PcgScalars* sc, const double* part, int g, int criterion, double tol, hipStream_t s)
{
    MI_REQUIRE(sc && part, MI_ERR_STATE, "pcg_init_scalars: null buffer");
    pcg_init_scalars_kernel<<<1, kThreads, 0, s>>>(sc, part, g, criterion, tol);
}

__global__ void pcg_set_bb_kernel(PcgScalars* sc, double bb, double tol)
{
    sc->bb     = bb;
    sc->rr_tol = tol * tol * bb;
}

void pcg_set_bb(PcgScalars* sc, double bb, double tol, hipStream_t s)
{
    pcg_set_bb_kernel<<<1, 1, 0, s>>>(sc, bb, tol);
}

void pcg_update_xr(PcgScalars* sc, const double* pap_part, int gp, double* x, const double* p, double* r,
                   const double* Ap, int64_t n, hipStream_t s, double* zero, int nzero)
{
    MI_REQUIRE(sc && pap_part && x && p && r && Ap, MI_ERR_STATE, "pcg_update_xr: null buffer");
    pcg_update_xr_kernel<<<pcg_grid(n), kThreads, 0, s>>>(sc, pap_part, gp, x, p, r, Ap, n, zero,
                                                          zero ? nzero : 0);
}

int pcg_update_xr_diag3(PcgScalars* sc, const double* pap_part, int gp, double* x, const double* p, double* r,
                        const double* Ap, double* z, const double* D, int64_t blk_off, int64_t blk_cnt, int64_t n,
                        double* part, SpmvScratch& scr, hipStream_t s)
{
    ensure(scr);
    MI_REQUIRE(sc && pap_part && x && p && r && Ap && z && D && part, MI_ERR_STATE, "pcg_update_xr_diag3: null buffer");
    int64_t nb = n / 3;
    int     g  = std::min(scr.grid, std::max(1, ceil_div(nb, kThreads)));
    pcg_update_xr_diag3_kernel<<<g, kThreads, 0, s>>>(sc, pap_part, gp, x, p, r, Ap, z, D, blk_off, blk_cnt, nb,
                                                      part);
    return g;
}

void pcg_update_p(PcgScalars* sc, const double* part, int gd, double* p, const double* z, int64_t n, float4* p32,
                  hipStream_t s)
{
    MI_REQUIRE(sc && part && p && z, MI_ERR_STATE, "pcg_update_p: null buffer");
    pcg_update_p_kernel<<<pcg_grid(n), kThreads, 0, s>>>(sc, part, gd, p, z, n, p32);
}

void pack_f32x4(const double* x, float4* x32, int64_t nb, hipStream_t s)
{
    if(nb > 0)
        pack_f32x4_kernel<<<vec_grid(3 * nb), kThreads, 0, s>>>(x, x32, nb);
}

void pcg_copy_if_active(PcgScalars* sc, double* dst, const double* src, int64_t n, hipStream_t s)
{
    pcg_copy_if_active_kernel<<<pcg_grid(n), kThreads, 0, s>>>(sc, dst, src, n);
}
}  // namespace mi
