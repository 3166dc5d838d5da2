// Batched PCG kernels (gfx950): many independent systems on consecutive block
// rows of one block-diagonal mi_sys, solved together with per-system scalars.
//
// FusedPCG (linear_system/linear_fused_pcg.cu:1374-2004) runs once per system
// in the reference; here one captured iteration serves every system:
//   the SpMV runs over the union (no block couples two systems);
//   every dot is a per-system sum: the DoFs are cut into tiles of at most
//     kTile DoFs that never cross a system boundary (a host-built table), one
//     workgroup writes each tile's partial, and one wave per system sums its
//     tiles in order (deterministic, no atomics);
//   alpha, beta, the convergence test and the iteration cap are per system;
//     a system that has converged or reached its cap stops changing;
//   the x / r and p updates run tile by tile, each tile reading its system's
//     scalars.
#include "linsys.hpp"

namespace mi
{
namespace
{
constexpr int kThreads = 256;

// per-system sum of tile partials part[stride * t], tiles [t0, t1), in order
__device__ __forceinline__ double wave_tiles(const double* part, int stride, int32_t t0, int32_t t1)
{
    double acc = 0.0;
    for(int32_t t = t0 + (int32_t)__lane_id(); t < t1; t += kWave)
        acc += part[stride * (int64_t)t];
    return wave_sum(acc);
}

// tile partials of a . b (and of a . a when aa != nullptr)
__global__ __launch_bounds__(kThreads) void tile_dot_kernel(const BatchTiles tiles,
                                                           const double* __restrict__ a,
                                                           const double* __restrict__ b,
                                                           double* ab,
                                                           double* aa)
{
    __shared__ double lds[kThreads / kWave];
    for(int32_t t = blockIdx.x; t < tiles.n; t += gridDim.x)
    {
        const int64_t lo = tiles.lo[t], hi = tiles.lo[t + 1];
        double        s0 = 0.0, s1 = 0.0;
        for(int64_t i = lo + threadIdx.x; i < hi; i += kThreads)
        {
            const double ai = a[i];
            s0 += ai * b[i];
            s1 += ai * ai;
        }
        s0 = block_sum<kThreads>(s0, lds);
        if(aa)
            s1 = block_sum<kThreads>(s1, lds);
        if(threadIdx.x == 0)
        {
            ab[t] = s0;
            if(aa)
                aa[t] = s1;
        }
    }
}

// rz0, bb and the tolerances of every system
__global__ void batch_init_kernel(const BatchTiles tiles, BatchScalars* sc, int32_t nseg, const double* rz,
                                  const double* rr, int criterion, double tol, double max_iter_ratio,
                                  const int32_t* bypass)
{
    const int32_t s = (int32_t)((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
    if(s >= nseg)
        return;
    const double z0 = wave_tiles(rz, 1, tiles.sys_tile[s], tiles.sys_tile[s + 1]);
    const double b2 = wave_tiles(rr, 1, tiles.sys_tile[s], tiles.sys_tile[s + 1]);
    if(__lane_id() == 0)
    {
        BatchScalars& q = sc[s];
        q.rz = q.rz_pending = q.rz_new = z0;
        q.bb = q.rr = b2;
        q.pAp = q.alpha = q.beta = 0.0;
        q.rz_tol                 = tol * fabs(z0);
        q.rr_tol                 = tol * tol * b2;
        q.criterion              = criterion;
        q.iter                   = 0;
        q.conv_iter              = 0;
        const int64_t n_s        = tiles.sys_lo[s + 1] - tiles.sys_lo[s];
        q.max_iter               = (int64_t)(max_iter_ratio * (double)n_s);
        q.bypass                 = bypass ? bypass[s] : 0;
        // nothing to do (r0.z0 = 0), or no iteration allowed: FusedPCG runs at
        // most max_iter - 1 iterations (linear_fused_pcg.cu:1994).  A
        // non-finite r0.z0 stops the system too; the host then bypasses its
        // preconditioner or reports the error.
        q.converged = (z0 == 0.0) ? 1 : 0;
        q.stopped   = (z0 == 0.0 || !isfinite(z0) || q.max_iter < 2) ? 1 : 0;
    }
}

// z = r on the tiles of the systems flagged in `flags` (the solve's start), or
// else of the running systems whose scalars carry the bypass flag
__global__ __launch_bounds__(kThreads) void batch_bypass_z_kernel(const BatchTiles tiles,
                                                                 const BatchScalars* __restrict__ sc,
                                                                 const int32_t* __restrict__ flags,
                                                                 const double* __restrict__ r,
                                                                 double* z)
{
    for(int32_t t = blockIdx.x; t < tiles.n; t += gridDim.x)
    {
        const int32_t g = tiles.sys[t];
        if(flags ? !flags[g] : (!sc[g].bypass || sc[g].stopped))
            continue;
        const int64_t lo = tiles.lo[t], hi = tiles.lo[t + 1];
        for(int64_t i = lo + threadIdx.x; i < hi; i += kThreads)
            z[i] = r[i];
    }
}

// alpha = rz / pAp per system (fused_update_xr's scalar)
__global__ void batch_alpha_kernel(const BatchTiles tiles, BatchScalars* sc, int32_t nseg, const double* pap)
{
    const int32_t s = (int32_t)((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
    if(s >= nseg)
        return;
    const double v = wave_tiles(pap, 1, tiles.sys_tile[s], tiles.sys_tile[s + 1]);
    if(__lane_id() == 0)
    {
        BatchScalars& q = sc[s];
        if(q.stopped)
        {
            q.alpha = 0.0;
            return;
        }
        q.pAp   = v;
        q.rz    = q.rz_pending;
        q.alpha = q.rz / v;
    }
}

__global__ __launch_bounds__(kThreads) void batch_update_xr_kernel(const BatchTiles tiles,
                                                                  const BatchScalars* __restrict__ sc,
                                                                  double*       x,
                                                                  const double* __restrict__ p,
                                                                  double*       r,
                                                                  const double* __restrict__ Ap)
{
    for(int32_t t = blockIdx.x; t < tiles.n; t += gridDim.x)
    {
        const BatchScalars& q = sc[tiles.sys[t]];
        if(q.stopped)
            continue;
        const double  alpha = q.alpha;
        const int64_t lo = tiles.lo[t], hi = tiles.lo[t + 1];
        for(int64_t i = lo + threadIdx.x; i < hi; i += kThreads)
        {
            x[i] += alpha * p[i];
            r[i] -= alpha * Ap[i];
        }
    }
}

// convergence, beta and the cap per system (the scalar node of fused_pcg)
__global__ void batch_beta_kernel(const BatchTiles tiles, BatchScalars* sc, int32_t nseg, const double* rz,
                                  const double* rr)
{
    const int32_t s = (int32_t)((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
    if(s >= nseg)
        return;
    const double z = wave_tiles(rz, 1, tiles.sys_tile[s], tiles.sys_tile[s + 1]);
    const double e = wave_tiles(rr, 1, tiles.sys_tile[s], tiles.sys_tile[s + 1]);
    if(__lane_id() == 0)
    {
        BatchScalars& q = sc[s];
        if(q.stopped)
            return;
        bool conv = q.criterion == 0 ? (fabs(z) <= q.rz_tol) : (e <= q.rr_tol);
        if(!isfinite(z) || !isfinite(e))
            conv = false;
        const int64_t it = q.iter + 1;
        q.iter           = it;
        q.rz_new         = z;
        q.rr             = e;
        q.beta           = z / q.rz;
        q.rz_pending     = z;
        if(conv)
        {
            q.converged = 1;
            q.conv_iter = it;
            q.stopped   = 1;
        }
        else if(it >= q.max_iter - 1 || !isfinite(z))
            q.stopped = 1;  // the cap of max_iter - 1 iterations (linear_fused_pcg.cu:1994)
    }
}

__global__ __launch_bounds__(kThreads) void batch_update_p_kernel(const BatchTiles tiles,
                                                                 const BatchScalars* __restrict__ sc,
                                                                 double*       p,
                                                                 const double* __restrict__ z,
                                                                 float4*       p32)
{
    for(int32_t t = blockIdx.x; t < tiles.n; t += gridDim.x)
    {
        const BatchScalars& q = sc[tiles.sys[t]];
        if(q.stopped)
            continue;
        const double  beta = q.beta;
        const int64_t lo = tiles.lo[t], hi = tiles.lo[t + 1];
        if(p32)
        {
            // tiles hold whole block rows (lo and hi are multiples of 3)
            for(int64_t k = lo / 3 + threadIdx.x; k < hi / 3; k += kThreads)
            {
                const double q0 = z[3 * k] + beta * p[3 * k], q1 = z[3 * k + 1] + beta * p[3 * k + 1],
                             q2 = z[3 * k + 2] + beta * p[3 * k + 2];
                p[3 * k]        = q0;
                p[3 * k + 1]    = q1;
                p[3 * k + 2]    = q2;
                p32[k]          = make_float4((float)q0, (float)q1, (float)q2, 0.f);
            }
        }
        else
            for(int64_t i = lo + threadIdx.x; i < hi; i += kThreads)
                p[i] = z[i] + beta * p[i];
    }
}

// the system holding `row`: the last g with seg[g] <= row
__device__ __forceinline__ int32_t system_of(const int32_t* seg, int32_t nseg, int32_t row)
{
    int32_t lo = 0, hi = nseg - 1;
    while(lo < hi)
    {
        const int32_t mid = (lo + hi + 1) / 2;
        if(seg[mid] <= row)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// a SpMV chunk is still needed while one of the systems whose rows lie in
// its row range [lo[c], hi[c]) runs (the range may also hold rows of other
// chunks: a chunk of stopped systems is then computed again, which leaves
// its rows unchanged)
__device__ __forceinline__ bool chunk_needed(const int32_t* lo, const int32_t* hi, int32_t c, const int32_t* seg,
                                             int32_t nseg, const BatchScalars* sc)
{
    const int32_t r0 = lo[c], r1 = hi[c];
    if(r1 <= r0)
        return false;
    const int32_t g1 = system_of(seg, nseg, r1 - 1);
    for(int32_t g = system_of(seg, nseg, r0); g <= g1; ++g)
        if(!sc[g].stopped)
            return true;
    return false;
}

__global__ void chunk_flag_kernel(const int32_t* lo, const int32_t* hi, int32_t nchunks, const int32_t* seg,
                                  int32_t nseg, const BatchScalars* sc, uint32_t* flags)
{
    for(int32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nchunks; c += gridDim.x * blockDim.x)
        flags[c] = chunk_needed(lo, hi, c, seg, nseg, sc) ? 1u : 0u;
}

__global__ void chunk_list_kernel(const int32_t* lo, const int32_t* hi, int32_t nchunks, const int32_t* seg,
                                  int32_t nseg, const BatchScalars* sc, const uint32_t* offs, int32_t* list)
{
    for(int32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nchunks; c += gridDim.x * blockDim.x)
        if(chunk_needed(lo, hi, c, seg, nseg, sc))
            list[offs[c]] = c;
}

__global__ void cluster_flag_kernel(const int32_t* part_to_real, int64_t clusters, int bank, int32_t row_off,
                                    const int32_t* seg, int32_t nseg, const BatchScalars* sc, int32_t* on)
{
    for(int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < clusters;
        k += (int64_t)gridDim.x * blockDim.x)
    {
        int32_t f = 0;
        for(int i = 0; i < bank && !f; ++i)
        {
            const int32_t v = part_to_real[k * bank + i];
            if(v >= 0 && !sc[system_of(seg, nseg, row_off + v)].stopped)
                f = 1;
        }
        on[k] = f;
    }
}

int tile_grid(const BatchTiles& t)
{
    return std::max(1, std::min(t.n, device_cu_count() * 8));
}
int seg_grid(int32_t nseg)
{
    return std::max(1, ceil_div((int64_t)nseg * kWave, kThreads));
}
}  // namespace

void batch_active_chunks(const int32_t* lo, const int32_t* hi, int32_t nchunks, const int32_t* seg, int32_t nseg,
                         const BatchScalars* sc, DBuf<uint32_t>& flags, DBuf<uint32_t>& scan_chunk, int32_t* list,
                         int64_t* d_count, hipStream_t s)
{
    flags.resize((size_t)std::max(1, nchunks));
    const int g = std::max(1, std::min(ceil_div((int64_t)nchunks, kThreads), device_cu_count() * 4));
    chunk_flag_kernel<<<g, kThreads, 0, s>>>(lo, hi, nchunks, seg, nseg, sc, flags.data());
    exclusive_scan_u32(flags.data(), nchunks, scan_chunk, d_count, s);
    chunk_list_kernel<<<g, kThreads, 0, s>>>(lo, hi, nchunks, seg, nseg, sc, flags.data(), list);
    MI_HIP(hipGetLastError());
}

void batch_cluster_flags(const int32_t* part_to_real, int64_t clusters, int bank, int32_t row_off, const int32_t* seg,
                         int32_t nseg, const BatchScalars* sc, int32_t* on, hipStream_t s)
{
    if(clusters > 0)
    {
        const int g = (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(clusters, kThreads), device_cu_count() * 4));
        cluster_flag_kernel<<<g, kThreads, 0, s>>>(part_to_real, clusters, bank, row_off, seg, nseg, sc, on);
    }
    MI_HIP(hipGetLastError());
}

void batch_tile_dot(const BatchTiles& t, const double* a, const double* b, double* ab, double* aa, hipStream_t s)
{
    tile_dot_kernel<<<tile_grid(t), kThreads, 0, s>>>(t, a, b, ab, aa);
    MI_HIP(hipGetLastError());
}

void batch_init(const BatchTiles& t, BatchScalars* sc, int32_t nseg, const double* rz, const double* rr,
                int criterion, double tol, double max_iter_ratio, const int32_t* bypass, hipStream_t s)
{
    batch_init_kernel<<<seg_grid(nseg), kThreads, 0, s>>>(t, sc, nseg, rz, rr, criterion, tol, max_iter_ratio,
                                                          bypass);
    MI_HIP(hipGetLastError());
}

void batch_bypass_z(const BatchTiles& t, const BatchScalars* sc, const int32_t* flags, const double* r, double* z,
                    hipStream_t s)
{
    batch_bypass_z_kernel<<<tile_grid(t), kThreads, 0, s>>>(t, sc, flags, r, z);
    MI_HIP(hipGetLastError());
}

void batch_alpha_xr(const BatchTiles& t, BatchScalars* sc, int32_t nseg, const double* pap, double* x,
                    const double* p, double* r, const double* Ap, hipStream_t s)
{
    batch_alpha_kernel<<<seg_grid(nseg), kThreads, 0, s>>>(t, sc, nseg, pap);
    batch_update_xr_kernel<<<tile_grid(t), kThreads, 0, s>>>(t, sc, x, p, r, Ap);
    MI_HIP(hipGetLastError());
}

void batch_beta_p(const BatchTiles& t, BatchScalars* sc, int32_t nseg, const double* rz, const double* rr,
                  double* p, const double* z, float4* p32, hipStream_t s)
{
    batch_beta_kernel<<<seg_grid(nseg), kThreads, 0, s>>>(t, sc, nseg, rz, rr);
    batch_update_p_kernel<<<tile_grid(t), kThreads, 0, s>>>(t, sc, p, z, p32);
    MI_HIP(hipGetLastError());
}
}  // namespace mi
