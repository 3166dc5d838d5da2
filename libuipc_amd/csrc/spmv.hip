// Block-symmetric SpMV fused with x^T A x, on gfx950.
//
// Replaces Spmv::rbk_sym_spmv_dot (linear_system/spmv.cu:1023-1159; kernel
// Spmv_rbk_sym_spmv_dot_chunked_kernel<1,0>, 531-756) with the semantics of
// Spmv::cpu_sym_spmv (spmv.cu:1210-1296): y = (U + U^T - diag U) x.
//
// gfx950 design: the reference accumulates A^T x_i of every upper block into
// y_j with FP64 atomics (spmv.cu:683-723).  On MI355X the device-scope atomics
// of eight non-coherent XCD L2s serialise badly, so the assembly also builds
// the full BSR (both triangles, BsrMatrix::full_*; every row holds at least
// one block, an empty row gets a zero diagonal block) and the SpMV is a row
// gather over it, deterministic and free of atomics:
//   * a workgroup takes a contiguous range of chunks of whole rows (chunk c =
//     the rows whose first block lies in [c S, (c+1) S), kChunkStep S) and
//     streams the range's blocks in stages of 768 (three rounds of one block
//     per lane, every stage full but the range's last), with consecutive
//     lanes on consecutive blocks, so every value plane streams as 64
//     contiguous entries per wave instruction (16 lanes per row of ~40 blocks
//     reached ~60% of that rate; stages cut at chunk ends left a third of the
//     rounds' lanes idle);
//   * a stage's products w_k = A_k x_col(k) go to LDS and each row is summed
//     from there, 16 lanes per row in a fixed order; a row running past the
//     stage leaves its partial sum to the next stage; 18 KB of LDS and at
//     most 64 VGPRs keep 8 waves per SIMD resident to hide the load latency;
//   * FP64 planes gather x as doubles; FP32 planes gather the float4 copy of
//     x (the PCG keeps one next to p) and form the products in double;
//   * x^T A x = sum_i x_i . (A x)_i per workgroup into partials[blockIdx.x];
//     the consumer kernel (PCG update) sums the partials in a fixed order;
//   * a persistent grid; each XCD (workgroups are dispatched round-robin over
//     the 8 XCDs) takes a contiguous range of chunks, so x is re-read from its
//     own L2.
#include <cstdlib>
#include <mutex>
#include <type_traits>

#include "linsys.hpp"

namespace mi
{
namespace
{
constexpr int kThreads  = 256;
constexpr int kWaves    = kThreads / kWave;
constexpr int kXcds     = 8;
constexpr int kLanes    = 16;  // lanes per block row in the chunk kernel's row sums
constexpr int kGroups   = kThreads / kLanes;
// blocks staged in LDS at a time: three rounds of one block per lane; 18 KB
// of LDS per workgroup, so that LDS does not limit the occupancy below 8
// waves per SIMD
template <typename VT>
struct Stage
{
    // FP32 planes stage float products: twice the blocks in the same 18 KB
    static constexpr int kPer   = sizeof(VT) == 4 ? 6 : 3;
    static constexpr int kBlocks = kPer * kThreads;
    using W                      = std::conditional_t<sizeof(VT) == 4, float, double>;
    // blocks per lane and round: FP32 planes are read as pairs of consecutive
    // blocks (8-byte loads; the planes' stride is even and the arrays hold
    // one block of padding past the last)
    static constexpr int kVec = sizeof(VT) == 4 ? 2 : 1;
};

template <bool ACCUM>
__device__ __forceinline__ void put_row(double* y, int64_t row, double a, double v0, double v1, double v2)
{
    double* yr = y + 3 * row;
    if constexpr(ACCUM)
    {
        yr[0] += a * v0;
        yr[1] += a * v1;
        yr[2] += a * v2;
    }
    else
    {
        yr[0] = v0;
        yr[1] = v1;
        yr[2] = v2;
    }
}

// x_j as three doubles: FP64 planes read x, FP32 planes the float4 copy
template <typename VT>
__device__ __forceinline__ void gather_x(const double* __restrict__ x,
                                         const float4* __restrict__ x32,
                                         uint32_t j,
                                         double&  x0,
                                         double&  x1,
                                         double&  x2)
{
    if constexpr(sizeof(VT) == 4)
    {
        const float4 t = x32[j];
        x0             = t.x;
        x1             = t.y;
        x2             = t.z;
    }
    else
    {
        x0 = x[3u * j];
        x1 = x[3u * j + 1];
        x2 = x[3u * j + 2];
    }
}

// The matrix stream of the SpMV is read once per product and is larger than
// the caches: non-temporal loads keep it from evicting the vectors (x, and the
// PCG's r, z, p, Ap between kernels).  On the 64^3 + 2M contact system: full
// BSR, FP64 277 -> 260 us per MAS-PCG iteration (its FP32 stream measured no
// better, and keeps the default policy); window layout, FP32 planes 117.5 ->
// 108.9 us per iteration, while FP64 planes loaded with the default policy
// by a third to all of the windows took 151.6 -> 160-181 us.
template <typename VT, typename E, bool NTH = sizeof(VT) == 8>
__device__ __forceinline__ E stream_load(const E* p)
{
    if constexpr(NTH)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

template <typename VT, bool ACCUM>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(8))) void spmv_chunk_kernel(int32_t        nchunks,
                                                             const int32_t* __restrict__ chunk_row,
                                                             const int32_t* __restrict__ fptr,
                                                             const int32_t* __restrict__ fcols,
                                                             const VT* __restrict__ fvals,
                                                             int64_t       fstride,
                                                             const double* __restrict__ x,
                                                             const float4* __restrict__ x32,
                                                             double*       y,
                                                             double        a,
                                                             double*       partials,
                                                             const int32_t* d_conv,
                                                             const int32_t* __restrict__ active,
                                                             const int64_t* d_nactive)
{
    if(d_conv && *d_conv != 0)
        return;
    using W                  = typename Stage<VT>::W;
    constexpr int      kPer   = Stage<VT>::kPer;
    constexpr int      kVec   = Stage<VT>::kVec;
    constexpr int      kStage = Stage<VT>::kBlocks;
    __shared__ W       w[3][kStage];
    __shared__ double  lds[kThreads / kWave];
    __shared__ double  carry[2][3];  // partial sums of the row running past a stage, by stage parity
    __shared__ int32_t nxt[2][2];    // first row of the next stage, and whether it carries a partial

    const int G   = gridDim.x;
    const int b   = blockIdx.x;
    const int per = G / kXcds;
    const int lb  = (per > 0 && b < per * kXcds) ? (b % kXcds) * per + b / kXcds : b;
    // the chunks to do: all of them, or the listed ones (ascending); a
    // workgroup streams each run of consecutive chunks of its share
    const int64_t nitems = active ? *d_nactive : (int64_t)nchunks;
    const int     a0     = (int)(nitems * lb / G);
    const int     a1     = (int)(nitems * (lb + 1) / G);

    const int tid  = threadIdx.x;
    const int g    = tid / kLanes;
    const int lane = tid % kLanes;
    double    dot  = 0.0;
    for(int it = a0; it < a1;)
    {
        int c0 = it, c1 = a1;
        if(active)
        {
            c0 = active[it];
            c1 = c0 + 1;
            for(++it; it < a1 && active[it] == c1; ++it)
                ++c1;
        }
        else
            it = a1;
        // the workgroup's rows [r_first, R1) hold the blocks [s0, K1); they are
        // streamed in stages of kStage blocks, every stage full but the last
        const int32_t R1      = chunk_row[c1];
        const int32_t K1      = fptr[R1];
        int32_t       r_first = chunk_row[c0];
        int32_t       s0      = fptr[r_first];
        bool          carried = false;
        int           par     = 0;
        for(int32_t s1; s0 < K1; s0 = s1, par ^= 1)
        {
            // the stage's slots start at o0 (s0 rounded down to a multiple of
            // kVec, so the first slot may precede s0); it covers [s0, s1)
            const int32_t o0 = s0 & ~(kVec - 1);
            s1               = o0 + kStage < K1 ? o0 + kStage : K1;
            // products, consecutive lanes on consecutive blocks (kVec of them
            // per lane and round: load, gather, product)
#pragma unroll
            for(int h = 0; h < kPer / kVec; ++h)
            {
                const int i = (h * kThreads + tid) * kVec;
                if constexpr(kVec == 2)
                {
                    // k >= s0 - 1, so the second block is at or after s0
                    const int32_t k   = o0 + i;
                    const bool    ok0 = k >= s0 && k < s1, ok1 = k + 1 < s1;
                    const bool    any = ok0 || ok1;
                    int2          jj  = {0, 0};
                    float2        A[9];
                    if(any)
                    {
                        jj = *reinterpret_cast<const int2*>(&fcols[k]);
#pragma unroll
                        for(int q = 0; q < 9; ++q)
                            A[q] = *reinterpret_cast<const float2*>(&fvals[q * fstride + k]);
                    }
                    else
                    {
#pragma unroll
                        for(int q = 0; q < 9; ++q)
                            A[q] = float2{0.f, 0.f};
                    }
                    double x0, x1, x2, z0, z1, z2;
                    gather_x<VT>(x, x32, ok0 ? (uint32_t)jj.x : 0u, x0, x1, x2);
                    gather_x<VT>(x, x32, ok1 ? (uint32_t)jj.y : 0u, z0, z1, z2);
                    if(ok0)
                    {
                        w[0][i] = (W)((double)A[0].x * x0 + (double)A[3].x * x1 + (double)A[6].x * x2);
                        w[1][i] = (W)((double)A[1].x * x0 + (double)A[4].x * x1 + (double)A[7].x * x2);
                        w[2][i] = (W)((double)A[2].x * x0 + (double)A[5].x * x1 + (double)A[8].x * x2);
                    }
                    if(ok1)
                    {
                        w[0][i + 1] = (W)((double)A[0].y * z0 + (double)A[3].y * z1 + (double)A[6].y * z2);
                        w[1][i + 1] = (W)((double)A[1].y * z0 + (double)A[4].y * z1 + (double)A[7].y * z2);
                        w[2][i + 1] = (W)((double)A[2].y * z0 + (double)A[5].y * z1 + (double)A[8].y * z2);
                    }
                }
                else
                {
                    const int32_t  k  = o0 + i;
                    const bool     ok = k < s1;
                    const uint32_t j  = ok ? (uint32_t)stream_load<VT>(&fcols[k]) : 0u;
                    VT             A[9];
#pragma unroll
                    for(int q = 0; q < 9; ++q)
                        A[q] = ok ? stream_load<VT>(&fvals[q * fstride + k]) : VT(0);
                    double x0, x1, x2;
                    gather_x<VT>(x, x32, j, x0, x1, x2);
                    if(ok)
                    {
                        w[0][i] = (W)((double)A[0] * x0 + (double)A[3] * x1 + (double)A[6] * x2);
                        w[1][i] = (W)((double)A[1] * x0 + (double)A[4] * x1 + (double)A[7] * x2);
                        w[2][i] = (W)((double)A[2] * x0 + (double)A[5] * x1 + (double)A[8] * x2);
                    }
                }
            }
            __syncthreads();
            // the rows meeting the stage, 16 lanes each: a row that ends here is
            // finished (plus the partial carried into the stage, for its first
            // row); a row that runs past the stage leaves its partial for the
            // next one.  Every row holds at least one block, so exactly one row
            // either crosses s1 or starts at it: its group names the next stage's
            // first row.
            for(int32_t row = r_first + g; row < R1; row += kGroups)
            {
                const int32_t rs = fptr[row];
                if(rs >= s1)
                {
                    if(rs == s1 && lane == 0)
                    {
                        nxt[par][0] = row;
                        nxt[par][1] = 0;
                    }
                    break;
                }
                const int32_t re = fptr[row + 1];
                const int32_t lo = (rs > s0 ? rs : s0) - o0, hi = (re < s1 ? re : s1) - o0;
                double        acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
                for(int32_t i = lo + lane; i < hi; i += kLanes)
                {
                    acc0 += (double)w[0][i];
                    acc1 += (double)w[1][i];
                    acc2 += (double)w[2][i];
                }
                // the row's 16 lanes (a DPP row) in a fixed order
                acc0 = row16_sum(acc0);
                acc1 = row16_sum(acc1);
                acc2 = row16_sum(acc2);
                if(lane == 0)
                {
                    if(row == r_first && carried)
                    {
                        acc0 = carry[par ^ 1][0] + acc0;
                        acc1 = carry[par ^ 1][1] + acc1;
                        acc2 = carry[par ^ 1][2] + acc2;
                    }
                    if(re <= s1)
                    {
                        put_row<ACCUM>(y, row, a, acc0, acc1, acc2);
                        dot += x[3 * (int64_t)row] * acc0 + x[3 * (int64_t)row + 1] * acc1
                               + x[3 * (int64_t)row + 2] * acc2;
                    }
                    else
                    {
                        carry[par][0] = acc0;
                        carry[par][1] = acc1;
                        carry[par][2] = acc2;
                        nxt[par][0]   = row;
                        nxt[par][1]   = 1;
                    }
                }
            }
            __syncthreads();  // w is rewritten by the next stage
            if(s1 < K1)
            {
                r_first = nxt[par][0];
                carried = nxt[par][1] != 0;
            }
        }
    }
    if(partials)
    {
        double bs = block_sum<kThreads>(dot, lds);
        if(threadIdx.x == 0)
            partials[blockIdx.x] = bs;
    }
}

// ---- the window-layout SpMV (BsrMatrix::w*): one window of KW consecutive
// row positions per work item.  The window's x rows are staged in LDS and its
// y rows accumulate in LDS:
//   * products: consecutive lanes on consecutive blocks; an inner block adds
//     A^T x_i into the LDS row j (LDS FP64 atomics) and stages A x_j; an outer
//     block stages A x_j with x_j gathered from global memory (its word is
//     the row position of j);
//   * row sums: the staged products of each row meeting the stage, 16 lanes
//     per row in a fixed order, added to the row's LDS entry;
//   * the window's rows are written once, with x . (A x).
// With ORD the vectors are in the layout's row order (x, y indexed by row
// position: the window's rows are contiguous); else in the caller's vertex
// order, reached through real (position -> vertex).
template <typename VT, int PER>
struct WinStage
{
    // blocks per lane and stage (FP32 products are staged as floats: twice
    // the blocks in the same LDS)
    static constexpr int kPer = sizeof(VT) == 4 ? 2 * PER : PER;
    using W                   = std::conditional_t<sizeof(VT) == 4, float, double>;
};

// one block of a window stage: stage A x_j in slot i of w; an inner block
// (both rows in the window) also adds A^T x_i into the LDS row j
template <typename VT, typename W, bool ORD, int KW, int STAGE>
__device__ __forceinline__ void win_block(int32_t wd,
                                          const VT (&A)[9],
                                          const double* __restrict__ x,
                                          const int32_t* __restrict__ real,
                                          double (&xs)[3][KW],
                                          double (&ys)[3][KW],
                                          W (&w)[3][STAGE],
                                          int i)
{
    double x0, x1, x2;
    if(wd < 0)
    {
        const int li = (wd >> 15) & 0x7fff, lj = wd & 0x7fff;
        x0 = xs[0][lj];
        x1 = xs[1][lj];
        x2 = xs[2][lj];
        if(li != lj)
        {
            const double u0 = xs[0][li], u1 = xs[1][li], u2 = xs[2][li];
            atomicAdd(&ys[0][lj], (double)A[0] * u0 + (double)A[1] * u1 + (double)A[2] * u2);
            atomicAdd(&ys[1][lj], (double)A[3] * u0 + (double)A[4] * u1 + (double)A[5] * u2);
            atomicAdd(&ys[2][lj], (double)A[6] * u0 + (double)A[7] * u1 + (double)A[8] * u2);
        }
    }
    else
    {
        const int64_t j = ORD ? wd : real[wd];
        x0              = x[3 * j];
        x1              = x[3 * j + 1];
        x2              = x[3 * j + 2];
    }
    w[0][i] = (W)((double)A[0] * x0 + (double)A[3] * x1 + (double)A[6] * x2);
    w[1][i] = (W)((double)A[1] * x0 + (double)A[4] * x1 + (double)A[7] * x2);
    w[2][i] = (W)((double)A[2] * x0 + (double)A[5] * x1 + (double)A[8] * x2);
}

typedef int    int2v __attribute__((ext_vector_type(2)));
typedef int    int4v __attribute__((ext_vector_type(4)));
typedef float  float2v __attribute__((ext_vector_type(2)));
typedef float  float4v __attribute__((ext_vector_type(4)));
typedef double double2v __attribute__((ext_vector_type(2)));
template <typename VT>
struct Vec2;
template <>
struct Vec2<double>
{
    using T = double2v;
};
template <>
struct Vec2<float>
{
    using T = float2v;
};

// VEC blocks per lane and load: with VEC = 2 a lane reads two consecutive
// blocks with 8-byte (FP32) or 16-byte (FP64) loads, with VEC = 4 four FP32
// blocks with 16-byte loads (where the stage holds a multiple of four blocks
// per lane, else two); a stage's slots then start at a multiple of VEC (the
// first slot may precede the stage, the last may pass the window: both are
// masked; the plane stride is a multiple of 4 and the arrays hold four
// blocks of padding past the last)
template <typename VT, bool ACCUM, bool ORD, int NT, int KW, int PER, int VEC>
__global__ __launch_bounds__(NT) void spmv_win_kernel(int32_t nchunks,
                                                      int32_t B,
                                                      const int32_t* __restrict__ wptr,
                                                      const int32_t* __restrict__ words,
                                                      const VT* __restrict__ vals,
                                                      int64_t stride,
                                                      const int32_t* __restrict__ real,
                                                      const double* __restrict__ x,
                                                      double*       y,
                                                      double        a,
                                                      double*       partials,
                                                      const int32_t* d_conv,
                                                      const int32_t* __restrict__ active,
                                                      const int64_t* d_nactive)
{
    static_assert(VEC == 1 || VEC == 2 || VEC == 4, "one, two or four blocks per load");
    static_assert(!(ORD && ACCUM), "row-order products overwrite y");
    if(d_conv && *d_conv != 0)
        return;
    using W                   = typename WinStage<VT, PER>::W;
    constexpr int kPer        = WinStage<VT, PER>::kPer;
    constexpr int kStage      = kPer * NT;
    constexpr int kGroupsW    = NT / kLanes;
    constexpr int kVec        = (VEC == 4 && (kPer % 4 != 0 || sizeof(VT) != 4)) ? 2 : VEC;
    static_assert(kPer % kVec == 0, "whole load groups per lane");
    __shared__ double  ys[3][KW];
    __shared__ double  xs[3][KW];
    __shared__ int32_t rp[KW + 1];
    __shared__ W       w[3][kStage];
    __shared__ double  red[NT / kWave];

    const int G   = gridDim.x;
    const int b   = blockIdx.x;
    const int per = G / kXcds;
    const int lb  = (per > 0 && b < per * kXcds) ? (b % kXcds) * per + b / kXcds : b;
    const int64_t nitems = active ? *d_nactive : (int64_t)nchunks;
    const int     a0     = (int)(nitems * lb / G);
    const int     a1     = (int)(nitems * (lb + 1) / G);

    const int tid  = threadIdx.x;
    const int g    = tid / kLanes;
    const int lane = tid % kLanes;
    double    dot  = 0.0;
    for(int it = a0; it < a1; ++it)
    {
        const int     c  = active ? active[it] : it;
        const int32_t R0 = c * KW;
        const int     nr = (int)std::min<int64_t>(KW, (int64_t)B - R0);
        // the stage's blocks in registers: lane slot j = kVec * h + e holds
        // block o0 + kVec * (h * NT + tid) + e.  The next stage's loads are
        // issued as soon as this stage's products are out, so they are in
        // flight through the barriers and the row sums.
        int32_t wd[kPer];
        VT      A[kPer][9];
        auto    load_stage = [&](int32_t o0, int32_t s0, int32_t s1)
        {
#pragma unroll
            for(int h = 0; h < kPer / kVec; ++h)
            {
                const int32_t k = o0 + kVec * (h * NT + tid);
                if constexpr(kVec == 1)
                {
                    if(k < s1)
                    {
                        wd[h] = stream_load<VT, int32_t, true>(&words[k]);
#pragma unroll
                        for(int q = 0; q < 9; ++q)
                            A[h][q] = stream_load<VT, VT, true>(&vals[q * stride + k]);
                    }
                }
                else if constexpr(kVec == 2)
                {
                    using V2 = typename Vec2<VT>::T;
                    if((k >= s0 && k < s1) || k + 1 < s1)
                    {
                        const int2v w2 = stream_load<VT, int2v, true>(reinterpret_cast<const int2v*>(&words[k]));
                        wd[2 * h]      = w2.x;
                        wd[2 * h + 1]  = w2.y;
#pragma unroll
                        for(int q = 0; q < 9; ++q)
                        {
                            const V2 a2      = stream_load<VT, V2, true>(reinterpret_cast<const V2*>(&vals[q * stride + k]));
                            A[2 * h][q]      = a2.x;
                            A[2 * h + 1][q]  = a2.y;
                        }
                    }
                }
                else
                {
                    // k >= o0 >= s0 - 3: one of the four blocks is in the stage when k < s1
                    if(k < s1)
                    {
                        const int4v w4 = stream_load<VT, int4v, true>(reinterpret_cast<const int4v*>(&words[k]));
                        wd[4 * h]      = w4.x;
                        wd[4 * h + 1]  = w4.y;
                        wd[4 * h + 2]  = w4.z;
                        wd[4 * h + 3]  = w4.w;
#pragma unroll
                        for(int q = 0; q < 9; ++q)
                        {
                            const float4v a4 = stream_load<VT, float4v, true>(reinterpret_cast<const float4v*>(&vals[q * stride + k]));
                            A[4 * h][q]      = a4.x;
                            A[4 * h + 1][q]  = a4.y;
                            A[4 * h + 2][q]  = a4.z;
                            A[4 * h + 3][q]  = a4.w;
                        }
                    }
                }
            }
        };
        auto compute_stage = [&](int32_t o0, int32_t s0, int32_t s1)
        {
#pragma unroll
            for(int j = 0; j < kPer; ++j)
            {
                const int     i = kVec * ((j / kVec) * NT + tid) + j % kVec;
                const int32_t k = o0 + i;
                if(k >= s0 && k < s1)
                    win_block<VT, W, ORD, KW, kStage>(wd[j], A[j], x, real, xs, ys, w, i);
            }
        };
        // the window's blocks [K0, K1), read from wptr: the first stage's
        // loads go out before x is staged, so that their latency overlaps it
        const int32_t K0 = wptr[R0], K1 = wptr[R0 + nr];
        int32_t       s0 = K0;
        int32_t       o0 = s0 & ~(kVec - 1);  // slots [o0, o0 + kStage) cover the stage's blocks [s0, s1)
        int32_t       s1 = o0 + kStage < K1 ? o0 + kStage : K1;
        if(s0 < K1)
            load_stage(o0, s0, s1);
        for(int l = tid; l < nr; l += NT)
        {
            const int64_t v = ORD ? R0 + l : real[R0 + l];
            xs[0][l]        = x[3 * v];
            xs[1][l]        = x[3 * v + 1];
            xs[2][l]        = x[3 * v + 2];
            ys[0][l] = ys[1][l] = ys[2][l] = 0.0;
            rp[l]                          = wptr[R0 + l];
        }
        if(tid == 0)
            rp[nr] = wptr[R0 + nr];
        __syncthreads();
        int row0 = 0;  // the row holding the stage's first block
        while(s0 < K1)
        {
            compute_stage(o0, s0, s1);
            const int32_t n0  = s1;
            const int32_t no0 = n0 & ~(kVec - 1);
            const int32_t n1  = no0 + kStage < K1 ? no0 + kStage : K1;
            if(n0 < K1)
                load_stage(no0, n0, n1);
            __syncthreads();
            // the rows meeting the stage: from the one holding block s0 (the
            // last row starting at or before it).  It is at or after the last
            // stage's (row0): the wave tests the next 64 row starts at once
            // (rp is non-decreasing, so the hits are a prefix), else a binary
            // search past them
            int lo;
            {
                const int      cand = row0 + (int)(tid % kWave);
                const uint64_t hits = __ballot(cand < nr && rp[cand < nr ? cand : nr] <= s0);
                lo                  = row0 + __popcll(hits) - 1;  // bit 0 holds: row0 starts at or before s0
                if(hits == ~0ull && row0 + kWave < nr)
                {
                    int hi = nr - 1;
                    while(lo < hi)
                    {
                        const int mid = (lo + hi + 1) >> 1;
                        if(rp[mid] <= s0)
                            lo = mid;
                        else
                            hi = mid - 1;
                    }
                }
            }
            row0 = lo;
            for(int row = lo + g; row < nr; row += kGroupsW)
            {
                const int32_t rs = rp[row];
                if(rs >= s1)
                    break;
                const int32_t re = rp[row + 1];
                const int32_t b0 = (rs > s0 ? rs : s0) - o0, b1 = (re < s1 ? re : s1) - o0;
                double        acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
                for(int32_t i = b0 + lane; i < b1; i += kLanes)
                {
                    acc0 += (double)w[0][i];
                    acc1 += (double)w[1][i];
                    acc2 += (double)w[2][i];
                }
                // the row's 16 lanes (a DPP row) in a fixed order
                acc0 = row16_sum(acc0);
                acc1 = row16_sum(acc1);
                acc2 = row16_sum(acc2);
                if(lane == 0 && b1 > b0)
                {
                    ys[0][row] += acc0;
                    ys[1][row] += acc1;
                    ys[2][row] += acc2;
                }
            }
            __syncthreads();  // w is rewritten by the next stage
            s0 = n0;
            o0 = no0;
            s1 = n1;
        }
        for(int l = tid; l < nr; l += NT)
        {
            const double v0 = ys[0][l], v1 = ys[1][l], v2 = ys[2][l];
            put_row<ACCUM>(y, ORD ? R0 + l : real[R0 + l], a, v0, v1, v2);
            dot += xs[0][l] * v0 + xs[1][l] * v1 + xs[2][l] * v2;
        }
        __syncthreads();  // the next window restages xs, ys, rp
    }
    if(partials)
    {
        double bs = block_sum<NT>(dot, red);
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

// launch shapes of the window kernel: threads per workgroup, log2 of the
// window rows (the layout is built for one of them), FP64 blocks per lane and
// stage, and blocks per lane and load; MI_SPMV_WIN picks one by index (the
// first is the default).  PCG products on the 64^3 + 2M contact system, us
// FP64 / FP32 planes (threads, rows, blocks per stage, per load):
// 512, 512, 2, 2: 104.3 / 76.0;  512, 512, 2, 1: 109.8 / 75.3;
// 256, 256, 4, 2: 108.5 / 76.5;  256, 256, 2, 2: 110.7 / 79.2;
// 256, 256, 3, 1: 111.9 / 77.7;  256, 256, 2, 1: 112.9 / 79.1;
// 512, 512, 1, 1: 114.9 / 82.4;  256, 128, 2, 1: 124.0 / 79.5;
// 1024, 1024, 1, 1: 141.3 / 88.2.  Small windows at high occupancy beat large
// windows with fewer stored blocks (the LDS window and stage buffer bound the
// workgroups per CU; 512 threads over 1024 rows: 185 us).
struct WinShape
{
    int threads, shift, per, vec;
};
#define MI_WIN_SHAPES(X)                                                                              \
    X(0, 512, 9, 2, 2) X(1, 256, 8, 3, 1) X(2, 512, 9, 2, 1) X(3, 256, 8, 2, 2) X(4, 256, 8, 2, 1) \
    X(5, 512, 9, 1, 1) X(6, 1024, 10, 1, 1) X(7, 256, 7, 2, 1) X(8, 256, 8, 4, 2) X(9, 512, 9, 4, 2) \
    X(10, 1024, 9, 2, 2) X(11, 512, 9, 2, 4)  /* 11: FP32 systems' 512-row default */
#define MI_WIN_ENTRY(i, nt, sh, per, vec) {nt, sh, per, vec},
constexpr WinShape kWinShapes[] = {MI_WIN_SHAPES(MI_WIN_ENTRY)};
#undef MI_WIN_ENTRY
constexpr int kNumWinShapes = (int)(sizeof(kWinShapes) / sizeof(kWinShapes[0]));

// MI_SPMV_WIN overrides the default shape (for sweeps); -1 when unset
int env_shape_index()
{
    static int idx = -2;
    if(idx == -2)
    {
        idx = -1;
        if(const char* e = std::getenv("MI_SPMV_WIN"))
            idx = std::max(0, std::min(kNumWinShapes - 1, std::atoi(e)));
    }
    return idx;
}
int default_shape_index()
{
    return env_shape_index() >= 0 ? env_shape_index() : 0;
}

template <typename VT, bool ACCUM>
const void* win_kernel(int idx)
{
    switch(idx)
    {
#define MI_WIN_FN(i, nt, sh, per, vec) \
    case i: return (const void*)spmv_win_kernel<VT, ACCUM, false, nt, (1 << sh), per, vec>;
        MI_WIN_SHAPES(MI_WIN_FN)
#undef MI_WIN_FN
    }
    return nullptr;
}

template <typename VT, bool ACCUM, bool ORD, int NT, int KW, int PER, int VEC>
void launch_win_t(int g, hipStream_t s, const BsrMatrix& A, const VT* vals, const double* x, double* y, double a,
                  double* partials, const int32_t* d_conv, const int32_t* active, const int64_t* d_nactive)
{
    spmv_win_kernel<VT, ACCUM, ORD, NT, KW, PER, VEC><<<g, NT, 0, s>>>(A.wnchunks, A.B, A.wptr.data(),
                                                                      A.wwords.data(), vals, A.wstride,
                                                                      A.wreal.data(), x, y, a, partials, d_conv,
                                                                      active, d_nactive);
}

// launches the SpMV on the layout the matrix holds (the window layout when
// built, else the full BSR); returns how many partials it wrote from partials[0].
// ORD: x and y in the window layout's row order (window layout only)
template <typename VT, bool ACCUM, bool ORD = false>
int launch_spmv(int g, hipStream_t s, const BsrMatrix& A, const VT* vals, const double* x, const float4* x32,
                double* y, double a, double* partials, const int32_t* d_conv, const int32_t* active = nullptr,
                const int64_t* d_nactive = nullptr)
{
    MI_REQUIRE(!ORD || (A.has_win && !active), MI_ERR_STATE, "spmv: row-order product without a window layout");
    if(A.has_win)
    {
        const int idx = A.wshape;
        MI_REQUIRE(idx >= 0 && idx < kNumWinShapes && A.wshift == kWinShapes[idx].shift, MI_ERR_STATE,
                   "spmv: window layout built for another shape");
        switch(idx)
        {
#define MI_WIN_LAUNCH(i, nt, sh, per, vec)                                                                    \
    case i:                                                                                                      \
        launch_win_t<VT, ACCUM, ORD, nt, (1 << sh), per, vec>(g, s, A, vals, x, y, a, partials, d_conv, active, \
                                                              d_nactive);                                        \
        break;
            MI_WIN_SHAPES(MI_WIN_LAUNCH)
#undef MI_WIN_LAUNCH
        }
        return g;
    }
    spmv_chunk_kernel<VT, ACCUM><<<g, kThreads, 0, s>>>(A.nchunks, A.chunk_row.data(), A.full_ptr.data(),
                                                        A.fcols.data(), vals, A.fstride, x, x32, y, a, partials,
                                                        d_conv, active, d_nactive);
    return g;
}

template <typename VT>
const VT* values_of(const BsrMatrix& A);
template <>
const double* values_of<double>(const BsrMatrix& A)
{
    return A.has_win ? A.wvals.data() : A.fvals.data();
}
template <>
const float* values_of<float>(const BsrMatrix& A)
{
    return A.has_win ? A.wvals32.data() : A.fvals32.data();
}

void check_matrix(const BsrMatrix& A, bool use32)
{
    if(A.has_win)
    {
        MI_REQUIRE(A.wptr.data() && A.wwords.data() && A.wreal.data(), MI_ERR_STATE, "spmv: matrix not assembled");
        MI_REQUIRE(use32 ? A.wvals32.data() != nullptr : A.wvals.data() != nullptr, MI_ERR_STATE,
                   "spmv: values of this precision were not assembled");
        return;
    }
    MI_REQUIRE(A.full_ptr.data() && A.fcols.data() && A.chunk_row.data() && A.nchunks > 0, MI_ERR_STATE,
               "spmv: matrix not assembled");
    MI_REQUIRE(use32 ? A.has32 : (A.fvals.data() != nullptr), MI_ERR_STATE,
               "spmv: values of this precision were not assembled");
}

// workgroups per CU that fit at once: the full-BSR kernel (shape -1) or a
// window shape, per value precision; looked up once each
int occupancy_per_cu(int shape, bool use32)
{
    static int cache[kNumWinShapes + 1][2] = {};
    int&       c = cache[shape + 1][use32 ? 1 : 0];
    if(c == 0)
    {
        const void* fn = shape >= 0 ? (use32 ? win_kernel<float, false>(shape) : win_kernel<double, false>(shape))
                                    : (use32 ? (const void*)spmv_chunk_kernel<float, false>
                                             : (const void*)spmv_chunk_kernel<double, false>);
        const int   nt     = shape >= 0 ? kWinShapes[shape].threads : kThreads;
        int         per_cu = 0;
        MI_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, nt, 0));
        c = std::max(1, per_cu);
    }
    return c;
}

// the persistent grid of a layout: as many workgroups per CU as fit at once
// (MI_SPMV_BLOCKS_PER_CU overrides for sweeps), a multiple of the XCD count
int grid_of_shape(int shape, bool use32, int cu)
{
    static int over = -1;
    if(over < 0)
    {
        over = 0;
        if(const char* e = std::getenv("MI_SPMV_BLOCKS_PER_CU"))
            over = std::max(1, std::min(32, std::atoi(e)));
    }
    const int g = cu * (over > 0 ? over : occupancy_per_cu(shape, use32));
    return std::max(kXcds, (g / kXcds) * kXcds);
}
constexpr int kMaxGridPerCu = 32;  // workgroups per CU of any shape (the partials bound)
}  // namespace

int spmv_shape_count()
{
    return kNumWinShapes;
}

int spmv_default_shape()
{
    return default_shape_index();
}

// the layout of a system of B block rows when none is set: the largest
// window (512, 256 or 128 rows) whose count fills the chip, at least 1.5
// windows per CU (one workgroup streams a window, so fewer windows leave
// CUs idle: 8 windows of 512 rows ran a 4,096-row system on 8 CUs), else the
// full BSR (spmv_shape_full)
int spmv_auto_shape(int64_t B, bool use32)
{
    if(env_shape_index() >= 0)
        return env_shape_index();
    const int64_t need   = (3 * (int64_t)device_cu_count() + 1) / 2;
    // FP32 value planes: the 512-row shape with four blocks per 16-byte load
    const int     cand[] = {use32 ? 11 : default_shape_index(), 3, 7};
    for(int k : cand)
        if((B + (int64_t(1) << kWinShapes[k].shift) - 1) >> kWinShapes[k].shift >= need)
            return k;
    return spmv_shape_full;
}

bool spmv_shape_info(int shape, int* threads, int* rows, int* per, int* vec)
{
    if(shape < 0 || shape >= kNumWinShapes)
        return false;
    if(vec)
        *vec = kWinShapes[shape].vec;
    if(threads)
        *threads = kWinShapes[shape].threads;
    if(rows)
        *rows = 1 << kWinShapes[shape].shift;
    if(per)
        *per = kWinShapes[shape].per;
    return true;
}

int spmv_shape_shift(int shape)
{
    return kWinShapes[shape].shift;
}

int spmv_threads(const BsrMatrix& A)
{
    return A.has_win ? kWinShapes[A.wshape].threads : kThreads;
}

bool spmv_use_window()
{
    static int v = -1;
    if(v < 0)
    {
        const char* e = std::getenv("MI_SPMV_LAYOUT");
        v             = (e && std::string(e) == "full") ? 0 : 1;
    }
    return v == 1;
}

void spmv_prepare(SpmvScratch& scr, bool use32)
{
    if(scr.grid == 0 || scr.use32 != use32)
    {
        scr.use32 = use32;
        scr.grid  = device_cu_count();  // marks the scratch prepared; the grids depend on the layout
        scr.partials.resize((size_t)2 * device_cu_count() * kMaxGridPerCu);
    }
}

int spmv_grid(const SpmvScratch& scr, const BsrMatrix& A)
{
    const int cu = device_cu_count();
    if(A.has_win)
        return (int)std::max<int64_t>(1, std::min<int64_t>(grid_of_shape(A.wshape, scr.use32, cu), A.wnchunks));
    return (int)std::max<int64_t>(1, std::min<int64_t>(grid_of_shape(-1, scr.use32, cu), A.nchunks));
}

// the FP32 kernel gathers x from a float4 copy: the caller's (the PCG keeps
// one next to p), else packed here into the scratch (not under capture)
const float4* x32_of(const BsrMatrix& A, const double* x, const float4* x32, SpmvScratch& scr, hipStream_t s)
{
    if(x32)
        return x32;
    scr.x32.resize((size_t)std::max(1, A.B));
    pack_f32x4(x, scr.x32.data(), A.B, s);
    return scr.x32.data();
}

int spmv_full(const BsrMatrix& A, bool use32, const double* x, const float4* x32, double* y, double* partials,
              const int32_t* d_conv, SpmvScratch& scr, hipStream_t s, const int32_t* active,
              const int64_t* d_nactive, bool ordered)
{
    spmv_prepare(scr, use32);
    if(A.B == 0)
        return 0;
    check_matrix(A, use32);
    const int g = spmv_grid(scr, A);
    MI_REQUIRE(x && y && partials, MI_ERR_STATE, "spmv_full: null buffer");
    MI_REQUIRE(partials >= scr.partials.data() && partials + g <= scr.partials.data() + scr.partials.size(),
               MI_ERR_STATE, "spmv_full: partials outside the scratch");
    int np;
    if(ordered)
        np = use32 ? launch_spmv<float, false, true>(g, s, A, values_of<float>(A), x, nullptr, y, 1.0, partials,
                                                     d_conv, active, d_nactive)
                   : launch_spmv<double, false, true>(g, s, A, values_of<double>(A), x, nullptr, y, 1.0, partials,
                                                      d_conv, active, d_nactive);
    else
        np = use32 ? launch_spmv<float, false>(g, s, A, values_of<float>(A), x,
                                               A.has_win ? nullptr : x32_of(A, x, x32, scr, s), y, 1.0, partials,
                                               d_conv, active, d_nactive)
                   : launch_spmv<double, false>(g, s, A, values_of<double>(A), x, nullptr, y, 1.0, partials, d_conv,
                                                active, d_nactive);
    MI_HIP(hipGetLastError());
    return np;
}

void sum_partials(const double* partials, int n, double* out, hipStream_t s)
{
    sum_partials_kernel<<<1, kThreads, 0, s>>>(partials, n, out);
    MI_HIP(hipGetLastError());
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
    check_matrix(A, use32);
    const int g = spmv_grid(scr, A);
    double*   p = d_dot ? scr.partials.data() : nullptr;
    const int np = use32 ? launch_spmv<float, true>(g, s, A, values_of<float>(A), x,
                                                    A.has_win ? nullptr : x32_of(A, x, nullptr, scr, s), y, a, p,
                                                    nullptr)
                         : launch_spmv<double, true>(g, s, A, values_of<double>(A), x, nullptr, y, a, p, nullptr);
    if(d_dot)
        sum_partials_kernel<<<1, kThreads, 0, s>>>(scr.partials.data(), np, d_dot);
    MI_HIP(hipGetLastError());
}

void spmv_axpby(const BsrMatrix& A, bool use32, double a, const double* x, double b, double* y, SpmvScratch& scr,
                hipStream_t s)
{
    int64_t n = 3 * (int64_t)A.B;
    if(a == 1.0 && b == 0.0 && A.B > 0)
    {
        // y = A x: the PCG's product (one launch, y overwritten)
        spmv_prepare(scr, use32);
        check_matrix(A, use32);
        const int g = spmv_grid(scr, A);
        if(use32)
            launch_spmv<float, false>(g, s, A, values_of<float>(A), x,
                                      A.has_win ? nullptr : x32_of(A, x, nullptr, scr, s), y, 1.0, nullptr, nullptr);
        else
            launch_spmv<double, false>(g, s, A, values_of<double>(A), x, nullptr, y, 1.0, nullptr, nullptr);
        MI_HIP(hipGetLastError());
        return;
    }
    if(n > 0 && b != 1.0)
        scale_kernel<<<std::max(1, std::min(ceil_div(n, 256), device_cu_count() * 8)), 256, 0, s>>>(y, b, n);
    if(a != 0.0)
        spmv_accumulate_dot(A, use32, a, x, y, nullptr, scr, s);
}

// ---- y += A x and x^T A x without allocating: the workgroup partials (and,
// for FP32 value planes, the float4 copy of x) live in a fixed device array.
// Every call on one GPU shares that GPU's copy, so calls that overlap on one
// GPU race; nothing in it has to start at zero (each launch writes every
// partial it later sums, and the whole copy of x).
namespace
{
constexpr size_t kNoAllocParts = 8192;         // workgroup partials (doubles): 64 KiB
constexpr size_t kNoAllocX32   = size_t(1) << 18;  // float4 rows of x for FP32 planes: 4 MiB
constexpr size_t kNoAllocBytes = kNoAllocParts * sizeof(double) + kNoAllocX32 * sizeof(float4);
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kNoAllocDevices = 64;
std::mutex    g_noalloc_mu;
void*         g_noalloc_ptr[kNoAllocDevices]     = {};
int           g_noalloc_grid[kNoAllocDevices][kNumWinShapes + 1][2] = {};  // [device][1 + shape (0: full BSR)][FP32]
}  // namespace

void spmv_noalloc_init(int device)
{
    MI_REQUIRE(device >= 0 && device < kNoAllocDevices, MI_ERR_ARG, "spmv_noalloc_init: device id out of range");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if(g_noalloc_ptr[device])
        return;
    int cur = -1;
    MI_HIP(hipGetDevice(&cur));
    MI_REQUIRE(cur == device, MI_ERR_STATE, "spmv_noalloc_init: the device is not current");
    // the array's name is a host-side stand-in: its device address comes from
    // the runtime, per device
    void* p = nullptr;
    MI_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    MI_REQUIRE(p != nullptr, MI_ERR_HIP, "spmv_noalloc_init: no device address for the scratch array");
    // the launch shapes of spmv_prepare (per layout and value precision),
    // without its buffer
    int cu = 0;
    MI_HIP(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device));
    for(int shape = -1; shape < kNumWinShapes; ++shape)
        for(int k = 0; k < 2; ++k)
            g_noalloc_grid[device][shape + 1][k] =
                (int)std::min<size_t>((size_t)grid_of_shape(shape, k == 1, std::max(1, cu)), kNoAllocParts);
    g_noalloc_ptr[device] = p;
}

void spmv_accumulate_dot_noalloc(const BsrMatrix& A, bool use32, const double* x, double* y, double* d_dot, int device,
                                 hipStream_t s)
{
    MI_REQUIRE(device >= 0 && device < kNoAllocDevices && g_noalloc_ptr[device], MI_ERR_STATE,
               "spmv no-alloc: scratch not set up on this device");
    MI_REQUIRE(!use32 || A.has_win || (size_t)A.B <= kNoAllocX32, MI_ERR_ARG,
               "spmv no-alloc: an FP32 system of more than 262144 block rows exceeds the fixed scratch");
    if(A.B == 0)
    {
        if(d_dot)
            fill_f64(d_dot, 0.0, 1, s);
        return;
    }
    check_matrix(A, use32);
    auto*     base     = static_cast<unsigned char*>(g_noalloc_ptr[device]);
    double*   partials = reinterpret_cast<double*>(base);
    float4*   x32      = reinterpret_cast<float4*>(base + kNoAllocParts * sizeof(double));
    const int g = (int)std::max<int64_t>(
        1, std::min<int64_t>(g_noalloc_grid[device][A.has_win ? A.wshape + 1 : 0][use32 ? 1 : 0],
                             A.has_win ? A.wnchunks : A.nchunks));
    double* p = d_dot ? partials : nullptr;
    int     np;
    if(use32)
    {
        if(!A.has_win)
            pack_f32x4(x, x32, A.B, s);
        np = launch_spmv<float, true>(g, s, A, values_of<float>(A), x, A.has_win ? nullptr : x32, y, 1.0, p, nullptr);
    }
    else
        np = launch_spmv<double, true>(g, s, A, values_of<double>(A), x, nullptr, y, 1.0, p, nullptr);
    if(d_dot)
        sum_partials_kernel<<<1, kThreads, 0, s>>>(partials, np, d_dot);
    MI_HIP(hipGetLastError());
}
}  // namespace mi
