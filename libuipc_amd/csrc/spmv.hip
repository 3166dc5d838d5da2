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
//   * work comes in chunks of whole rows (the rows whose first block lies in
//     [c S, (c+1) S), kChunkStep S), read with consecutive lanes on
//     consecutive blocks, so every value plane streams as 64 contiguous
//     entries per wave instruction (16 lanes per row of ~40 blocks reached
//     ~60% of that rate);
//   * FP64 planes (spmv_chunk_kernel): a workgroup stages a chunk's products
//     w_k = A_k x_col(k) in LDS and sums each row from there, 16 lanes per row;
//   * FP32 planes (spmv_seg2_kernel): a wave walks a chunk in windows of 128
//     blocks (two per lane), finds each lane's rows by a shuffle binary search
//     over the row starts, sums rows with a segmented shuffle scan (carrying a
//     row across windows) and loads the next window while this one's x
//     gathers are outstanding;
//   * x^T A x = sum_i x_i . (A x)_i per workgroup into partials[blockIdx.x];
//     the consumer kernel (PCG update) sums the partials in a fixed order;
//   * a persistent grid; each XCD (workgroups are dispatched round-robin over
//     the 8 XCDs) takes a contiguous range of chunks, so x is re-read from its
//     own L2.
#include <cstdlib>

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
constexpr int kChunkCap = 2 * kChunkStep;  // blocks a chunk may hold to be staged in LDS
constexpr int kPer      = kChunkCap / kThreads;
static_assert(kChunkCap % (2 * kThreads) == 0, "chunk capacity must be a multiple of two workgroups");

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

// FP32 planes, two blocks per lane: a wave walks a chunk in windows of 128
// blocks, lane l on blocks k + 2l and k + 2l + 1 (8-byte loads: 512 bytes
// per plane per wave instruction).  Windows start at an even block; a chunk
// that starts at an odd block masks the block before it.  A lane's blocks are
// first combined in the lane (the sum of its last segment), then the lanes
// run the segmented scan; the segment that ends at a lane's first block (the
// second one starts a row) is the previous lane's inclusive value plus that
// block.
typedef int   v2i __attribute__((ext_vector_type(2)));
typedef float v2f __attribute__((ext_vector_type(2)));

struct Window2
{
    v2i j;
    v2f A[9];
};

__device__ __forceinline__ void load_window2(Window2& w,
                                             int32_t  kw,
                                             int32_t  kEnd,
                                             int      lane,
                                             const int32_t* __restrict__ fcols,
                                             const float* __restrict__ fvals,
                                             int64_t fstride)
{
    const int32_t kb = kw + 2 * lane;
    const bool    ok = kb < kEnd;
    w.j              = ok ? *reinterpret_cast<const v2i*>(fcols + kb) : v2i{0, 0};
#pragma unroll
    for(int q = 0; q < 9; ++q)
        w.A[q] = ok ? *reinterpret_cast<const v2f*>(fvals + q * fstride + kb) : v2f{0.f, 0.f};
}

// row start i of the 128 that can meet a window: i < 64 in fa, else in fb
__device__ __forceinline__ int32_t row_start(int32_t fa, int32_t fb, int i)
{
    const int32_t va = __shfl(fa, i & (kWave - 1), kWave);
    const int32_t vb = __shfl(fb, i & (kWave - 1), kWave);
    return i < kWave ? va : vb;
}

template <bool ACCUM>
__global__ __launch_bounds__(kThreads) void spmv_seg2_kernel(int32_t        nchunks,
                                                            const int32_t* __restrict__ chunk_row,
                                                            const int32_t* __restrict__ fptr,
                                                            const int32_t* __restrict__ fcols,
                                                            const float* __restrict__ fvals,
                                                            int64_t       fstride,
                                                            const double* __restrict__ x,
                                                            const float4* __restrict__ x32,
                                                            double*       y,
                                                            double        a,
                                                            double*       partials,
                                                            const int32_t* d_conv)
{
    if(d_conv && *d_conv != 0)
        return;
    __shared__ double lds[kThreads / kWave];
    constexpr int     kWin = 2 * kWave;

    const int G    = gridDim.x;
    const int b    = blockIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const int nx   = (G % kXcds == 0) ? kXcds : 1;
    const int xcd  = b % nx;
    const int Wx   = (G / nx) * kWaves;
    const int t    = (b / nx) * kWaves + wave;
    const int cx1  = (int)((int64_t)nchunks * (xcd + 1) / nx);
    int       c    = (int)((int64_t)nchunks * xcd / nx) + t;

    const int lane = (int)(threadIdx.x % kWave);
    double    dot  = 0.0;
    if(c < cx1)
    {
        int32_t row    = chunk_row[c];
        int32_t rowEnd = chunk_row[c + 1];
        int32_t kStart = fptr[row];
        int32_t kEnd   = fptr[rowEnd];
        int32_t k      = kStart & ~1;
        double  carry0 = 0.0, carry1 = 0.0, carry2 = 0.0;
        Window2 nw;
        load_window2(nw, k, kEnd, lane, fcols, fvals, fstride);
        for(;;)
        {
            const Window2 cw  = nw;
            const int32_t kb0 = k + 2 * lane, kb1 = kb0 + 1;
            const bool    lead = kb0 < kStart;  // the block before an odd chunk start
            const bool    ok0 = !lead && kb0 < kEnd, ok1 = kb1 < kEnd;
            // columns of blocks outside [kStart, kEnd) are never used (the pair's
            // second half past a chunk's end may be unwritten slack)
            // x is gathered from its FP32 copy, one 16-byte load per block
            const uint32_t j0 = ok0 ? (uint32_t)cw.j.x : 0u, j1 = ok1 ? (uint32_t)cw.j.y : 0u;
            const float4   g0 = x32[j0], g1 = x32[j1];
            const double   x00 = g0.x, x01 = g0.y, x02 = g0.z;
            const double   x10 = g1.x, x11 = g1.y, x12 = g1.z;
            const int32_t fa = fptr[min(row + lane, rowEnd)];
            const int32_t fb = fptr[min(row + kWave + lane, rowEnd)];
            const bool    same = k + kWin < kEnd;
            int32_t       nrow = 0, nrowEnd = 0, nk = 0, nkStart = 0, nkEnd = 0;
            const int     cn   = c + Wx;
            if(same)
            {
                nk      = k + kWin;
                nkStart = kStart;
                nkEnd   = kEnd;
            }
            else if(cn < cx1)
            {
                nrow    = chunk_row[cn];
                nrowEnd = chunk_row[cn + 1];
                nkStart = fptr[nrow];
                nkEnd   = fptr[nrowEnd];
                nk      = nkStart & ~1;
            }
            if(same || cn < cx1)
                load_window2(nw, nk, nkEnd, lane, fcols, fvals, fstride);
            // rows of the lane's blocks: search the first, the second is the
            // same row or the next one (every row holds a block)
            const int32_t kq  = lead ? kStart : kb0;
            int           idx = 0;
#pragma unroll
            for(int step = kWave; step > 0; step >>= 1)
                if(row_start(fa, fb, idx + step < kWin ? idx + step : kWin - 1) <= kq && idx + step < kWin)
                    idx += step;
            int32_t       f0cur  = row_start(fa, fb, idx);
            int32_t       f0next = row_start(fa, fb, min(idx + 1, kWin - 1));
            const int32_t row0   = row + idx;
            if(idx + 1 >= kWin)
                f0next = fptr[min(row0 + 1, rowEnd)];
            const bool    h1   = ok1 && (lead || kb1 == f0next);
            const int32_t row1 = (h1 && !lead) ? row0 + 1 : row0;
            const int     idx1 = row1 - row;
            int32_t       f1next = row_start(fa, fb, min(idx1 + 1, kWin - 1));
            if(idx1 + 1 >= kWin)
                f1next = fptr[min(row1 + 1, rowEnd)];
            const bool h0 = ok0 && f0cur == kb0;
            // block products
            double w00 = 0.0, w01 = 0.0, w02 = 0.0, w10 = 0.0, w11 = 0.0, w12 = 0.0;
            if(ok0)
            {
                w00 = (double)cw.A[0].x * x00 + (double)cw.A[3].x * x01 + (double)cw.A[6].x * x02;
                w01 = (double)cw.A[1].x * x00 + (double)cw.A[4].x * x01 + (double)cw.A[7].x * x02;
                w02 = (double)cw.A[2].x * x00 + (double)cw.A[5].x * x01 + (double)cw.A[8].x * x02;
            }
            if(ok1)
            {
                w10 = (double)cw.A[0].y * x10 + (double)cw.A[3].y * x11 + (double)cw.A[6].y * x12;
                w11 = (double)cw.A[1].y * x10 + (double)cw.A[4].y * x11 + (double)cw.A[7].y * x12;
                w12 = (double)cw.A[2].y * x10 + (double)cw.A[5].y * x11 + (double)cw.A[8].y * x12;
            }
            // the lane's last segment inside the lane, and whether it starts here
            double v0, v1, v2;
            bool   brk;
            if(ok1 && !h1)
            {
                v0  = w00 + w10;
                v1  = w01 + w11;
                v2  = w02 + w12;
                brk = h0;
            }
            else if(ok1)
            {
                v0  = w10;
                v1  = w11;
                v2  = w12;
                brk = true;
            }
            else
            {
                v0  = w00;
                v1  = w01;
                v2  = w02;
                brk = h0 || !ok0;
            }
            const uint64_t bm = __ballot(brk);
            int            fl = brk ? 1 : 0;
#pragma unroll
            for(int d = 1; d < kWave; d <<= 1)
            {
                const double t0 = __shfl_up(v0, d, kWave), t1 = __shfl_up(v1, d, kWave), t2 = __shfl_up(v2, d, kWave);
                const int    tf = __shfl_up(fl, d, kWave);
                if(lane >= d)
                {
                    if(!fl)
                    {
                        v0 += t0;
                        v1 += t1;
                        v2 += t2;
                    }
                    fl |= tf;
                }
            }
            // the previous lane's inclusive value (the segment the first block continues)
            double p0 = __shfl_up(v0, 1, kWave), p1 = __shfl_up(v1, 1, kWave), p2 = __shfl_up(v2, 1, kWave);
            if(lane == 0)
                p0 = p1 = p2 = 0.0;
            const bool none_below = (bm & ((1ull << lane) - 1ull)) == 0ull;   // no break in lanes [0, l)
            const bool none_upto  = (bm & ((2ull << lane) - 1ull)) == 0ull;   // ... in lanes [0, l]
            if(none_below && !h0)
            {
                p0 += carry0;
                p1 += carry1;
                p2 += carry2;
            }
            if(none_upto)
            {
                v0 += carry0;
                v1 += carry1;
                v2 += carry2;
            }
            const bool end0 = ok0 && (h1 || !ok1) && kb0 + 1 == f0next;
            const bool end1 = ok1 && kb1 + 1 == f1next;
            if(end0)
            {
                const double e0 = (h0 ? 0.0 : p0) + w00, e1 = (h0 ? 0.0 : p1) + w01, e2 = (h0 ? 0.0 : p2) + w02;
                put_row<ACCUM>(y, row0, a, e0, e1, e2);
                dot += x[3 * (int64_t)row0] * e0 + x[3 * (int64_t)row0 + 1] * e1 + x[3 * (int64_t)row0 + 2] * e2;
            }
            if(end1)
            {
                put_row<ACCUM>(y, row1, a, v0, v1, v2);
                dot += x[3 * (int64_t)row1] * v0 + x[3 * (int64_t)row1 + 1] * v1 + x[3 * (int64_t)row1 + 2] * v2;
            }
            if(same)
            {
                // lane 63 hands its row on: its second block is valid (the
                // window is not the chunk's last); the row continues or ends
                const int  e63   = __shfl(end1 ? 1 : 0, kWave - 1, kWave);
                const int  r63   = __shfl(row1, kWave - 1, kWave);
                const bool cont  = e63 == 0;
                carry0           = cont ? __shfl(v0, kWave - 1, kWave) : 0.0;
                carry1           = cont ? __shfl(v1, kWave - 1, kWave) : 0.0;
                carry2           = cont ? __shfl(v2, kWave - 1, kWave) : 0.0;
                row              = cont ? r63 : r63 + 1;
                k                = nk;
            }
            else
            {
                if(cn >= cx1)
                    break;
                c      = cn;
                row    = nrow;
                rowEnd = nrowEnd;
                kStart = nkStart;
                k      = nk;
                kEnd   = nkEnd;
                carry0 = carry1 = carry2 = 0.0;
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

// FP64 planes: a workgroup takes whole chunks (rows whose first block lies in
// [c S, (c+1) S)), loads the chunk's blocks with consecutive lanes on
// consecutive blocks, stages w_k = A_k x_col(k) in LDS and sums each row from
// LDS with 16 lanes in a fixed order.  A chunk longer than kChunkCap blocks
// (a row of more than S blocks) is done row by row straight from memory.
template <typename VT, bool ACCUM>
__global__ __launch_bounds__(kThreads) void spmv_chunk_kernel(int32_t        nchunks,
                                                             const int32_t* __restrict__ chunk_row,
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
    __shared__ double w[3][kChunkCap];
    __shared__ double lds[kThreads / kWave];

    const int G   = gridDim.x;
    const int b   = blockIdx.x;
    const int per = G / kXcds;
    const int lb  = (per > 0 && b < per * kXcds) ? (b % kXcds) * per + b / kXcds : b;
    const int c0  = (int)((int64_t)nchunks * lb / G);
    const int c1  = (int)((int64_t)nchunks * (lb + 1) / G);

    const int tid  = threadIdx.x;
    const int g    = tid / kLanes;
    const int lane = tid % kLanes;
    double    dot  = 0.0;
    for(int c = c0; c < c1; ++c)
    {
        const int32_t row0 = chunk_row[c], row1 = chunk_row[c + 1];
        const int32_t k0 = fptr[row0], k1 = fptr[row1];
        if(k1 - k0 <= kChunkCap)
        {
            // products of the chunk's blocks, consecutive lanes on consecutive blocks
#pragma unroll
            for(int h = 0; h < kPer; h += 2)
            {
                uint32_t jj[2];
                bool     ok[2];
                VT       A[2][9];
#pragma unroll
                for(int u = 0; u < 2; ++u)
                {
                    const int32_t k = k0 + (h + u) * kThreads + tid;
                    ok[u]           = k < k1;
                    jj[u]           = ok[u] ? (uint32_t)fcols[k] : 0u;
#pragma unroll
                    for(int q = 0; q < 9; ++q)
                        A[u][q] = ok[u] ? fvals[q * fstride + k] : VT(0);
                }
#pragma unroll
                for(int u = 0; u < 2; ++u)
                {
                    const double x0 = x[3u * jj[u]], x1 = x[3u * jj[u] + 1], x2 = x[3u * jj[u] + 2];
                    if(ok[u])
                    {
                        const int i = (h + u) * kThreads + tid;
                        w[0][i]     = (double)A[u][0] * x0 + (double)A[u][3] * x1 + (double)A[u][6] * x2;
                        w[1][i]     = (double)A[u][1] * x0 + (double)A[u][4] * x1 + (double)A[u][7] * x2;
                        w[2][i]     = (double)A[u][2] * x0 + (double)A[u][5] * x1 + (double)A[u][8] * x2;
                    }
                }
            }
            __syncthreads();
            for(int32_t row = row0 + g; row < row1; row += kGroups)
            {
                const int32_t s = fptr[row] - k0, e = fptr[row + 1] - k0;
                double        acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
                for(int32_t i = s + lane; i < e; i += kLanes)
                {
                    acc0 += w[0][i];
                    acc1 += w[1][i];
                    acc2 += w[2][i];
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
                    put_row<ACCUM>(y, row, a, acc0, acc1, acc2);
                    dot += x[3 * (int64_t)row] * acc0 + x[3 * (int64_t)row + 1] * acc1
                           + x[3 * (int64_t)row + 2] * acc2;
                }
            }
            __syncthreads();  // w is rewritten by the next chunk
        }
        else
        {
            // a chunk with a very long row: 16 lanes per row straight from memory
            for(int32_t row = row0 + g; row < row1; row += kGroups)
            {
                const int32_t s = fptr[row], e = fptr[row + 1];
                double        acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
                for(int32_t k = s + lane; k < e; k += kLanes)
                {
                    const uint32_t j = (uint32_t)fcols[k];
                    double         A[9];
#pragma unroll
                    for(int q = 0; q < 9; ++q)
                        A[q] = (double)fvals[q * fstride + k];
                    const double x0 = x[3u * j], x1 = x[3u * j + 1], x2 = x[3u * j + 2];
                    acc0 += A[0] * x0 + A[3] * x1 + A[6] * x2;
                    acc1 += A[1] * x0 + A[4] * x1 + A[7] * x2;
                    acc2 += A[2] * x0 + A[5] * x1 + A[8] * x2;
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
                    put_row<ACCUM>(y, row, a, acc0, acc1, acc2);
                    dot += x[3 * (int64_t)row] * acc0 + x[3 * (int64_t)row + 1] * acc1
                           + x[3 * (int64_t)row + 2] * acc2;
                }
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

// launches the SpMV; returns how many partials it wrote from partials[0].
// FP64 planes go through the LDS chunk kernel, FP32 planes through the
// two-blocks-per-lane segmented-scan kernel.
template <typename VT, bool ACCUM>
int launch_spmv(const SpmvScratch&, int g, hipStream_t s, const BsrMatrix& A, const VT* vals, const double* x,
                const float4* x32, double* y, double a, double* partials, const int32_t* d_conv)
{
    if constexpr(sizeof(VT) == 8)
        spmv_chunk_kernel<VT, ACCUM><<<g, kThreads, 0, s>>>(A.nchunks, A.chunk_row.data(), A.full_ptr.data(),
                                                            A.fcols.data(), vals, A.fstride, x, y, a, partials,
                                                            d_conv);
    else
        spmv_seg2_kernel<ACCUM><<<g, kThreads, 0, s>>>(A.nchunks, A.chunk_row.data(), A.full_ptr.data(),
                                                       A.fcols.data(), vals, A.fstride, x, x32, y, a, partials,
                                                       d_conv);
    return g;
}

void check_matrix(const BsrMatrix& A, bool use32)
{
    MI_REQUIRE(A.full_ptr.data() && A.fcols.data() && A.chunk_row.data() && A.nchunks > 0, MI_ERR_STATE,
               "spmv: matrix not assembled");
    MI_REQUIRE(use32 ? A.has32 : (A.fvals.data() != nullptr), MI_ERR_STATE,
               "spmv: values of this precision were not assembled");
}
}  // namespace

void spmv_prepare(SpmvScratch& scr, bool use32)
{
    if(scr.grid == 0 || scr.use32 != use32)
    {
        scr.use32 = use32;
        // a persistent grid of as many workgroups per CU as fit at once (the
        // kernel's occupancy); MI_SPMV_BLOCKS_PER_CU overrides for sweeps
        int per_cu = 0;
        if(const char* e = std::getenv("MI_SPMV_BLOCKS_PER_CU"))
            per_cu = std::max(1, std::min(32, std::atoi(e)));
        if(per_cu == 0)
        {
            const void* fn = use32 ? (const void*)spmv_seg2_kernel<false>
                                   : (const void*)spmv_chunk_kernel<double, false>;
            MI_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kThreads, 0));
            per_cu = std::max(1, per_cu);
        }
        int grid = device_cu_count() * per_cu;
        grid     = std::max(kXcds, (grid / kXcds) * kXcds);
        scr.partials.resize((size_t)2 * grid);
        scr.grid = grid;
    }
}

int spmv_grid(const SpmvScratch& scr, const BsrMatrix& A)
{
    const int64_t per_wg = scr.use32 ? kWaves : 1;  // chunks a workgroup takes at once
    return (int)std::max<int64_t>(1, std::min<int64_t>(scr.grid, (A.nchunks + per_wg - 1) / per_wg));
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
              const int32_t* d_conv, SpmvScratch& scr, hipStream_t s)
{
    spmv_prepare(scr, use32);
    if(A.B == 0)
        return 0;
    check_matrix(A, use32);
    const int g = spmv_grid(scr, A);
    MI_REQUIRE(x && y && partials, MI_ERR_STATE, "spmv_full: null buffer");
    MI_REQUIRE(partials >= scr.partials.data() && partials + g <= scr.partials.data() + scr.partials.size(),
               MI_ERR_STATE, "spmv_full: partials outside the scratch");
    const int np = use32 ? launch_spmv<float, false>(scr, g, s, A, A.fvals32.data(), x, x32_of(A, x, x32, scr, s), y,
                                                     1.0, partials, d_conv)
                         : launch_spmv<double, false>(scr, g, s, A, A.fvals.data(), x, nullptr, y, 1.0, partials,
                                                      d_conv);
    MI_HIP(hipGetLastError());
    return np;
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
    const int np = use32 ? launch_spmv<float, true>(scr, g, s, A, A.fvals32.data(), x, x32_of(A, x, nullptr, scr, s),
                                                    y, a, p, nullptr)
                         : launch_spmv<double, true>(scr, g, s, A, A.fvals.data(), x, nullptr, y, a, p, nullptr);
    if(d_dot)
        sum_partials_kernel<<<1, kThreads, 0, s>>>(scr.partials.data(), np, d_dot);
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
