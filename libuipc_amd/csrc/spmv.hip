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
//   * FP32 planes (spmv_seg_kernel): a wave walks a chunk in windows of 64
//     blocks, finds each lane's row by a shuffle binary search over the row
//     starts, sums rows with a segmented shuffle scan (carrying a row across
//     windows) and loads the next window while this one's x gathers are
//     outstanding;
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

template <typename VT>
struct Window
{
    uint32_t j;
    VT       A[9];
};

template <typename VT>
__device__ __forceinline__ void load_window(Window<VT>& w,
                                            int32_t     k,
                                            int32_t     kEnd,
                                            int         lane,
                                            const int32_t* __restrict__ fcols,
                                            const VT* __restrict__ fvals,
                                            int64_t fstride)
{
    const int32_t kb = k + lane;
    const bool    ok = kb < kEnd;
    w.j              = ok ? (uint32_t)fcols[kb] : 0u;
#pragma unroll
    for(int q = 0; q < 9; ++q)
        w.A[q] = ok ? fvals[q * fstride + kb] : VT(0);
}

template <typename VT, bool ACCUM>
__global__ __launch_bounds__(kThreads) void spmv_seg_kernel(int32_t        nchunks,
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
    __shared__ double lds[kThreads / kWave];

    // chunks of one XCD are dealt to its waves round-robin, so the rows in
    // flight on an XCD form a narrow band whose slice of x stays in its L2
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
        int32_t k      = fptr[row];
        int32_t kEnd   = fptr[rowEnd];
        double  carry0 = 0.0, carry1 = 0.0, carry2 = 0.0;
        // double buffer: the next window's columns and values are loaded while
        // this window's x gathers are outstanding, so a window costs one memory
        // round trip instead of two (columns, then x)
        Window<VT> nw;
        load_window<VT>(nw, k, kEnd, lane, fcols, fvals, fstride);
        for(;;)
        {
            const Window<VT> cw = nw;
            const int32_t    kb = k + lane;
            const bool       ok = kb < kEnd;
            // x gathers and the starts of the rows that can meet the window
            // (every row holds a block, so at most 64 of them; past rowEnd the
            // start is kEnd)
            const double  x0 = x[3u * cw.j], x1 = x[3u * cw.j + 1], x2 = x[3u * cw.j + 2];
            const int32_t f  = fptr[min(row + lane, rowEnd)];
            // where the next window is: further on in this chunk, or the start
            // of the wave's next chunk
            const bool same = k + kWave < kEnd;
            int32_t    nrow = 0, nrowEnd = 0, nk = 0, nkEnd = 0;
            const int  cn   = c + Wx;
            if(same)
            {
                nk    = k + kWave;
                nkEnd = kEnd;
            }
            else if(cn < cx1)
            {
                nrow    = chunk_row[cn];
                nrowEnd = chunk_row[cn + 1];
                nk      = fptr[nrow];
                nkEnd   = fptr[nrowEnd];
            }
            if(same || cn < cx1)
                load_window<VT>(nw, nk, nkEnd, lane, fcols, fvals, fstride);
            // row of block kb: the last i with f_i <= kb
            int idx = 0;
#pragma unroll
            for(int step = kWave / 2; step > 0; step >>= 1)
            {
                const int32_t fc = __shfl(f, idx + step, kWave);
                if(fc <= kb)
                    idx += step;
            }
            const int32_t fcur  = __shfl(f, idx, kWave);
            int32_t       fnext = __shfl(f, min(idx + 1, kWave - 1), kWave);
            const int32_t myrow = row + idx;
            if(idx == kWave - 1)
                fnext = fptr[min(myrow + 1, rowEnd)];
            const bool head = ok && fcur == kb;
            double     v0 = 0.0, v1 = 0.0, v2 = 0.0;
            if(ok)
            {
                v0 = (double)cw.A[0] * x0 + (double)cw.A[3] * x1 + (double)cw.A[6] * x2;
                v1 = (double)cw.A[1] * x0 + (double)cw.A[4] * x1 + (double)cw.A[7] * x2;
                v2 = (double)cw.A[2] * x0 + (double)cw.A[5] * x1 + (double)cw.A[8] * x2;
            }
            // segmented inclusive scan: each lane ends with the sum of its row's
            // blocks from the row's head (or lane 0) up to itself
            const uint64_t hmask = __ballot(head || !ok);
            int            fl    = (head || !ok) ? 1 : 0;
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
            // the row continuing from the previous window adds the carry
            if((hmask & ((2ull << lane) - 1ull)) == 0ull)
            {
                v0 += carry0;
                v1 += carry1;
                v2 += carry2;
            }
            const bool row_end = ok && kb + 1 == fnext;
            if(row_end)
            {
                double* yr = y + 3 * (int64_t)myrow;
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
                dot += x[3 * (int64_t)myrow] * v0 + x[3 * (int64_t)myrow + 1] * v1 + x[3 * (int64_t)myrow + 2] * v2;
            }
            if(same)
            {
                // lane 63 hands its row on: either the row continues (carry)
                // or the next window starts the next row
                const int  end63 = __shfl(row_end ? 1 : 0, kWave - 1, kWave);
                const int  row63 = __shfl(myrow, kWave - 1, kWave);
                const bool cont  = end63 == 0;
                carry0           = cont ? __shfl(v0, kWave - 1, kWave) : 0.0;
                carry1           = cont ? __shfl(v1, kWave - 1, kWave) : 0.0;
                carry2           = cont ? __shfl(v2, kWave - 1, kWave) : 0.0;
                row              = cont ? row63 : row63 + 1;
                k                = nk;
            }
            else
            {
                // a chunk ends with a whole row: nothing carries over
                if(cn >= cx1)
                    break;
                c      = cn;
                row    = nrow;
                rowEnd = nrowEnd;
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
// segmented-scan kernel (each measured fastest for its precision on the 64^3
// FEM + 2M contact system).
template <typename VT, bool ACCUM>
int launch_spmv(const SpmvScratch&, int g, hipStream_t s, const BsrMatrix& A, const VT* vals, const double* x,
                double* y, double a, double* partials, const int32_t* d_conv)
{
    if constexpr(sizeof(VT) == 8)
        spmv_chunk_kernel<VT, ACCUM><<<g, kThreads, 0, s>>>(A.nchunks, A.chunk_row.data(), A.full_ptr.data(),
                                                            A.fcols.data(), vals, A.fstride, x, y, a, partials,
                                                            d_conv);
    else
        spmv_seg_kernel<VT, ACCUM><<<g, kThreads, 0, s>>>(A.nchunks, A.chunk_row.data(), A.full_ptr.data(),
                                                          A.fcols.data(), vals, A.fstride, x, y, a, partials, d_conv);
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
            const void* fn = use32 ? (const void*)spmv_seg_kernel<float, false>
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

int spmv_full(const BsrMatrix& A, bool use32, const double* x, double* y, double* partials, const int32_t* d_conv,
              SpmvScratch& scr, hipStream_t s)
{
    spmv_prepare(scr, use32);
    if(A.B == 0)
        return 0;
    check_matrix(A, use32);
    const int g = spmv_grid(scr, A);
    MI_REQUIRE(x && y && partials, MI_ERR_STATE, "spmv_full: null buffer");
    MI_REQUIRE(partials >= scr.partials.data() && partials + g <= scr.partials.data() + scr.partials.size(),
               MI_ERR_STATE, "spmv_full: partials outside the scratch");
    const int np = use32 ? launch_spmv<float, false>(scr, g, s, A, A.fvals32.data(), x, y, 1.0, partials, d_conv)
                         : launch_spmv<double, false>(scr, g, s, A, A.fvals.data(), x, y, 1.0, partials, d_conv);
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
    const int np = use32 ? launch_spmv<float, true>(scr, g, s, A, A.fvals32.data(), x, y, a, p, nullptr)
                         : launch_spmv<double, true>(scr, g, s, A, A.fvals.data(), x, y, a, p, nullptr);
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
