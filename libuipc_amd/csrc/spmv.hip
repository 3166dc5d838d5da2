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
// gather over it, streamed as one flat sequence of blocks:
//   * work comes in chunks of whole rows (the rows whose first block lies in
//     [c S, (c+1) S), kChunkStep S); a wave walks a chunk's blocks in windows
//     of 64, lane l on block k + l, so every value plane is
//     read as 64 contiguous entries per instruction (the rate of a plain
//     stream; 16 lanes per row of ~40 blocks reached ~60% of it);
//   * a lane's row is found by a binary search over the 64 row starts that
//     can meet the window (fptr[row .. row + 63], held one per lane, read
//     with shuffles); w = A_k x_col(k) is summed per row by a segmented
//     inclusive scan over the wave (xor-free shuffle-up steps, fixed order:
//     deterministic, no LDS, no barrier); a row that runs past the window
//     carries its partial sum into the next one;
//   * the lane holding a row's last block stores y (no atomics) and adds
//     x_row . (A x)_row to the wave's x^T A x; workgroup partials go to
//     partials[blockIdx.x], which the consumer kernel (PCG update) sums in a
//     fixed order;
//   * a persistent grid; each XCD (workgroups are dispatched round-robin over
//     the 8 XCDs) takes one contiguous eighth of the chunks and deals them to
//     its waves round-robin, so the rows in flight form a narrow band that
//     sweeps the eighth and x is re-read from the XCD's L2, not from HBM;
//   * each wave loads its next window while the current one's x gathers are
//     outstanding (double buffer).
#include <cstdlib>
#include <string>

#include "linsys.hpp"

namespace mi
{
namespace
{
constexpr int kThreads = 256;
constexpr int kWaves   = kThreads / kWave;
constexpr int kXcds    = 8;

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

// ---- front variant: the 64-block windows of the whole matrix are dealt
// round-robin to all waves, so the blocks in flight form one contiguous front
// (the access pattern of a plain stream).  A window's rows that start and end
// in it are stored at once; a row crossing windows leaves its head part (S,
// the segment from its first block to the window end) and every later window
// its part (P, flagged when the row ends there), and spmv_fixup_kernel adds
// them up in window order.
enum : int32_t
{
    kWinHasP  = 1,  // the window's first segment belongs to a row begun earlier
    kWinPEnds = 2,  // ... and that row ends in this window
    kWinHasS  = 4   // a row begun in this window runs past its end
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

template <typename VT, bool ACCUM>
__global__ __launch_bounds__(kThreads) void spmv_front_kernel(int32_t        nwin,
                                                             int32_t        B,
                                                             const int32_t* __restrict__ win_row,
                                                             const int32_t* __restrict__ fptr,
                                                             const int32_t* __restrict__ fcols,
                                                             const VT* __restrict__ fvals,
                                                             int64_t       fstride,
                                                             const double* __restrict__ x,
                                                             double*       y,
                                                             double        a,
                                                             double*       P,
                                                             double*       S,
                                                             int32_t*      srow,
                                                             int32_t*      flags,
                                                             double*       partials,
                                                             const int32_t* d_conv)
{
    if(d_conv && *d_conv != 0)
        return;
    __shared__ double lds[kThreads / kWave];
    const int     lane = (int)(threadIdx.x % kWave);
    const int     wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const int64_t W    = (int64_t)gridDim.x * kWaves;
    const int32_t nnz  = fptr[B];
    double        dot  = 0.0;
    for(int64_t w = (int64_t)blockIdx.x * kWaves + wave; w < nwin; w += W)
    {
        const int32_t k = (int32_t)(w * kWave);
        if(k >= nnz)
            break;
        const int32_t  kb = k + lane;
        const bool     ok = kb < nnz;
        const uint32_t j  = ok ? (uint32_t)fcols[kb] : 0u;
        VT             A[9];
#pragma unroll
        for(int q = 0; q < 9; ++q)
            A[q] = ok ? fvals[q * fstride + kb] : VT(0);
        const int32_t row = win_row[w];
        const int32_t f   = fptr[min(row + lane, B)];
        const double  x0 = x[3u * j], x1 = x[3u * j + 1], x2 = x[3u * j + 2];
        int           idx = 0;
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
            fnext = fptr[min(myrow + 1, B)];
        const bool head = ok && fcur == kb;
        double     v0 = 0.0, v1 = 0.0, v2 = 0.0;
        if(ok)
        {
            v0 = (double)A[0] * x0 + (double)A[3] * x1 + (double)A[6] * x2;
            v1 = (double)A[1] * x0 + (double)A[4] * x1 + (double)A[7] * x2;
            v2 = (double)A[2] * x0 + (double)A[5] * x1 + (double)A[8] * x2;
        }
        const uint64_t hm = __ballot(head);
        const uint64_t bm = __ballot(head || !ok);
        int            fl = (head || !ok) ? 1 : 0;
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
        const bool first_seg = (bm & ((2ull << lane) - 1ull)) == 0ull;  // row begun before the window
        const bool row_end   = ok && kb + 1 == fnext;
        if(row_end && !first_seg)
        {
            put_row<ACCUM>(y, myrow, a, v0, v1, v2);
            dot += x[3 * (int64_t)myrow] * v0 + x[3 * (int64_t)myrow + 1] * v1 + x[3 * (int64_t)myrow + 2] * v2;
        }
        const int      last_valid = min(kWave - 1, nnz - 1 - k);
        const uint64_t rem        = bm & ~1ull;
        const int      fend       = rem ? (__builtin_ctzll(rem) - 1) : last_valid;  // end of the first segment
        const int      fend_row   = __shfl(row_end ? 1 : 0, fend, kWave);
        const int      last_end   = __shfl(row_end ? 1 : 0, last_valid, kWave);
        int            flg        = 0;
        if(!(hm & 1ull))
        {
            flg |= kWinHasP | (fend_row ? kWinPEnds : 0);
            if(lane == fend)
            {
                P[3 * w]     = v0;
                P[3 * w + 1] = v1;
                P[3 * w + 2] = v2;
            }
        }
        if(!last_end && hm != 0ull && last_valid == kWave - 1)
        {
            flg |= kWinHasS;
            if(lane == kWave - 1)
            {
                S[3 * w]     = v0;
                S[3 * w + 1] = v1;
                S[3 * w + 2] = v2;
                srow[w]      = myrow;
            }
        }
        if(lane == 0)
            flags[w] = flg;
    }
    if(partials)
    {
        double bs = block_sum<kThreads>(dot, lds);
        if(threadIdx.x == 0)
            partials[blockIdx.x] = bs;
    }
}

// rows that cross windows: S of the window where the row begins plus P of every
// later window up to the one where it ends, in window order
template <bool ACCUM>
__global__ __launch_bounds__(kThreads) void spmv_fixup_kernel(int32_t        nwin,
                                                             int32_t        B,
                                                             const int32_t* __restrict__ fptr,
                                                             const double* __restrict__ P,
                                                             const double* __restrict__ S,
                                                             const int32_t* __restrict__ srow,
                                                             const int32_t* __restrict__ flags,
                                                             const double* __restrict__ x,
                                                             double*       y,
                                                             double        a,
                                                             double*       partials,
                                                             const int32_t* d_conv)
{
    if(d_conv && *d_conv != 0)
        return;
    __shared__ double lds[kThreads / kWave];
    const int32_t nw  = (int32_t)std::min<int64_t>(nwin, ((int64_t)fptr[B] + kWave - 1) / kWave);
    double        dot = 0.0;
    for(int32_t w = blockIdx.x * kThreads + threadIdx.x; w < nw; w += gridDim.x * kThreads)
    {
        if(!(flags[w] & kWinHasS))
            continue;
        double s0 = S[3 * (int64_t)w], s1 = S[3 * (int64_t)w + 1], s2 = S[3 * (int64_t)w + 2];
        for(int32_t u = w + 1; u < nw; ++u)
        {
            s0 += P[3 * (int64_t)u];
            s1 += P[3 * (int64_t)u + 1];
            s2 += P[3 * (int64_t)u + 2];
            if(flags[u] & kWinPEnds)
                break;
        }
        const int64_t r = srow[w];
        put_row<ACCUM>(y, r, a, s0, s1, s2);
        dot += x[3 * r] * s0 + x[3 * r + 1] * s1 + x[3 * r + 2] * s2;
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

int fix_grid(const BsrMatrix& A)
{
    return std::max(1, std::min(ceil_div(A.nwin, kThreads), device_cu_count() * 4));
}

// launches the SpMV; returns how many partials it wrote from partials[0]
template <typename VT, bool ACCUM>
int launch_spmv(const SpmvScratch& scr, int g, hipStream_t s, const BsrMatrix& A, const VT* vals, const double* x,
                double* y, double a, double* partials, const int32_t* d_conv)
{
    if(scr.front)
    {
        const int gf = fix_grid(A);
        spmv_front_kernel<VT, ACCUM><<<g, kThreads, 0, s>>>(A.nwin, A.B, A.win_row.data(), A.full_ptr.data(),
                                                            A.fcols.data(), vals, A.fstride, x, y, a, A.win_P.data(),
                                                            A.win_S.data(), A.win_srow.data(), A.win_flags.data(),
                                                            partials, d_conv);
        spmv_fixup_kernel<ACCUM><<<gf, kThreads, 0, s>>>(A.nwin, A.B, A.full_ptr.data(), A.win_P.data(),
                                                         A.win_S.data(), A.win_srow.data(), A.win_flags.data(), x, y,
                                                         a, partials ? partials + g : nullptr, d_conv);
        return g + gf;
    }
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
        // MI_SPMV_KERNEL=front selects the whole-matrix front variant
        const char* kv = std::getenv("MI_SPMV_KERNEL");
        scr.front      = kv && std::string(kv) == "front";
        // a persistent grid of as many workgroups per CU as fit at once (the
        // kernel's occupancy); MI_SPMV_BLOCKS_PER_CU overrides for sweeps
        int per_cu = 0;
        if(const char* e = std::getenv("MI_SPMV_BLOCKS_PER_CU"))
            per_cu = std::max(1, std::min(32, std::atoi(e)));
        if(per_cu == 0)
        {
            const void* fn = use32 ? (const void*)spmv_seg_kernel<float, false>
                                   : (const void*)spmv_seg_kernel<double, false>;
            MI_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kThreads, 0));
            per_cu = std::max(1, per_cu);
        }
        int grid = device_cu_count() * per_cu;
        grid     = std::max(kXcds, (grid / kXcds) * kXcds);
        // the front variant's fix-up adds up to 4 partials per CU after the grid's
        scr.partials.resize((size_t)2 * grid + 4 * device_cu_count());
        scr.grid = grid;
    }
}

int spmv_grid(const SpmvScratch& scr, const BsrMatrix& A)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>(scr.grid, (A.nchunks + kWaves - 1) / kWaves));
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
