// Triplet -> sorted upper-triangular 3x3 BCOO + BSR row offsets on gfx950.
//
// Replaces MatrixConverter<Float,3>::convert_sym
// (algorithm/details/matrix_converter.inl:1549-1699) and the segmented reduce
// it calls (algorithm/fast_segmental_reduce.h:33-103).  Same contract:
//   keep (row >= 0 && row <= col), key = row * B + col, stable sort on
//   bit_width(B*B - 1) bits (matrix_converter.inl:726-735), run-length encode,
//   sum each run of equal keys.
// Differences by design:
//   * the sort is a hand-written LSD radix sort (8-bit digits, wave64 ballot
//     ranking, LDS-staged stable scatter) instead of CUB;
//   * duplicates are summed sequentially in ascending source order by one
//     thread per output block, so the values are deterministic (the reference
//     sums in atomic arrival order, fast_segmental_reduce.h:23-26);
//   * values land in 9 SoA planes (see BsrMatrix);
//   * one host read-back (the unique count, to size the outputs) instead of two.
#include "linsys.hpp"

namespace mi
{
namespace
{
constexpr int kThreads = 256;
constexpr int kItems   = 16;
constexpr int kTile    = kThreads * kItems;  // 4096 keys per tile
constexpr int kWaves   = kThreads / kWave;

__device__ __forceinline__ bool keep_triplet(int32_t r, int32_t c, int32_t B)
{
    // the reference's flag (matrix_converter.inl:142) plus a bound on col,
    // which the reference leaves to its producers
    return r >= 0 && r <= c && c < B;
}

// ---- 1. count kept triplets per tile
__global__ __launch_bounds__(kThreads) void cv_count_kernel(int64_t        T,
                                                           const int32_t* rows,
                                                           const int32_t* cols,
                                                           int32_t        B,
                                                           uint32_t*      tile_cnt)
{
    __shared__ uint32_t lds[kWaves];
    const int64_t       base = (int64_t)blockIdx.x * kTile;
    uint32_t            c    = 0;
#pragma unroll
    for(int k = 0; k < kItems; ++k)
    {
        int64_t i = base + (int64_t)k * kThreads + threadIdx.x;
        if(i < T && keep_triplet(rows[i], cols[i], B))
            ++c;
    }
    uint32_t s = block_sum<kThreads>(c, lds);
    if(threadIdx.x == 0)
        tile_cnt[blockIdx.x] = s;
}

// ---- 2. stable compaction of (key, source index); thread t owns the
// contiguous items [t * kItems, (t + 1) * kItems) of the tile so that the
// block prefix follows index order
__global__ __launch_bounds__(kThreads) void cv_compact_kernel(int64_t         T,
                                                             const int32_t*  rows,
                                                             const int32_t*  cols,
                                                             int32_t         B,
                                                             const uint32_t* tile_off,
                                                             uint64_t*       keys,
                                                             int32_t*        srcs)
{
    __shared__ uint32_t lds[kWaves + 1];
    const int64_t       first = (int64_t)blockIdx.x * kTile + (int64_t)threadIdx.x * kItems;
    uint32_t            flags = 0;
    uint32_t            c     = 0;
#pragma unroll
    for(int k = 0; k < kItems; ++k)
    {
        int64_t i = first + k;
        if(i < T && keep_triplet(rows[i], cols[i], B))
        {
            flags |= 1u << k;
            ++c;
        }
    }
    uint32_t total;
    uint32_t o = block_exclusive_scan<kThreads>(c, lds, &total) + tile_off[blockIdx.x];
#pragma unroll
    for(int k = 0; k < kItems; ++k)
        if(flags & (1u << k))
        {
            int64_t i = first + k;
            keys[o]   = (uint64_t)rows[i] * (uint64_t)B + (uint64_t)cols[i];
            srcs[o]   = (int32_t)i;
            ++o;
        }
}

// ---- single-workgroup exclusive scan (n up to a few 100k), total -> *total
__global__ __launch_bounds__(1024) void scan_small_kernel(uint32_t* data, int64_t n, int64_t* total)
{
    __shared__ uint32_t lds[1024 / kWave + 1];
    __shared__ uint32_t carry_s;
    uint32_t            carry = 0;
    for(int64_t base = 0; base < n; base += 1024)
    {
        int64_t  i = base + threadIdx.x;
        uint32_t v = i < n ? data[i] : 0u;
        uint32_t tot;
        uint32_t e = block_exclusive_scan<1024>(v, lds, &tot);
        if(i < n)
            data[i] = carry + e;
        carry += tot;
    }
    if(threadIdx.x == 0)
    {
        carry_s = carry;
        if(total)
            *total = (int64_t)carry;
    }
}

// ---- large exclusive scan: per-chunk sums, scan of sums, down-sweep
constexpr int kScanChunk = kThreads * kItems;
__global__ __launch_bounds__(kThreads) void scan_reduce_kernel(const uint32_t* data, int64_t n, uint32_t* sums)
{
    __shared__ uint32_t lds[kWaves];
    const int64_t       base = (int64_t)blockIdx.x * kScanChunk;
    uint32_t            s    = 0;
#pragma unroll
    for(int k = 0; k < kItems; ++k)
    {
        int64_t i = base + (int64_t)k * kThreads + threadIdx.x;
        if(i < n)
            s += data[i];
    }
    s = block_sum<kThreads>(s, lds);
    if(threadIdx.x == 0)
        sums[blockIdx.x] = s;
}
__global__ __launch_bounds__(kThreads) void scan_down_kernel(uint32_t* data, int64_t n, const uint32_t* sums)
{
    __shared__ uint32_t lds[kWaves + 1];
    const int64_t       first = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kItems;
    uint32_t            v[kItems];
    uint32_t            c = 0;
#pragma unroll
    for(int k = 0; k < kItems; ++k)
    {
        int64_t i = first + k;
        v[k]      = i < n ? data[i] : 0u;
        c += v[k];
    }
    uint32_t o = block_exclusive_scan<kThreads>(c, lds, (uint32_t*)nullptr) + sums[blockIdx.x];
#pragma unroll
    for(int k = 0; k < kItems; ++k)
    {
        int64_t i = first + k;
        if(i < n)
            data[i] = o;
        o += v[k];
    }
}

// ---- radix sort: tile histogram (digit-major layout hist[d * ntiles + tile])
__global__ __launch_bounds__(kThreads) void rs_hist_kernel(const int64_t*  d_m,
                                                          const uint64_t* keys,
                                                          int             shift,
                                                          uint32_t        dmask,
                                                          uint32_t*       hist,
                                                          int             ntiles)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t m    = *d_m;
    const int64_t base = (int64_t)blockIdx.x * kTile;
#pragma unroll
    for(int k = 0; k < kItems; ++k)
    {
        int64_t i = base + (int64_t)k * kThreads + threadIdx.x;
        if(i < m)
            atomicAdd(&h[(uint32_t)(keys[i] >> shift) & dmask], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

__device__ __forceinline__ uint64_t lanemask_lt64()
{
    return (1ull << __lane_id()) - 1ull;
}

// ---- radix sort: stable scatter of one tile.  Wave w ranks the 64-key rounds
// of its contiguous quarter of the tile in order (ballot-matched peers +
// per-wave digit counters), the tile is reordered by digit in LDS, then each
// digit run is written out contiguously.
__global__ __launch_bounds__(kThreads) void rs_scatter_kernel(const int64_t*  d_m,
                                                             const uint64_t* kin,
                                                             const int32_t*  vin,
                                                             uint64_t*       kout,
                                                             int32_t*        vout,
                                                             int             shift,
                                                             uint32_t        dmask,
                                                             const uint32_t* hist_scan,
                                                             int             ntiles)
{
    __shared__ uint64_t sk[kTile];
    __shared__ int32_t  sv[kTile];
    __shared__ uint32_t wcnt[kWaves][256];
    __shared__ uint32_t dstart[256];
    __shared__ uint32_t gbase[256];
    __shared__ uint32_t lds[kWaves + 1];

    const int64_t m    = *d_m;
    const int64_t base = (int64_t)blockIdx.x * kTile;
    if(base >= m)
        return;  // uniform per block
    const int w    = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
#pragma unroll
    for(int q = 0; q < kWaves; ++q)
        wcnt[q][threadIdx.x] = 0;
    __syncthreads();

    uint64_t key[kItems];
    int32_t  val[kItems];
    uint32_t rank[kItems];
    uint32_t dig[kItems];
    const uint64_t lt = lanemask_lt64();
#pragma unroll
    for(int k = 0; k < kItems; ++k)
    {
        int64_t  i     = base + (int64_t)w * (kWave * kItems) + (int64_t)k * kWave + lane;
        bool     valid = i < m;
        key[k]         = valid ? kin[i] : 0ull;
        val[k]         = valid ? vin[i] : 0;
        uint32_t d     = valid ? ((uint32_t)(key[k] >> shift) & dmask) : 0u;
        dig[k]         = valid ? d : 0xFFFFFFFFu;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for(int b = 0; b < 8; ++b)
        {
            uint64_t bm = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bm : ~bm;
        }
        uint32_t r = 0;
        if(valid)
        {
            uint32_t cnt = wcnt[w][d];
            r            = cnt + (uint32_t)__popcll(peers & lt);
            int leader   = 63 - __clzll(peers);
            if(lane == leader)
                wcnt[w][d] = cnt + (uint32_t)__popcll(peers);
        }
        rank[k] = r;
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // per digit: wave prefixes (in place) and the tile total
    {
        const int d   = threadIdx.x;
        uint32_t  acc = 0;
#pragma unroll
        for(int q = 0; q < kWaves; ++q)
        {
            uint32_t c = wcnt[q][d];
            wcnt[q][d] = acc;
            acc += c;
        }
        uint32_t e = block_exclusive_scan<kThreads>(acc, lds, (uint32_t*)nullptr);
        dstart[d]  = e;
        gbase[d]   = hist_scan[(int64_t)d * ntiles + blockIdx.x];
    }
    __syncthreads();
#pragma unroll
    for(int k = 0; k < kItems; ++k)
        if(dig[k] != 0xFFFFFFFFu)
        {
            uint32_t pos = dstart[dig[k]] + wcnt[w][dig[k]] + rank[k];
            sk[pos]      = key[k];
            sv[pos]      = val[k];
        }
    __syncthreads();
    const int64_t cnt = (m - base) < kTile ? (m - base) : kTile;
#pragma unroll
    for(int k = 0; k < kItems; ++k)
    {
        int i = k * kThreads + threadIdx.x;
        if(i < cnt)
        {
            uint64_t kk  = sk[i];
            uint32_t d   = (uint32_t)(kk >> shift) & dmask;
            uint32_t out = gbase[d] + ((uint32_t)i - dstart[d]);
            kout[out]    = kk;
            vout[out]    = sv[i];
        }
    }
}

// ---- run-length encode: heads per tile
__global__ __launch_bounds__(kThreads) void rle_count_kernel(const int64_t* d_m, const uint64_t* keys, uint32_t* tile_cnt)
{
    __shared__ uint32_t lds[kWaves];
    const int64_t       m    = *d_m;
    const int64_t       base = (int64_t)blockIdx.x * kTile;
    uint32_t            c    = 0;
#pragma unroll
    for(int k = 0; k < kItems; ++k)
    {
        int64_t i = base + (int64_t)k * kThreads + threadIdx.x;
        if(i < m && (i == 0 || keys[i] != keys[i - 1]))
            ++c;
    }
    uint32_t s = block_sum<kThreads>(c, lds);
    if(threadIdx.x == 0)
        tile_cnt[blockIdx.x] = s;
}

__global__ __launch_bounds__(kThreads) void rle_write_kernel(const int64_t*  d_m,
                                                            const uint64_t* keys,
                                                            const uint32_t* tile_off,
                                                            int32_t         B,
                                                            int32_t*        out_rows,
                                                            int32_t*        out_cols,
                                                            int32_t*        seg_start)
{
    __shared__ uint32_t lds[kWaves + 1];
    const int64_t       m     = *d_m;
    const int64_t       first = (int64_t)blockIdx.x * kTile + (int64_t)threadIdx.x * kItems;
    uint32_t            flags = 0, c = 0;
#pragma unroll
    for(int k = 0; k < kItems; ++k)
    {
        int64_t i = first + k;
        if(i < m && (i == 0 || keys[i] != keys[i - 1]))
        {
            flags |= 1u << k;
            ++c;
        }
    }
    uint32_t o = block_exclusive_scan<kThreads>(c, lds, (uint32_t*)nullptr) + tile_off[blockIdx.x];
#pragma unroll
    for(int k = 0; k < kItems; ++k)
        if(flags & (1u << k))
        {
            int64_t  i   = first + k;
            uint64_t key = keys[i];
            int32_t  r   = (int32_t)(key / (uint64_t)B);
            out_rows[o]  = r;
            out_cols[o]  = (int32_t)(key - (uint64_t)r * (uint64_t)B);
            seg_start[o] = (int32_t)i;
            ++o;
        }
}

// ---- segmented reduce: one thread per unique block, duplicates summed in
// ascending source order (bit-identical to a sequential left fold)
__global__ __launch_bounds__(kThreads) void seg_reduce_kernel(int64_t        U,
                                                             const int64_t* d_m,
                                                             const int32_t* seg_start,
                                                             const int32_t* srcs,
                                                             const double*  vals,
                                                             double*        out,
                                                             float*         out32,
                                                             int64_t        stride)
{
    const int64_t m = *d_m;
    for(int64_t u = (int64_t)blockIdx.x * kThreads + threadIdx.x; u < U; u += (int64_t)gridDim.x * kThreads)
    {
        const int64_t b = seg_start[u];
        const int64_t e = (u + 1 < U) ? (int64_t)seg_start[u + 1] : m;
        double        acc[9];
        const double* v0 = vals + 9 * (int64_t)srcs[b];
#pragma unroll
        for(int q = 0; q < 9; ++q)
            acc[q] = v0[q];
        for(int64_t k = b + 1; k < e; ++k)
        {
            const double* v = vals + 9 * (int64_t)srcs[k];
#pragma unroll
            for(int q = 0; q < 9; ++q)
                acc[q] = acc[q] + v[q];
        }
#pragma unroll
        for(int q = 0; q < 9; ++q)
            out[q * stride + u] = acc[q];
        if(out32)
        {
#pragma unroll
            for(int q = 0; q < 9; ++q)
                out32[q * stride + u] = (float)acc[q];
        }
    }
}

// ---- BSR row offsets: row_ptr[r] = first u with rows[u] >= r
__global__ __launch_bounds__(kThreads) void row_ptr_kernel(int64_t U, const int32_t* rows, int32_t* row_ptr, int32_t B)
{
    for(int64_t u = (int64_t)blockIdx.x * kThreads + threadIdx.x; u <= U; u += (int64_t)gridDim.x * kThreads)
    {
        int32_t prev = (u == 0) ? -1 : rows[u - 1];
        int32_t cur  = (u == U) ? B : rows[u];
        for(int32_t r = prev + 1; r <= cur; ++r)
            row_ptr[r] = (int32_t)u;
    }
}

int key_bits_for(uint64_t max_key)
{
    int bits = 0;
    while(max_key > 0)
    {
        ++bits;
        max_key >>= 1;
    }
    return bits > 0 ? bits : 1;
}

void scan_exclusive(uint32_t* data, int64_t n, DBuf<uint32_t>& chunk, int64_t* total, hipStream_t s)
{
    if(n <= 64 * 1024)
    {
        scan_small_kernel<<<1, 1024, 0, s>>>(data, n, total);
        return;
    }
    int nch = ceil_div(n, kScanChunk);
    chunk.resize((size_t)nch + 1);
    scan_reduce_kernel<<<nch, kThreads, 0, s>>>(data, n, chunk.data());
    scan_small_kernel<<<1, 1024, 0, s>>>(chunk.data(), nch, total);
    scan_down_kernel<<<nch, kThreads, 0, s>>>(data, n, chunk.data());
}
}  // namespace

MatrixConverter::~MatrixConverter()
{
    if(pinned_)
        (void)hipHostFree(pinned_);
}

void MatrixConverter::convert_sym(int32_t        B,
                                  int64_t        T,
                                  const int32_t* rows,
                                  const int32_t* cols,
                                  const double*  vals,
                                  BsrMatrix&     out,
                                  bool           want32,
                                  double         ratio,
                                  hipStream_t    s)
{
    MI_REQUIRE(B >= 0 && T >= 0, MI_ERR_ARG, "convert_sym: negative size");
    MI_REQUIRE(T < ((int64_t)1 << 31), MI_ERR_ARG, "convert_sym: more than 2^31-1 triplets");
    MI_REQUIRE(T == 0 || (rows && cols && vals), MI_ERR_ARG, "convert_sym: null triplet arrays");
    out.B = B;
    out.row_ptr.resize((size_t)B + 1, ratio);
    if(!pinned_)
        MI_HIP(hipHostMalloc((void**)&pinned_, 2 * sizeof(int64_t)));
    counts_.resize(2);
    MI_HIP(hipMemsetAsync(counts_.data(), 0, 2 * sizeof(int64_t), s));
    if(T == 0 || B == 0)
    {
        out.U = 0;
        out.stride = 0;
        MI_HIP(hipMemsetAsync(out.row_ptr.data(), 0, sizeof(int32_t) * ((size_t)B + 1), s));
        return;
    }
    const int ntiles = ceil_div(T, kTile);
    tile_cnt_.resize((size_t)ntiles + 1, ratio);
    keys_[0].resize((size_t)T, ratio);
    keys_[1].resize((size_t)T, ratio);
    srcs_[0].resize((size_t)T, ratio);
    srcs_[1].resize((size_t)T, ratio);
    hist_.resize((size_t)256 * ntiles, ratio);

    int64_t* d_m = counts_.data();
    int64_t* d_u = counts_.data() + 1;

    // 1-3. flag + scan + compact (matrix_converter.inl:1579-1595)
    cv_count_kernel<<<ntiles, kThreads, 0, s>>>(T, rows, cols, B, tile_cnt_.data());
    scan_exclusive(tile_cnt_.data(), ntiles, chunk_, d_m, s);
    cv_compact_kernel<<<ntiles, kThreads, 0, s>>>(
        T, rows, cols, B, tile_cnt_.data(), keys_[0].data(), srcs_[0].data());

    // 4. stable LSD radix sort (matrix_converter.inl:1610-1616)
    const int bits = key_bits_for((uint64_t)B * (uint64_t)B - 1ull);
    int       cur  = 0;
    for(int shift = 0; shift < bits; shift += 8)
    {
        int      w     = (bits - shift) < 8 ? (bits - shift) : 8;
        uint32_t dmask = (1u << w) - 1u;
        rs_hist_kernel<<<ntiles, kThreads, 0, s>>>(d_m, keys_[cur].data(), shift, dmask, hist_.data(), ntiles);
        scan_exclusive(hist_.data(), (int64_t)256 * ntiles, chunk_, nullptr, s);
        rs_scatter_kernel<<<ntiles, kThreads, 0, s>>>(d_m,
                                                      keys_[cur].data(),
                                                      srcs_[cur].data(),
                                                      keys_[cur ^ 1].data(),
                                                      srcs_[cur ^ 1].data(),
                                                      shift,
                                                      dmask,
                                                      hist_.data(),
                                                      ntiles);
        cur ^= 1;
    }

    // 5-6. run-length encode (1626-1650); the one host read-back sizes the outputs
    rle_count_kernel<<<ntiles, kThreads, 0, s>>>(d_m, keys_[cur].data(), tile_cnt_.data());
    scan_exclusive(tile_cnt_.data(), ntiles, chunk_, d_u, s);
    MI_HIP(hipMemcpyAsync(pinned_, counts_.data(), 2 * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    const int64_t U = pinned_[1];
    out.U           = U;
    out.rows.resize((size_t)(U > 0 ? U : 1), ratio);
    out.cols.resize((size_t)(U > 0 ? U : 1), ratio);
    if((int64_t)out.rows.cap > out.stride || out.stride < U)
    {
        out.stride = (int64_t)out.rows.cap;
    }
    out.vals.resize((size_t)9 * out.stride);
    if(want32)
    {
        out.vals32.resize((size_t)9 * out.stride);
        out.has32 = true;
    }
    seg_start_.resize((size_t)(U > 0 ? U : 1), ratio);
    rle_write_kernel<<<ntiles, kThreads, 0, s>>>(
        d_m, keys_[cur].data(), tile_cnt_.data(), B, out.rows.data(), out.cols.data(), seg_start_.data());

    // 7-8. segmented reduce over the sorted permutation (1652-1698)
    const int grid = std::max(1, std::min(ceil_div(U, kThreads), device_cu_count() * 16));
    if(U > 0)
        seg_reduce_kernel<<<grid, kThreads, 0, s>>>(U,
                                                    d_m,
                                                    seg_start_.data(),
                                                    srcs_[cur].data(),
                                                    vals,
                                                    out.vals.data(),
                                                    want32 ? out.vals32.data() : nullptr,
                                                    out.stride);
    row_ptr_kernel<<<std::max(1, std::min(ceil_div(U + 1, kThreads), device_cu_count() * 16)), kThreads, 0, s>>>(
        U, out.rows.data(), out.row_ptr.data(), B);
    MI_HIP(hipGetLastError());
}
}  // namespace mi
