// Triplet -> sorted upper-triangular 3x3 BCOO + BSR row offsets on gfx950,
// plus the full (both triangles) BSR the SpMV streams.
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
//   * one host read-back (the unique count, to size the outputs) instead of two;
//   * the upper BCOO is then transposed (the same radix sort on col * B + row
//     of the off-diagonal blocks) and merged row by row into a full BSR, so the
//     SpMV is a pure row gather without atomics.
#include "linsys.hpp"

namespace mi
{
namespace
{
constexpr int kThreads = 256;
constexpr int kItems   = 16;
constexpr int kTile    = kThreads * kItems;  // 4096 keys per tile
constexpr int kWaves   = kThreads / kWave;

// kept entries and their sort keys.  Mode 0: the triplet filter of
// matrix_converter.inl:142 (plus a bound on col, which the reference leaves
// to its producers), key (row, col).  Mode 1: the off-diagonal blocks of the
// upper BCOO, key (col, row) (the transpose).
template <int MODE>
__device__ __forceinline__ bool keep_entry(int32_t r, int32_t c, int32_t B)
{
    if constexpr(MODE == 0)
        return r >= 0 && r <= c && c < B;
    else
        return r != c;
}
template <int MODE>
__device__ __forceinline__ uint64_t entry_key(int32_t r, int32_t c, const KeyCodec& kc)
{
    if constexpr(MODE == 0)
        return kc.up(r, c);
    else
        return kc.lo(r, c);
}

// ---- 1. count kept entries per tile
// (and, with d_band, the largest col - row of the kept entries)
template <int MODE>
__global__ __launch_bounds__(kThreads) void cv_count_kernel(int64_t        T,
                                                           const int32_t* rows,
                                                           const int32_t* cols,
                                                           int32_t        B,
                                                           uint32_t*      tile_cnt,
                                                           int32_t*       d_band)
{
    __shared__ uint32_t lds[kWaves];
    __shared__ int32_t  bmax[kWaves];
    const int64_t       base = (int64_t)blockIdx.x * kTile;
    uint32_t            c    = 0;
    int32_t             band = 0;
#pragma unroll
    for(int k = 0; k < kItems; ++k)
    {
        int64_t i = base + (int64_t)k * kThreads + threadIdx.x;
        if(i < T)
        {
            const int32_t r = rows[i], cc = cols[i];
            if(keep_entry<MODE>(r, cc, B))
            {
                ++c;
                band = max(band, cc - r);
            }
        }
    }
    uint32_t s = block_sum<kThreads>(c, lds);
    if(threadIdx.x == 0)
        tile_cnt[blockIdx.x] = s;
    if(d_band)
    {
#pragma unroll
        for(int o = kWave / 2; o > 0; o >>= 1)
            band = max(band, __shfl_xor(band, o, kWave));
        if(threadIdx.x % kWave == 0)
            bmax[threadIdx.x / kWave] = band;
        __syncthreads();
        if(threadIdx.x == 0)
        {
            int32_t m = bmax[0];
            for(int w = 1; w < kWaves; ++w)
                m = max(m, bmax[w]);
            atomicMax(d_band, m);
        }
    }
}

// ---- single-workgroup exclusive scan (n up to a few 100k), total -> *total
__global__ __launch_bounds__(1024) void scan_small_kernel(uint32_t* data, int64_t n, int64_t* total)
{
    __shared__ uint32_t lds[1024 / kWave + 1];
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
    if(threadIdx.x == 0 && total)
        *total = (int64_t)carry;
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
    __shared__ uint32_t sv[kScanChunk];
    __shared__ uint32_t lds[kWaves + 1];
    const int64_t       base = (int64_t)blockIdx.x * kScanChunk;
#pragma unroll
    for(int k = 0; k < kItems; ++k)
    {
        int     t = k * kThreads + threadIdx.x;
        int64_t i = base + t;
        sv[t]     = i < n ? data[i] : 0u;
    }
    __syncthreads();
    const int first = threadIdx.x * kItems;
    uint32_t  c     = 0;
#pragma unroll
    for(int k = 0; k < kItems; ++k)
        c += sv[first + k];
    uint32_t o = block_exclusive_scan<kThreads>(c, lds, (uint32_t*)nullptr) + sums[blockIdx.x];
#pragma unroll
    for(int k = 0; k < kItems; ++k)
    {
        uint32_t v      = sv[first + k];
        sv[first + k] = o;
        o += v;
    }
    __syncthreads();
#pragma unroll
    for(int k = 0; k < kItems; ++k)
    {
        int     t = k * kThreads + threadIdx.x;
        int64_t i = base + t;
        if(i < n)
            data[i] = sv[t];
    }
}

// ---- radix sort inputs.  PairSource: the (key, source) pairs of the last
// pass, m of them.  TripletSource: the first pass straight from the
// triplets' (row, col) -- entries that keep_entry drops are skipped and the
// key is built on the fly, so the compaction writes no pairs of its own (a
// stable pass over the kept entries in index order, as compacting first
// would give).
template <typename KeyT>
struct PairSource
{
    const int64_t* d_m;
    const KeyT*    keys;
    const int32_t* srcs;
    __device__ int64_t count() const { return *d_m; }
    __device__ bool    key_at(int64_t i, int64_t m, KeyT& k) const
    {
        if(i >= m)
            return false;
        k = keys[i];
        return true;
    }
    __device__ bool pair_at(int64_t i, int64_t m, KeyT& k, int32_t& v) const
    {
        if(i >= m)
            return false;
        k = keys[i];
        v = srcs[i];
        return true;
    }
};
template <int MODE, typename KeyT>
struct TripletSource
{
    const int32_t* rows;
    const int32_t* cols;
    int64_t        n;
    int32_t        B;
    KeyCodec       kc;
    __device__ int64_t count() const { return n; }
    __device__ bool    key_at(int64_t i, int64_t m, KeyT& k) const
    {
        if(i >= m)
            return false;
        const int32_t r = rows[i], c = cols[i];
        if(!keep_entry<MODE>(r, c, B))
            return false;
        k = (KeyT)entry_key<MODE>(r, c, kc);
        return true;
    }
    __device__ bool pair_at(int64_t i, int64_t m, KeyT& k, int32_t& v) const
    {
        v = (int32_t)i;
        return key_at(i, m, k);
    }
};

// ---- radix sort: tile histogram (digit-major layout hist[d * ntiles + tile])
template <typename KeyT, typename Src>
__global__ __launch_bounds__(kThreads) void rs_hist_kernel(Src           src,
                                                          int           shift,
                                                          uint32_t      dmask,
                                                          uint32_t*     hist,
                                                          int           ntiles)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t m    = src.count();
    const int64_t base = (int64_t)blockIdx.x * kTile;
#pragma unroll
    for(int k = 0; k < kItems; ++k)
    {
        const int64_t i = base + (int64_t)k * kThreads + threadIdx.x;
        KeyT          key;
        if(src.key_at(i, m, key))
            atomicAdd(&h[(uint32_t)(key >> shift) & dmask], 1u);
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
template <typename KeyT, typename Src>
__global__ __launch_bounds__(kThreads) void rs_scatter_kernel(Src             src,
                                                             KeyT*           kout,
                                                             int32_t*        vout,
                                                             int             shift,
                                                             uint32_t        dmask,
                                                             const uint32_t* hist_scan,
                                                             int             ntiles)
{
    __shared__ KeyT     sk[kTile];
    __shared__ int32_t  sv[kTile];
    __shared__ uint32_t wcnt[kWaves][256];
    __shared__ uint32_t dstart[256];
    __shared__ uint32_t gbase[256];
    __shared__ uint32_t lds[kWaves + 1];

    const int64_t m    = src.count();
    const int64_t base = (int64_t)blockIdx.x * kTile;
    if(base >= m)
        return;  // uniform per block
    const int w    = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
#pragma unroll
    for(int q = 0; q < kWaves; ++q)
        wcnt[q][threadIdx.x] = 0;
    __syncthreads();

    KeyT           key[kItems];
    int32_t        val[kItems];
    uint32_t       rank[kItems];
    uint32_t       dig[kItems];
    const uint64_t lt = lanemask_lt64();
#pragma unroll
    for(int k = 0; k < kItems; ++k)
    {
        int64_t  i     = base + (int64_t)w * (kWave * kItems) + (int64_t)k * kWave + lane;
        key[k]         = KeyT(0);
        val[k]         = 0;
        const bool valid = src.pair_at(i, m, key[k], val[k]);
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
    uint32_t tile_total;  // the tile's kept entries
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
        uint32_t e = block_exclusive_scan<kThreads>(acc, lds, &tile_total);
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
    const int64_t cnt = tile_total;
#pragma unroll
    for(int k = 0; k < kItems; ++k)
    {
        int i = k * kThreads + threadIdx.x;
        if(i < cnt)
        {
            KeyT     kk  = sk[i];
            uint32_t d   = (uint32_t)(kk >> shift) & dmask;
            uint32_t out = gbase[d] + ((uint32_t)i - dstart[d]);
            kout[out]    = kk;
            vout[out]    = sv[i];
        }
    }
}

// stage one tile of sorted keys into LDS (plus the key before the tile)
template <typename KeyT>
__device__ __forceinline__ void stage_keys(const KeyT* keys, int64_t m, int64_t base, KeyT* sk)
{
#pragma unroll
    for(int k = 0; k < kItems; ++k)
    {
        int     t = k * kThreads + threadIdx.x;
        int64_t i = base + t;
        sk[t + 1] = i < m ? keys[i] : ~KeyT(0);
    }
    if(threadIdx.x == 0)
        sk[0] = base > 0 ? keys[base - 1] : ~KeyT(0);
    __syncthreads();
}

// ---- run-length encode: heads per tile
template <typename KeyT>
__global__ __launch_bounds__(kThreads) void rle_count_kernel(const int64_t* d_m, const KeyT* keys, uint32_t* tile_cnt)
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

template <typename KeyT>
__global__ __launch_bounds__(kThreads) void rle_write_kernel(const int64_t*  d_m,
                                                            const KeyT*     keys,
                                                            const uint32_t* tile_off,
                                                            KeyCodec        kc,
                                                            int32_t*        out_rows,
                                                            int32_t*        out_cols,
                                                            int32_t*        seg_start)
{
    __shared__ KeyT     sk[kTile + 1];
    __shared__ int32_t  hp[kTile];
    __shared__ uint32_t lds[kWaves + 1];
    const int64_t       m    = *d_m;
    const int64_t       base = (int64_t)blockIdx.x * kTile;
    stage_keys(keys, m, base, sk);
    const int first = threadIdx.x * kItems;
    uint32_t  flags = 0, c = 0;
    KeyT      kk[kItems];
#pragma unroll
    for(int k = 0; k < kItems; ++k)
    {
        int64_t i = base + first + k;
        kk[k]     = sk[first + k + 1];
        if(i < m && (i == 0 || kk[k] != sk[first + k]))
        {
            flags |= 1u << k;
            ++c;
        }
    }
    uint32_t heads = 0;
    uint32_t o     = block_exclusive_scan<kThreads>(c, lds, &heads);  // its barriers end every read of sk
    // the heads' keys (in place, in order) and positions, then written out by
    // consecutive lanes
#pragma unroll
    for(int k = 0; k < kItems; ++k)
        if(flags & (1u << k))
        {
            sk[o] = kk[k];
            hp[o] = first + k;
            ++o;
        }
    __syncthreads();
    const uint32_t g = tile_off[blockIdx.x];
    for(uint32_t i = threadIdx.x; i < heads; i += kThreads)
    {
        int32_t r, c;
        kc.up_rc((uint64_t)sk[i], r, c);
        out_rows[g + i]    = r;
        out_cols[g + i]    = c;
        seg_start[g + i]   = (int32_t)(base + hp[i]);
    }
}

// one 72-byte FP64 triplet record (index src) by five 16-byte buffer loads
// from the 16-byte boundary at or below it (odd records start 8 bytes past
// one); the buffer's range check returns zeros past the array's end
__device__ __forceinline__ void load_record64(__amdgpu_buffer_rsrc_t rs, int32_t src, double (&v)[9])
{
    typedef double d2v __attribute__((ext_vector_type(2)));
    const uint32_t off = 72u * (uint32_t)src;
    const uint32_t b   = off & ~15u;
    const d2v      t0  = __builtin_bit_cast(d2v, __builtin_amdgcn_raw_buffer_load_b128(rs, b, 0, 0));
    const d2v      t1  = __builtin_bit_cast(d2v, __builtin_amdgcn_raw_buffer_load_b128(rs, b + 16u, 0, 0));
    const d2v      t2  = __builtin_bit_cast(d2v, __builtin_amdgcn_raw_buffer_load_b128(rs, b + 32u, 0, 0));
    const d2v      t3  = __builtin_bit_cast(d2v, __builtin_amdgcn_raw_buffer_load_b128(rs, b + 48u, 0, 0));
    const d2v      t4  = __builtin_bit_cast(d2v, __builtin_amdgcn_raw_buffer_load_b128(rs, b + 64u, 0, 0));
    const bool     odd = (off & 8u) != 0;
    v[0]               = odd ? t0.y : t0.x;
    v[1]               = odd ? t1.x : t0.y;
    v[2]               = odd ? t1.y : t1.x;
    v[3]               = odd ? t2.x : t1.y;
    v[4]               = odd ? t2.y : t2.x;
    v[5]               = odd ? t3.x : t2.y;
    v[6]               = odd ? t3.y : t3.x;
    v[7]               = odd ? t4.x : t3.y;
    v[8]               = odd ? t4.y : t4.x;
}

// ---- segmented reduce: one thread per unique block, duplicates summed in
// FP64 in ascending source order (bit-identical to a sequential left fold of
// the values as doubles).  The gathers of a group of 4 are issued before
// their (ordered) additions.  TV: the triplet values' type;
// kWide: FP64 records by 16-byte loads (T records, 72 T < 2^32).
template <typename TV, bool kWide = false>
__global__ __launch_bounds__(kThreads) void seg_reduce_kernel(int64_t        U,
                                                             const int64_t* d_m,
                                                             const int32_t* seg_start,
                                                             const int32_t* srcs,
                                                             const TV*      vals,
                                                             double*        out,
                                                             int64_t        T = 0)
{
    const int64_t m = *d_m;
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc((void*)vals, 0, kWide ? (int)(uint32_t)(72 * T) : 0, 0x00020000);
    auto load = [&](int32_t src, double (&v)[9]) {
        if constexpr(kWide)
            load_record64(rs, src, v);
        else
        {
            const TV* p = vals + 9 * (int64_t)src;
#pragma unroll
            for(int q = 0; q < 9; ++q)
                v[q] = (double)p[q];
        }
    };
    for(int64_t u = (int64_t)blockIdx.x * kThreads + threadIdx.x; u < U; u += (int64_t)gridDim.x * kThreads)
    {
        const int64_t b = seg_start[u];
        const int64_t e = (u + 1 < U) ? (int64_t)seg_start[u + 1] : m;
        double        acc[9];
        load(srcs[b], acc);
        int64_t k = b + 1;
        for(; k + 4 <= e; k += 4)
        {
            double t[4][9];
#pragma unroll
            for(int g = 0; g < 4; ++g)
                load(srcs[k + g], t[g]);
#pragma unroll
            for(int g = 0; g < 4; ++g)
#pragma unroll
                for(int q = 0; q < 9; ++q)
                    acc[q] = acc[q] + t[g][q];
        }
        for(; k < e; ++k)
        {
            double t[9];
            load(srcs[k], t);
#pragma unroll
            for(int q = 0; q < 9; ++q)
                acc[q] = acc[q] + t[q];
        }
#pragma unroll
        for(int q = 0; q < 9; ++q)
            out[upper_at(u, q)] = acc[q];
    }
}

// ---- BSR row offsets from sorted rows: row_ptr[r] = first u with rows[u] >= r
__global__ __launch_bounds__(kThreads) void row_ptr_kernel(const int64_t* d_n, int64_t nmax, const int32_t* rows, int32_t* row_ptr, int32_t B)
{
    const int64_t U = d_n ? *d_n : nmax;
    for(int64_t u = (int64_t)blockIdx.x * kThreads + threadIdx.x; u <= U; u += (int64_t)gridDim.x * kThreads)
    {
        int32_t prev = (u == 0) ? -1 : rows[u - 1];
        int32_t cur  = (u == U) ? B : rows[u];
        for(int32_t r = prev + 1; r <= cur; ++r)
            row_ptr[r] = (int32_t)u;
    }
}

// ---- full BSR: lower part (transposed upper blocks, sorted by (col, row)) then
// upper part of each row
template <typename KeyT>
__global__ void unpack_lower_kernel(const int64_t* d_n, const KeyT* keys, KeyCodec kc, int32_t* lrow)
{
    const int64_t n = *d_n;
    for(int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    {
        int32_t lr, lc;
        kc.lo_rc((uint64_t)keys[k], lr, lc);
        lrow[k] = lr;
    }
}

// chunk_row[c] = the first row whose first full-BSR block is at or after
// c kChunkStep (binary search in full_ptr); chunk_row[nchunks] = B
__global__ void chunk_row_kernel(const int32_t* __restrict__ fptr, int32_t B, int32_t nchunks, int32_t* chunk_row)
{
    for(int32_t c = blockIdx.x * blockDim.x + threadIdx.x; c <= nchunks; c += gridDim.x * blockDim.x)
    {
        if(c == nchunks)
        {
            chunk_row[c] = B;
            continue;
        }
        const int64_t target = (int64_t)c * kChunkStep;
        int32_t       lo = 0, hi = B;  // answer in [0, B]
        while(lo < hi)
        {
            const int32_t mid = lo + (hi - lo) / 2;
            if((int64_t)fptr[mid] >= target)
                hi = mid;
            else
                lo = mid + 1;
        }
        chunk_row[c] = lo;
    }
}

// blocks per full-BSR row (lower + upper; an empty row gets one zero diagonal
// block, so every row holds at least one), scanned into full_ptr afterwards
__global__ void full_count_kernel(const int32_t* lower_ptr, const int32_t* upper_ptr, uint32_t* cnt, int32_t B)
{
    for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= B; i += (int64_t)gridDim.x * blockDim.x)
    {
        if(i == B)
        {
            cnt[i] = 0;
            continue;
        }
        const int32_t c = (lower_ptr[i + 1] - lower_ptr[i]) + (upper_ptr[i + 1] - upper_ptr[i]);
        cnt[i]          = c > 0 ? (uint32_t)c : 1u;
    }
}

template <typename VT>
__global__ void full_empty_kernel(const int32_t* lower_ptr,
                                  const int32_t* upper_ptr,
                                  const int32_t* full_ptr,
                                  int32_t        B,
                                  int32_t*       fcols,
                                  VT*            fvals,
                                  int64_t        fstride)
{
    for(int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < B; i += gridDim.x * blockDim.x)
        if(lower_ptr[i + 1] == lower_ptr[i] && upper_ptr[i + 1] == upper_ptr[i])
        {
            const int64_t d = full_ptr[i];
            fcols[d]        = i;
#pragma unroll
            for(int q = 0; q < 9; ++q)
                fvals[q * fstride + d] = VT(0);
        }
}

template <typename VT, typename KeyT>
__global__ void full_lower_kernel(const int64_t*  d_n,
                                  const KeyT*     keys,
                                  KeyCodec        kc,
                                  const int32_t*  src,
                                  const int32_t*  lower_ptr,
                                  const int32_t*  full_ptr,
                                  const double*   uvals,
                                  int32_t         B,
                                  int32_t*        fcols,
                                  VT*             fvals,
                                  int64_t         fstride)
{
    const int64_t n = *d_n;
    for(int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    {
        int32_t i, j;  // lower row = upper col, lower col = upper row
        kc.lo_rc((uint64_t)keys[k], i, j);
        const int64_t  u   = src[k];
        const int64_t  d   = (int64_t)full_ptr[i] + (k - lower_ptr[i]);
        fcols[d]           = j;
#pragma unroll
        for(int r = 0; r < 3; ++r)
#pragma unroll
            for(int c = 0; c < 3; ++c)
                fvals[(r + 3 * c) * fstride + d] = (VT)uvals[upper_at(u, c + 3 * r)];
    }
}

template <typename VT>
__global__ void full_upper_kernel(int64_t        U,
                                  const int32_t* rows,
                                  const int32_t* cols,
                                  const int32_t* upper_ptr,
                                  const int32_t* lower_ptr,
                                  const int32_t* full_ptr,
                                  const double*  uvals,
                                  int32_t*       fcols,
                                  VT*            fvals,
                                  int64_t        fstride)
{
    for(int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < U; u += (int64_t)gridDim.x * blockDim.x)
    {
        const int32_t i = rows[u];
        const int64_t d = (int64_t)full_ptr[i] + (lower_ptr[i + 1] - lower_ptr[i]) + (u - upper_ptr[i]);
        fcols[d]        = cols[u];
#pragma unroll
        for(int q = 0; q < 9; ++q)
            fvals[q * fstride + d] = (VT)uvals[upper_at(u, q)];
    }
}

// ---- window layout (BsrMatrix::w*): blocks per row position; a row holds its
// upper blocks and the lower blocks whose other row is in another window.
// 16 lanes per row walk the row's lower entries together (coalesced reads of
// the transposed keys; one lane per row read them 16 times over).
constexpr int kRowLanes = 16;

template <typename KeyT>
__global__ __launch_bounds__(kThreads) void win_count_kernel(int32_t        B,
                                                            int            shift,
                                                            const int32_t* __restrict__ pos,
                                                            const int32_t* __restrict__ real,
                                                            const int32_t* __restrict__ row_ptr,
                                                            const int32_t* __restrict__ lower_ptr,
                                                            const KeyT* __restrict__ lkeys,
                                                            KeyCodec       kc,
                                                            uint32_t*      cnt)
{
    const int64_t groups = (int64_t)gridDim.x * (kThreads / kRowLanes);
    const int     lane   = threadIdx.x % kRowLanes;
    for(int64_t p = ((int64_t)blockIdx.x * kThreads + threadIdx.x) / kRowLanes; p <= B; p += groups)
    {
        if(p == B)
        {
            if(lane == 0)
                cnt[p] = 0;
            continue;
        }
        const int32_t v  = real[p];
        const int32_t ch = (int32_t)p >> shift;
        uint32_t      c  = 0;
        for(int32_t k = lower_ptr[v] + lane; k < lower_ptr[v + 1]; k += kRowLanes)
        {
            int32_t lr, i;  // i: the upper row
            kc.lo_rc((uint64_t)lkeys[k], lr, i);
            c += (pos[i] >> shift) != ch ? 1u : 0u;
        }
#pragma unroll
        for(int o = kRowLanes / 2; o > 0; o >>= 1)
            c += __shfl_xor(c, o, kWave);
        if(lane == 0)
            cnt[p] = c + (uint32_t)(row_ptr[v + 1] - row_ptr[v]);
    }
}

__host__ __device__ __forceinline__ int32_t win_inner_word(int32_t li, int32_t lj)
{
    return (int32_t)(0x80000000u | ((uint32_t)li << 15) | (uint32_t)lj);
}

template <typename VT>
__global__ void win_upper_kernel(int64_t        U,
                                 int            shift,
                                 const int32_t* __restrict__ rows,
                                 const int32_t* __restrict__ cols,
                                 const int32_t* __restrict__ row_ptr,
                                 const int32_t* __restrict__ pos,
                                 const int32_t* __restrict__ wptr,
                                 const double* __restrict__ uvals,
                                 int32_t*       words,
                                 VT*            wvals,
                                 int64_t        stride)
{
    const int32_t mask = (1 << shift) - 1;
    for(int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < U; u += (int64_t)gridDim.x * blockDim.x)
    {
        const int32_t v = rows[u], j = cols[u];
        const int32_t p = pos[v], pj = pos[j];
        const int64_t d = (int64_t)wptr[p] + (u - row_ptr[v]);
        words[d]        = (p >> shift) == (pj >> shift) ? win_inner_word(p & mask, pj & mask) : pj;
#pragma unroll
        for(int q = 0; q < 9; ++q)
            wvals[q * stride + d] = (VT)uvals[upper_at(u, q)];
    }
}

template <typename VT, typename KeyT>
__global__ __launch_bounds__(kThreads) void win_lower_kernel(int32_t        B,
                                                            int            shift,
                                                            const int32_t* __restrict__ pos,
                                                            const int32_t* __restrict__ real,
                                                            const int32_t* __restrict__ row_ptr,
                                                            const int32_t* __restrict__ lower_ptr,
                                                            const KeyT* __restrict__ lkeys,
                                                            KeyCodec       kc,
                                                            const int32_t* __restrict__ lsrc,
                                                            const int32_t* __restrict__ wptr,
                                                            const double* __restrict__ uvals,
                                                            int32_t*       words,
                                                            VT*            wvals,
                                                            int64_t        stride)
{
    const int64_t groups = (int64_t)gridDim.x * (kThreads / kRowLanes);
    const int     lane   = threadIdx.x % kRowLanes;
    const int     gsh    = (int)(threadIdx.x % kWave) & ~(kRowLanes - 1);  // the group's first lane in the wave
    // the group's rows run in lockstep: every group of a wave takes the same
    // number of rounds (the wave-wide ballots below need all 64 lanes)
    const int64_t first  = ((int64_t)blockIdx.x * kThreads + threadIdx.x) / kRowLanes;
    const int64_t rounds = (B + groups - 1) / groups;
    for(int64_t it = 0; it < rounds; ++it)
    {
        const int64_t p    = first + it * groups;
        const bool    live = p < B;
        const int32_t v    = live ? real[p] : 0;
        const int32_t ch   = (int32_t)p >> shift;
        const int32_t lo = live ? lower_ptr[v] : 0, hi = live ? lower_ptr[v + 1] : 0;
        int64_t       d  = live ? (int64_t)wptr[p] + (row_ptr[v + 1] - row_ptr[v]) : 0;
        // the wave's longest lower list bounds its rounds
        int32_t len = hi - lo;
#pragma unroll
        for(int o = kWave / 2; o > 0; o >>= 1)
            len = max(len, __shfl_xor(len, o, kWave));
        for(int32_t b = 0; b < len; b += kRowLanes)
        {
            const int32_t k   = lo + b + lane;
            int32_t       i   = 0, pi = 0;
            bool          out = false;
            if(k < hi)
            {
                int32_t lr;
                kc.lo_rc((uint64_t)lkeys[k], lr, i);
                pi  = pos[i];
                out = (pi >> shift) != ch;  // an inner block is stored once, with row i
            }
            const uint64_t  bal  = __ballot(out);
            const uint32_t  mine = (uint32_t)(bal >> gsh) & ((1u << kRowLanes) - 1u);
            if(out)
            {
                const int64_t dk = d + __popc(mine & ((1u << lane) - 1u));
                const int64_t u  = lsrc[k];
                words[dk]        = pi;
#pragma unroll
                for(int r = 0; r < 3; ++r)
#pragma unroll
                    for(int c = 0; c < 3; ++c)
                        wvals[(r + 3 * c) * stride + dk] = (VT)uvals[upper_at(u, c + 3 * r)];
            }
            d += __popc(mine);
        }
    }
}

// lowest row and highest row + 1 of every window (batched solves skip the
// windows that hold no row of a running system): one workgroup per window
__global__ __launch_bounds__(kThreads) void win_range_kernel(int32_t B, int shift, int32_t nchunks,
                                                            const int32_t* __restrict__ real, int32_t* lo,
                                                            int32_t* hi)
{
    __shared__ int32_t smin[kThreads / kWave], smax[kThreads / kWave];
    for(int32_t c = blockIdx.x; c < nchunks; c += gridDim.x)
    {
        const int32_t p0 = c << shift, p1 = (int32_t)std::min<int64_t>((int64_t)(c + 1) << shift, B);
        int32_t       a = B, b = -1;
        for(int32_t p = p0 + (int32_t)threadIdx.x; p < p1; p += kThreads)
        {
            a = std::min(a, real[p]);
            b = std::max(b, real[p]);
        }
#pragma unroll
        for(int o = kWave / 2; o > 0; o >>= 1)
        {
            a = std::min(a, __shfl_xor(a, o, kWave));
            b = std::max(b, __shfl_xor(b, o, kWave));
        }
        if(threadIdx.x % kWave == 0)
        {
            smin[threadIdx.x / kWave] = a;
            smax[threadIdx.x / kWave] = b;
        }
        __syncthreads();
        if(threadIdx.x == 0)
        {
            for(int w = 1; w < kThreads / kWave; ++w)
            {
                a = std::min(a, smin[w]);
                b = std::max(b, smax[w]);
            }
            lo[c] = a;
            hi[c] = b + 1;
        }
        __syncthreads();
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

int grid_of(int64_t n)
{
    return std::max(1, std::min(ceil_div(n, kThreads), device_cu_count() * 16));
}
}  // namespace

MatrixConverter::~MatrixConverter()
{
    if(pinned_)
        (void)hipHostFree(pinned_);
}

template <typename KeyT, typename First>
int radix_sort_from(First           first,
                    KeyT*           k0,
                    int32_t*        v0,
                    KeyT*           k1,
                    int32_t*        v1,
                    const int64_t*  d_n,
                    int64_t         nmax,
                    int             bits,
                    DBuf<uint32_t>& hist,
                    DBuf<uint32_t>& chunk,
                    hipStream_t     s);

// compaction + stable radix sort of (key, source) pairs; returns the buffer
// index holding the result, *d_count = number of kept entries (device)
template <int MODE, typename KeyT>
int MatrixConverter::compact_sort(int64_t        n,
                                  const int32_t* rows,
                                  const int32_t* cols,
                                  int32_t        B,
                                  int64_t*       d_count,
                                  double         ratio,
                                  hipStream_t    s)
{
    const int ntiles = ceil_div(n, kTile);
    tile_cnt_.resize((size_t)ntiles + 1, ratio);
    keys_[0].resize((size_t)n, ratio);
    keys_[1].resize((size_t)n, ratio);
    srcs_[0].resize((size_t)n, ratio);
    srcs_[1].resize((size_t)n, ratio);
    hist_.resize((size_t)256 * ntiles, ratio);
    if(!(MODE == 0 && counted_))
        cv_count_kernel<MODE><<<ntiles, kThreads, 0, s>>>(n, rows, cols, B, tile_cnt_.data(), nullptr);
    counted_ = false;
    scan_exclusive(tile_cnt_.data(), ntiles, chunk_, d_count, s);
    KeyT* k0 = reinterpret_cast<KeyT*>(keys_[0].data());
    KeyT* k1 = reinterpret_cast<KeyT*>(keys_[1].data());
    // compaction (the kept count, *d_count, from the tile counts) and a stable
    // LSD radix sort (matrix_converter.inl:1579-1616) on the key's bits, whose
    // first pass reads the triplets' (row, col) and skips the dropped entries
    TripletSource<MODE, KeyT> first{rows, cols, n, B, codec_};
    return radix_sort_from<KeyT>(first, k0, srcs_[0].data(), k1, srcs_[1].data(), d_count, n, key_bits_, hist_,
                                 chunk_, s);
}

void MatrixConverter::build_window(BsrMatrix& A, int shape, bool want32, double ratio, hipStream_t s)
{
    MI_REQUIRE(shape >= 0 && shape < spmv_shape_count(), MI_ERR_ARG, "window layout: unknown kernel shape");
    const int shift = spmv_shape_shift(shape);
    MI_REQUIRE(shift >= 4 && shift <= kWinShiftMax, MI_ERR_ARG, "window layout: window size out of range");
    A.wshape = shape;
    const int32_t B = A.B;
    MI_REQUIRE(A.wpos.data() && A.wreal.data() && A.wpos.size() >= (size_t)B && A.wreal.size() >= (size_t)B,
               MI_ERR_STATE, "window layout: no row order");
    A.wshift   = shift;
    A.wnchunks = (int32_t)(((int64_t)B + (1 << shift) - 1) >> shift);
    A.wptr.resize((size_t)B + 1, ratio);
    A.wchunk_lo.resize((size_t)std::max(1, A.wnchunks), ratio);
    A.wchunk_hi.resize((size_t)std::max(1, A.wnchunks), ratio);
    A.has_win = false;
    if(B == 0)
    {
        A.wblocks = 0;
        A.has_win = true;
        return;
    }
    const int64_t U = A.U;
    uint32_t*     cnt = reinterpret_cast<uint32_t*>(A.wptr.data());  // scanned in place
    const bool narrow = codec_.bb > 0;
    if(narrow)
        win_count_kernel<uint32_t><<<grid_of(((int64_t)B + 1) * kRowLanes), kThreads, 0, s>>>(
            B, shift, A.wpos.data(), A.wreal.data(), A.row_ptr.data(), lower_ptr_.data(),
            U > 0 ? key_ptr<uint32_t>(tcur_) : nullptr, codec(), cnt);
    else
        win_count_kernel<uint64_t><<<grid_of(((int64_t)B + 1) * kRowLanes), kThreads, 0, s>>>(
            B, shift, A.wpos.data(), A.wreal.data(), A.row_ptr.data(), lower_ptr_.data(),
            U > 0 ? key_ptr<uint64_t>(tcur_) : nullptr, codec(), cnt);
    scan_exclusive(cnt, (int64_t)B + 1, chunk_, counts_.data() + 4, s);
    win_range_kernel<<<std::max(1, std::min(A.wnchunks, device_cu_count() * 8)), kThreads, 0, s>>>(
        B, shift, A.wnchunks, A.wreal.data(), A.wchunk_lo.data(), A.wchunk_hi.data());
    // the window layout holds at most the full BSR's 2U blocks; four blocks of
    // padding past the last, and a plane stride that is a multiple of 4 (the
    // SpMV reads up to four consecutive blocks per load)
    const int64_t wmax = 2 * U;
    A.wwords.resize((size_t)wmax + 4, ratio);
    A.wstride = ((int64_t)A.wwords.cap + 3) & ~(int64_t)3;
    auto win_lower = [&](auto* wv) {
        using VT = std::remove_pointer_t<decltype(wv)>;
        if(narrow)
            win_lower_kernel<VT, uint32_t><<<grid_of((int64_t)B * kRowLanes), kThreads, 0, s>>>(
                B, shift, A.wpos.data(), A.wreal.data(), A.row_ptr.data(), lower_ptr_.data(),
                key_ptr<uint32_t>(tcur_), codec(), srcs_[tcur_].data(), A.wptr.data(), A.vals.data(),
                A.wwords.data(), wv, A.wstride);
        else
            win_lower_kernel<VT, uint64_t><<<grid_of((int64_t)B * kRowLanes), kThreads, 0, s>>>(
                B, shift, A.wpos.data(), A.wreal.data(), A.row_ptr.data(), lower_ptr_.data(),
                key_ptr<uint64_t>(tcur_), codec(), srcs_[tcur_].data(), A.wptr.data(), A.vals.data(),
                A.wwords.data(), wv, A.wstride);
    };
    if(U > 0)
    {
        if(want32)
        {
            A.wvals32.resize((size_t)9 * A.wstride);
            win_upper_kernel<float><<<grid_of(U), kThreads, 0, s>>>(U, shift, A.rows.data(), A.cols.data(),
                                                                    A.row_ptr.data(), A.wpos.data(), A.wptr.data(),
                                                                    A.vals.data(), A.wwords.data(), A.wvals32.data(),
                                                                    A.wstride);
            win_lower(A.wvals32.data());
        }
        else
        {
            A.wvals.resize((size_t)9 * A.wstride);
            win_upper_kernel<double><<<grid_of(U), kThreads, 0, s>>>(U, shift, A.rows.data(), A.cols.data(),
                                                                     A.row_ptr.data(), A.wpos.data(), A.wptr.data(),
                                                                     A.vals.data(), A.wwords.data(), A.wvals.data(),
                                                                     A.wstride);
            win_lower(A.wvals.data());
        }
    }
    else if(want32)
        A.wvals32.resize((size_t)9 * A.wstride);
    else
        A.wvals.resize((size_t)9 * A.wstride);
    MI_HIP(hipGetLastError());
    A.wblocks = -1;  // read back lazily (window_blocks)
    A.has_win = true;
}

// stable LSD passes of 8 bits; the first reads `first` (pairs or triplets),
// the rest ping-pong between the buffers; returns the buffer holding the result
template <typename KeyT, typename First>
int radix_sort_from(First           first,
                    KeyT*           k0,
                    int32_t*        v0,
                    KeyT*           k1,
                    int32_t*        v1,
                    const int64_t*  d_n,
                    int64_t         nmax,
                    int             bits,
                    DBuf<uint32_t>& hist,
                    DBuf<uint32_t>& chunk,
                    hipStream_t     s)
{
    const int ntiles = ceil_div(nmax, kTile);
    if(ntiles == 0)
        return 0;
    hist.resize((size_t)256 * ntiles);
    KeyT*    kb[2] = {k0, k1};
    int32_t* vb[2] = {v0, v1};
    int      cur   = 0;
    for(int shift = 0; shift < bits; shift += 8)
    {
        int      w     = (bits - shift) < 8 ? (bits - shift) : 8;
        uint32_t dmask = (1u << w) - 1u;
        auto     pass  = [&](auto src) {
            using Src = decltype(src);
            rs_hist_kernel<KeyT, Src><<<ntiles, kThreads, 0, s>>>(src, shift, dmask, hist.data(), ntiles);
            scan_exclusive(hist.data(), (int64_t)256 * ntiles, chunk, nullptr, s);
            rs_scatter_kernel<KeyT, Src><<<ntiles, kThreads, 0, s>>>(src, kb[cur ^ 1], vb[cur ^ 1], shift, dmask,
                                                                     hist.data(), ntiles);
        };
        if(shift == 0)
            pass(first);
        else
            pass(PairSource<KeyT>{d_n, kb[cur], vb[cur]});
        cur ^= 1;
    }
    return cur;
}

template <typename KeyT>
int radix_sort_pairs_t(KeyT*           k0,
                       int32_t*        v0,
                       KeyT*           k1,
                       int32_t*        v1,
                       const int64_t*  d_n,
                       int64_t         nmax,
                       int             bits,
                       DBuf<uint32_t>& hist,
                       DBuf<uint32_t>& chunk,
                       hipStream_t     s)
{
    return radix_sort_from<KeyT>(PairSource<KeyT>{d_n, k0, v0}, k0, v0, k1, v1, d_n, nmax, bits, hist, chunk, s);
}

int radix_sort_pairs(uint64_t*       k0,
                     int32_t*        v0,
                     uint64_t*       k1,
                     int32_t*        v1,
                     const int64_t*  d_n,
                     int64_t         nmax,
                     int             bits,
                     DBuf<uint32_t>& hist,
                     DBuf<uint32_t>& chunk,
                     hipStream_t     s)
{
    return radix_sort_pairs_t<uint64_t>(k0, v0, k1, v1, d_n, nmax, bits, hist, chunk, s);
}

void exclusive_scan_u32(uint32_t* data, int64_t n, DBuf<uint32_t>& chunk, int64_t* d_total, hipStream_t s)
{
    scan_exclusive(data, n, chunk, d_total, s);
}

int radix_key_bits(uint64_t max_key)
{
    return key_bits_for(max_key);
}

void MatrixConverter::convert(int32_t        B,
                              int64_t        T,
                              const int32_t* rows,
                              const int32_t* cols,
                              const void*    vals,
                              bool           vals32,
                              BsrMatrix&     out,
                              bool           want32,
                              bool           build_full,
                              double         ratio,
                              hipStream_t    s)
{
    MI_REQUIRE(B >= 0 && T >= 0, MI_ERR_ARG, "convert_sym: negative size");
    MI_REQUIRE(T < ((int64_t)1 << 31), MI_ERR_ARG, "convert_sym: more than 2^31-1 triplets");
    MI_REQUIRE(T == 0 || (rows && cols && vals), MI_ERR_ARG, "convert_sym: null triplet arrays");
    out.B = B;
    out.row_ptr.resize((size_t)B + 1, ratio);
    out.full_ptr.resize((size_t)B + 1, ratio);
    out.has32 = want32;
    if(!pinned_)
        MI_HIP(hipHostMalloc((void**)&pinned_, 8 * sizeof(int64_t)));
    counts_.resize(6);
    MI_HIP(hipMemsetAsync(counts_.data(), 0, 6 * sizeof(int64_t), s));
    codec_ = KeyCodec{(uint64_t)std::max(B, 1), 0, 0};
    out.has_win  = false;
    out.has_full = false;
    if(T == 0 || B == 0)
    {
        out.U = 0;
        MI_HIP(hipMemsetAsync(out.row_ptr.data(), 0, sizeof(int32_t) * ((size_t)B + 1), s));
        MI_HIP(hipMemsetAsync(out.full_ptr.data(), 0, sizeof(int32_t) * ((size_t)B + 1), s));
        lower_ptr_.resize((size_t)B + 1, ratio);
        MI_HIP(hipMemsetAsync(lower_ptr_.data(), 0, sizeof(int32_t) * ((size_t)B + 1), s));
        return;
    }
    int64_t* d_m   = counts_.data();      // kept triplets
    int64_t* d_u   = counts_.data() + 1;  // unique upper blocks
    int64_t* d_off = counts_.data() + 2;  // off-diagonal upper blocks

    // 1. count the kept triplets and their largest col - row; one host read
    // of the band chooses the key (narrow 32-bit keys when they fit)
    const int ntiles = ceil_div(T, kTile);
    {
        tile_cnt_.resize((size_t)ntiles + 1, ratio);
        int32_t* d_band = reinterpret_cast<int32_t*>(counts_.data() + 5);
        cv_count_kernel<0><<<ntiles, kThreads, 0, s>>>(T, rows, cols, B, tile_cnt_.data(), d_band);
        MI_HIP(hipMemcpyAsync(pinned_ + 5, counts_.data() + 5, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        MI_HIP(hipStreamSynchronize(s));
        const int32_t band  = (int32_t)(pinned_[5] & 0x7fffffff);
        const int     bb    = key_bits_for((uint64_t)std::max(band, 1));
        const int     nbits = key_bits_for(((uint64_t)(B - 1) << bb) | (uint64_t)band);
        static const bool wide_only = [] {
            const char* e = std::getenv("MI_SORT_KEY");
            return e && std::string(e) == "wide";
        }();
        if(!wide_only && nbits <= 32)
        {
            codec_    = KeyCodec{(uint64_t)B, bb, band};
            key_bits_ = nbits;
        }
        else
        {
            codec_    = KeyCodec{(uint64_t)B, 0, 0};
            key_bits_ = key_bits_for((uint64_t)B * (uint64_t)B - 1ull);
        }
        counted_ = true;
    }
    const bool narrow = codec_.bb > 0;
    // 2-4. scan, compact, stable radix sort (matrix_converter.inl:1579-1616)
    const int cur = narrow ? compact_sort<0, uint32_t>(T, rows, cols, B, d_m, ratio, s)
                           : compact_sort<0, uint64_t>(T, rows, cols, B, d_m, ratio, s);

    // 5-6. run-length encode (1626-1650); the one host read-back sizes the outputs
    if(narrow)
        rle_count_kernel<uint32_t><<<ntiles, kThreads, 0, s>>>(d_m, key_ptr<uint32_t>(cur), tile_cnt_.data());
    else
        rle_count_kernel<uint64_t><<<ntiles, kThreads, 0, s>>>(d_m, key_ptr<uint64_t>(cur), tile_cnt_.data());
    scan_exclusive(tile_cnt_.data(), ntiles, chunk_, d_u, s);
    MI_HIP(hipMemcpyAsync(pinned_, counts_.data(), 2 * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    const int64_t U = pinned_[1];
    out.U           = U;
    out.rows.resize((size_t)(U > 0 ? U : 1), ratio);
    out.cols.resize((size_t)(U > 0 ? U : 1), ratio);
    out.vals.resize((size_t)9 * out.rows.cap);
    seg_start_.resize((size_t)(U > 0 ? U : 1), ratio);
    if(narrow)
        rle_write_kernel<uint32_t><<<ntiles, kThreads, 0, s>>>(d_m, key_ptr<uint32_t>(cur), tile_cnt_.data(),
                                                               codec(), out.rows.data(), out.cols.data(),
                                                               seg_start_.data());
    else
        rle_write_kernel<uint64_t><<<ntiles, kThreads, 0, s>>>(d_m, key_ptr<uint64_t>(cur), tile_cnt_.data(),
                                                               codec(), out.rows.data(), out.cols.data(),
                                                               seg_start_.data());

    // 7-8. segmented reduce over the sorted permutation (1652-1698)
    if(U > 0 && vals32)
        seg_reduce_kernel<float><<<grid_of(U), kThreads, 0, s>>>(
            U, d_m, seg_start_.data(), srcs_[cur].data(), static_cast<const float*>(vals), out.vals.data());
    else if(U > 0 && 72 * T + 16 < ((int64_t)1 << 32))
        seg_reduce_kernel<double, true><<<grid_of(U), kThreads, 0, s>>>(
            U, d_m, seg_start_.data(), srcs_[cur].data(), static_cast<const double*>(vals), out.vals.data(), T);
    else if(U > 0)
        seg_reduce_kernel<double><<<grid_of(U), kThreads, 0, s>>>(
            U, d_m, seg_start_.data(), srcs_[cur].data(), static_cast<const double*>(vals), out.vals.data());
    row_ptr_kernel<<<grid_of(U + 1), kThreads, 0, s>>>(nullptr, U, out.rows.data(), out.row_ptr.data(), B);

    // the transpose of the off-diagonal upper blocks (the lower triangle)
    const int tcur = narrow ? compact_sort<1, uint32_t>(U, out.rows.data(), out.cols.data(), B, d_off, ratio, s)
                            : compact_sort<1, uint64_t>(U, out.rows.data(), out.cols.data(), B, d_off, ratio, s);
    tcur_          = tcur;
    lower_rows_.resize((size_t)(U > 0 ? U : 1), ratio);
    lower_ptr_.resize((size_t)B + 1, ratio);
    if(narrow)
        unpack_lower_kernel<uint32_t><<<grid_of(U), kThreads, 0, s>>>(d_off, key_ptr<uint32_t>(tcur), codec(),
                                                                      lower_rows_.data());
    else
        unpack_lower_kernel<uint64_t><<<grid_of(U), kThreads, 0, s>>>(d_off, key_ptr<uint64_t>(tcur), codec(),
                                                                      lower_rows_.data());
    row_ptr_kernel<<<grid_of(U + 1), kThreads, 0, s>>>(d_off, U, lower_rows_.data(), lower_ptr_.data(), B);
    out.has_win  = false;
    out.has_full = false;
    if(build_full)
        this->build_full(out, want32, ratio, s);
    MI_HIP(hipGetLastError());
}

// the full BSR (both triangles) of the last conversion, from its upper BCOO
// and its sorted lower keys (both kept until the next conversion)
void MatrixConverter::build_full(BsrMatrix& out, bool want32, double ratio, hipStream_t s)
{
    const int32_t B      = out.B;
    const int64_t U      = out.U;
    const bool    narrow = codec_.bb > 0;
    const int     tcur   = tcur_;
    int64_t*      d_off  = counts_.data() + 2;
    if(B == 0)
        return;
    {
        uint32_t* cnt = reinterpret_cast<uint32_t*>(out.full_ptr.data());  // scanned in place
        full_count_kernel<<<grid_of(B + 1), kThreads, 0, s>>>(lower_ptr_.data(), out.row_ptr.data(), cnt, B);
        scan_exclusive(cnt, (int64_t)B + 1, chunk_, counts_.data() + 3, s);
    }
    // SpMV chunks over the full rows (the full BSR has at most 2U + B blocks)
    const int64_t fmax = 2 * U + B;
    out.nchunks = (int32_t)std::max<int64_t>(1, (fmax + kChunkStep - 1) / kChunkStep);
    out.chunk_row.resize((size_t)out.nchunks + 1, ratio);
    chunk_row_kernel<<<grid_of(out.nchunks + 1), kThreads, 0, s>>>(out.full_ptr.data(), B, out.nchunks,
                                                                   out.chunk_row.data());
    // two blocks past the end: the FP32 SpMV reads blocks in aligned pairs;
    // an even plane stride keeps every pair 8-byte aligned
    out.fcols.resize((size_t)fmax + 2, ratio);
    out.fstride = ((int64_t)out.fcols.cap + 1) & ~(int64_t)1;
    auto full_lower = [&](auto* fv) {
        using VT = std::remove_pointer_t<decltype(fv)>;
        if(narrow)
            full_lower_kernel<VT, uint32_t><<<grid_of(U), kThreads, 0, s>>>(
                d_off, key_ptr<uint32_t>(tcur), codec(), srcs_[tcur].data(), lower_ptr_.data(), out.full_ptr.data(),
                out.vals.data(), B, out.fcols.data(), fv, out.fstride);
        else
            full_lower_kernel<VT, uint64_t><<<grid_of(U), kThreads, 0, s>>>(
                d_off, key_ptr<uint64_t>(tcur), codec(), srcs_[tcur].data(), lower_ptr_.data(), out.full_ptr.data(),
                out.vals.data(), B, out.fcols.data(), fv, out.fstride);
    };
    if(want32)
    {
        out.fvals32.resize((size_t)9 * out.fstride);
        full_lower(out.fvals32.data());
        full_upper_kernel<float><<<grid_of(U), kThreads, 0, s>>>(U, out.rows.data(), out.cols.data(),
                                                                 out.row_ptr.data(), lower_ptr_.data(),
                                                                 out.full_ptr.data(), out.vals.data(),
                                                                 out.fcols.data(), out.fvals32.data(), out.fstride);
        full_empty_kernel<float><<<grid_of(B), kThreads, 0, s>>>(lower_ptr_.data(), out.row_ptr.data(),
                                                                 out.full_ptr.data(), B, out.fcols.data(),
                                                                 out.fvals32.data(), out.fstride);
    }
    else
    {
        out.fvals.resize((size_t)9 * out.fstride);
        full_lower(out.fvals.data());
        full_upper_kernel<double><<<grid_of(U), kThreads, 0, s>>>(U, out.rows.data(), out.cols.data(),
                                                                  out.row_ptr.data(), lower_ptr_.data(),
                                                                  out.full_ptr.data(), out.vals.data(),
                                                                  out.fcols.data(), out.fvals.data(), out.fstride);
        full_empty_kernel<double><<<grid_of(B), kThreads, 0, s>>>(lower_ptr_.data(), out.row_ptr.data(),
                                                                  out.full_ptr.data(), B, out.fcols.data(),
                                                                  out.fvals.data(), out.fstride);
    }
    out.has_full = true;
    MI_HIP(hipGetLastError());
}
}  // namespace mi
