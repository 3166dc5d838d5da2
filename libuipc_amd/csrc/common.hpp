// Shared helpers of the MI355X linear-solve library: error handling across the
// C ABI, growable device buffers, and wave64 reduction primitives.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>
#include <stdexcept>
#include <string>

#include "../../include/mi_linsys.h"

namespace mi
{
// An error that crosses the C ABI as (code, message).
struct Error : std::runtime_error
{
    int code;
    Error(int c, const std::string& m)
        : std::runtime_error(m)
        , code(c)
    {
    }
};

void set_last_error(const std::string& msg);

#define MI_HIP(call)                                                                 \
    do                                                                               \
    {                                                                                \
        hipError_t mi_e_ = (call);                                                   \
        if(mi_e_ != hipSuccess)                                                      \
            throw ::mi::Error(MI_ERR_HIP,                                            \
                              std::string(#call) + " failed: " + hipGetErrorString(mi_e_) \
                                  + " (" __FILE__ ":" + std::to_string(__LINE__) + ")"); \
    } while(0)

#define MI_REQUIRE(cond, code, msg)                                                  \
    do                                                                               \
    {                                                                                \
        if(!(cond))                                                                  \
            throw ::mi::Error((code), (msg));                                        \
    } while(0)

// Growable device buffer: capacity grows by `ratio` when a resize exceeds it
// (the reserve semantics of GlobalLinearSystem, global_linear_system.cu:321-405).
template <typename T>
struct DBuf
{
    T*     ptr = nullptr;
    size_t n   = 0;
    size_t cap = 0;

    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }

    void release()
    {
        if(ptr)
            (void)hipFree(ptr);
        ptr = nullptr;
        n = cap = 0;
    }
    // discard contents, ensure room for `count`
    void resize(size_t count, double ratio = 1.0)
    {
        if(count > cap)
        {
            size_t want = (size_t)((double)count * (ratio > 1.0 ? ratio : 1.0));
            if(want < count)
                want = count;
            if(want == 0)
                want = 1;
            T* p = nullptr;
            hipError_t e = hipMalloc((void**)&p, want * sizeof(T));
            if(e != hipSuccess)
                throw Error(MI_ERR_NOMEM,
                            "hipMalloc of " + std::to_string(want * sizeof(T)) + " bytes failed: "
                                + hipGetErrorString(e));
            if(ptr)
                (void)hipFree(ptr);
            ptr = p;
            cap = want;
        }
        n = count;
    }
    T*       data() { return ptr; }
    const T* data() const { return ptr; }
    size_t   size() const { return n; }
    size_t   bytes() const { return n * sizeof(T); }
};

inline int ceil_div(int64_t a, int64_t b)
{
    return (int)((a + b - 1) / b);
}

// Persistent-grid size for grid-stride kernels: a few workgroups per CU.
int device_cu_count();

// ------------------------------------------------------------------ device
constexpr int kWave = 64;

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for(int o = kWave / 2; o > 0; o >>= 1)
        v += __shfl_xor(v, o, kWave);
    return v;
}

// ---- DPP cross-lane moves (VALU; a shuffle is an LDS round trip).  A DPP
// row is 16 lanes.  Every lane of the rows read must be active.
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ double dpp_f64(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, 0xf, false);
    return __hiloint2double(hi, lo);
}
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ float dpp_f32(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, ROWS, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_t(double v)
{
    return dpp_f64<CTRL>(v);
}
template <int CTRL>
__device__ __forceinline__ float dpp_t(float v)
{
    return dpp_f32<CTRL>(v);
}
constexpr int kDppQuadXor1   = 0xB1;   // quad_perm [1,0,3,2]
constexpr int kDppQuadXor2   = 0x4E;   // quad_perm [2,3,0,1]
constexpr int kDppHalfMirror = 0x141;  // row_half_mirror: lane i of 8 <- 7 - i
constexpr int kDppMirror     = 0x140;  // row_mirror: lane i of 16 <- 15 - i
constexpr int kDppBcast15    = 0x142;  // row_bcast:15: lane 15 of a row into the next row
constexpr int kDppBcast31    = 0x143;  // row_bcast:31: lane 31 into rows 2 and 3

// Sum over the lane's 16-lane row, the same bits in every lane of the row
// (each step adds two equal-ordered halves, and a + b == b + a exactly):
// ((x0 + x1) + (x2 + x3)) + ... quads, then halves, then the row.
__device__ __forceinline__ double row16_sum(double v)
{
    v += dpp_f64<kDppQuadXor1>(v);
    v += dpp_f64<kDppQuadXor2>(v);
    v += dpp_f64<kDppHalfMirror>(v);
    v += dpp_f64<kDppMirror>(v);
    return v;
}

// Sum over aligned groups of g lanes (g = 4, 8, ..., 64, uniform over the
// wave), every lane of a group getting it: the same bits as the ascending xor
// butterfly (after each step a half-group holds one value, so the mirror's
// partner lane and the xor partner carry the same bits); strides of 16 and up
// are shuffles.
template <typename T>
__device__ __forceinline__ T group_sum(T v, int g)
{
    v += dpp_t<kDppQuadXor1>(v);
    v += dpp_t<kDppQuadXor2>(v);
    if(g >= 8)
        v += dpp_t<kDppHalfMirror>(v);
    if(g >= 16)
        v += dpp_t<kDppMirror>(v);
    for(int o = 16; o < g; o <<= 1)
        v += __shfl_xor(v, o, kWave);
    return v;
}

// Sum over the wave, (R3 + R2) + (R1 + R0) of its row sums, uniform (read
// from lane 63).  All 64 lanes must be active.
__device__ __forceinline__ double wave_total(double v)
{
    v = row16_sum(v);
    v += dpp_f64<kDppBcast15, 0xa>(v);  // rows 1, 3: R1 + R0, R3 + R2
    v += dpp_f64<kDppBcast31, 0xc>(v);  // rows 2, 3: + (R1 + R0)
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane(__double2loint(v), 63);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double((int)hi, (int)lo);
}

// Sum over a workgroup of `NT` threads; every thread gets the result.
// `lds` must hold NT / 64 values.  Called by every thread of the workgroup.
template <int NT, typename T>
__device__ __forceinline__ T block_sum(T v, T* lds)
{
    if constexpr(std::is_same<T, double>::value)
        v = wave_total(v);
    else
        v = wave_sum(v);
    const int w   = threadIdx.x / kWave;
    const int l   = threadIdx.x % kWave;
    if(l == 0)
        lds[w] = v;
    __syncthreads();
    T acc = T(0);
#pragma unroll
    for(int k = 0; k < NT / kWave; ++k)
        acc += lds[k];
    __syncthreads();
    return acc;
}

// Two sums over a workgroup in one pass (one pair of barriers); each is added
// in the same order as block_sum's.  `lds` must hold 2 * NT / 64 values.
template <int NT, typename T>
__device__ __forceinline__ void block_sum2(T& a, T& b, T* lds)
{
    a           = wave_total(a);
    b           = wave_total(b);
    const int w = threadIdx.x / kWave;
    if(threadIdx.x % kWave == 0)
    {
        lds[2 * w]     = a;
        lds[2 * w + 1] = b;
    }
    __syncthreads();
    T s0 = T(0), s1 = T(0);
#pragma unroll
    for(int k = 0; k < NT / kWave; ++k)
    {
        s0 += lds[2 * k];
        s1 += lds[2 * k + 1];
    }
    __syncthreads();
    a = s0;
    b = s1;
}

// Exclusive prefix within a workgroup of NT threads (values >= 0).
// Returns the exclusive prefix of this thread; *total = workgroup total.
// `lds` must hold NT / 64 + 1 values.
template <int NT, typename T>
__device__ __forceinline__ T block_exclusive_scan(T v, T* lds, T* total)
{
    const int l = threadIdx.x % kWave;
    const int w = threadIdx.x / kWave;
    // inclusive wave scan
    T inc = v;
#pragma unroll
    for(int o = 1; o < kWave; o <<= 1)
    {
        T t = __shfl_up(inc, o, kWave);
        if(l >= o)
            inc += t;
    }
    if(l == kWave - 1)
        lds[w] = inc;
    __syncthreads();
    T woff = T(0), tot = T(0);
#pragma unroll
    for(int k = 0; k < NT / kWave; ++k)
    {
        T s = lds[k];
        if(k < w)
            woff += s;
        tot += s;
    }
    __syncthreads();
    if(total)
        *total = tot;
    return woff + inc - v;
}

__device__ __forceinline__ void atomic_add_f64(double* p, double v)
{
    unsafeAtomicAdd(p, v);
}
__device__ __forceinline__ void atomic_add_f32(float* p, float v)
{
    unsafeAtomicAdd(p, v);
}
}  // namespace mi
