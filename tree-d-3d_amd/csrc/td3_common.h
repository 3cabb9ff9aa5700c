// Shared device helpers of the gfx950 field-evaluation kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "td3.h"

#define TD3_WAVE 64

typedef int64_t i64;
typedef uint64_t u64;
typedef uint32_t u32;

// Error state of the library (host side, one string per thread).
void td3_set_error(const char* fmt, ...);

// Kernel timing (host side): TD3_TIMED(stream, "name", launch) brackets a launch with
// HIP events when timing is enabled (td3_timing_enable), and costs one branch otherwise.
void* td3_timing_begin(hipStream_t stream);
void td3_timing_end(void* token, const char* name, hipStream_t stream);
#define TD3_TIMED(stream, name, ...)                    \
    do {                                                \
        void* _tok = td3_timing_begin(stream);          \
        __VA_ARGS__;                                    \
        td3_timing_end(_tok, name, stream);             \
    } while (0)

#define TD3_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t _e = (call);                                                         \
        if (_e != hipSuccess) {                                                         \
            td3_set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(_e)); \
            return TD3_ERR_HIP;                                                         \
        }                                                                               \
    } while (0)

#define TD3_LAUNCHED()                                                                  \
    do {                                                                                \
        hipError_t _e = hipGetLastError();                                              \
        if (_e != hipSuccess) {                                                         \
            td3_set_error("%s:%d launch failed: %s", __FILE__, __LINE__, hipGetErrorString(_e)); \
            return TD3_ERR_HIP;                                                         \
        }                                                                               \
    } while (0)

#define TD3_ARG(cond, msg)                                                              \
    do {                                                                                \
        if (!(cond)) {                                                                  \
            td3_set_error("bad argument: %s", msg);                                     \
            return TD3_ERR_ARG;                                                         \
        }                                                                               \
    } while (0)

static inline unsigned td3_blocks(i64 items, int threads) {
    i64 b = (items + threads - 1) / threads;
    return (unsigned)(b < 1 ? 1 : b);
}

// MI355X deals a launch's blocks to its eight XCDs round robin (block b runs on XCD b % 8),
// and each XCD has its own L2.  A launch of td3_xcd_grid(blocks) blocks that takes its work
// from block td3_xcd_block() instead of blockIdx.x gives each XCD one contiguous eighth of
// the blocks, so neighbouring work (the same tree, nearby leaves) shares an L2.  Blocks
// past the original count come out of td3_xcd_block() too: their work is past the end.
constexpr unsigned kXcds = 8;
static inline unsigned td3_xcd_grid(unsigned blocks) { return (blocks + kXcds - 1) / kXcds * kXcds; }
__device__ __forceinline__ unsigned td3_xcd_block() {
    return (blockIdx.x % kXcds) * (gridDim.x / kXcds) + blockIdx.x / kXcds;
}

// status[0] keeps the smallest positive report (1 + an index), so the result does not
// depend on which wave reports first; a negative report (stack or depth overflow)
// wins over every positive one.
__device__ __forceinline__ void td3_report(i64* status, i64 value) {
    unsigned long long* s = reinterpret_cast<unsigned long long*>(status);
    unsigned long long old = __hip_atomic_load(s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
        const i64 o = (i64)old;
        if (o < 0) return;
        if (value > 0 && o > 0 && o <= value) return;
        unsigned long long assumed = old;
        old = atomicCAS(s, assumed, (unsigned long long)value);
        if (old == assumed) return;
    }
}

// Sum of a double over the 64 lanes of a wave (every lane gets the total).
__device__ __forceinline__ double td3_wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, TD3_WAVE);
    return v;
}

__device__ __forceinline__ float td3_wave_sumf(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, TD3_WAVE);
    return v;
}

__device__ __forceinline__ int td3_lane() { return (int)(threadIdx.x & (TD3_WAVE - 1)); }

// Sum over each row of 16 lanes by DPP moves (quad_perm [1,0,3,2], [2,3,0,1], then the
// half-row and row mirrors): four adds, no LDS traffic.  Every lane of the row gets the
// row's sum, the same terms in an order fixed per lane.
template <int CTRL>
__device__ __forceinline__ float td3_dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ double td3_dpp_d(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Sum of a double over each group of `lanes` (4, 8 or 16, uniform) consecutive lanes by DPP
// moves: every lane of the group gets the group's sum.
__device__ __forceinline__ double td3_group_sum(double v, int lanes) {
    v += td3_dpp_d<0xb1>(v);
    v += td3_dpp_d<0x4e>(v);
    if (lanes >= 8) v += td3_dpp_d<0x141>(v);
    if (lanes >= 16) v += td3_dpp_d<0x140>(v);
    return v;
}

__device__ __forceinline__ float td3_row16_sum(float v) {
    v += td3_dpp_f<0xb1>(v);
    v += td3_dpp_f<0x4e>(v);
    v += td3_dpp_f<0x141>(v);
    v += td3_dpp_f<0x140>(v);
    return v;
}

