// Test infrastructure: a CPU stand-in for <hip/hip_runtime.h> that lets g++ compile the
// library's unmodified .hip sources into a host library, so the kernels' indexing can be
// run under AddressSanitizer on small inputs before anything runs on a GPU.
//
// Blocks run in parallel on a few OS threads; inside a block every GPU thread is a fiber.
// __syncthreads is a block barrier; wave operations (ballot, shuffles, the wave barrier)
// exchange values through a 64-slot array between wave barriers.  Threads that return
// early drop out of both barriers, as exited waves do on the GPU.  Device memory is host
// memory; streams are ignored (everything is synchronous).  x86-64 only.
// Not a product path: nothing in tree-d-3d_amd/ includes or loads it.
#pragma once

#include <atomic>
#include <barrier>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static thread_local

// --------------------------------------------------------------------- vector types
struct int2 {
    int x, y;
};
struct int4 {
    int x, y, z, w;
};
struct alignas(16) float4 {
    float x, y, z, w;
};
inline float4 make_float4(float x, float y, float z, float w) { return {x, y, z, w}; }
struct alignas(16) double2 {
    double x, y;
};
inline double2 make_double2(double x, double y) { return {x, y}; }
struct alignas(32) double4 {
    double x, y, z, w;
};
inline double4 make_double4(double x, double y, double z, double w) { return {x, y, z, w}; }
inline int2 make_int2(int x, int y) { return {x, y}; }
inline int4 make_int4(int x, int y, int z, int w) { return {x, y, z, w}; }
struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

// --------------------------------------------------------------------- host API
typedef int hipError_t;
typedef void* hipStream_t;
typedef void* hipEvent_t;
#define hipSuccess 0
#define hipMemcpyDeviceToHost 0
#define hipMemcpyDeviceToDevice 0
#define hipMemcpyHostToDevice 0
inline const char* hipGetErrorString(hipError_t) { return "emulated"; }
// one emulated device; a __device__ variable is a host variable, its address its own
#define __align__(n) alignas(n)
#define HIP_SYMBOL(x) (&(x))
inline hipError_t hipGetDevice(int* d) {
    *d = 0;
    return 0;
}
inline hipError_t hipGetSymbolAddress(void** p, const void* symbol) {
    *p = const_cast<void*>(symbol);
    return 0;
}
inline hipError_t hipGetLastError() { return 0; }
inline hipError_t hipMemsetAsync(void* p, int v, size_t n, hipStream_t) {
    memset(p, v, n);
    return 0;
}
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, int, hipStream_t) {
    memmove(d, s, n);
    return 0;
}
inline hipError_t hipStreamSynchronize(hipStream_t) { return 0; }
inline hipError_t hipEventCreate(hipEvent_t* e) {
    *e = nullptr;
    return 0;
}
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return 0; }
inline hipError_t hipEventDestroy(hipEvent_t) { return 0; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return 0; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) {
    *ms = 0.0f;
    return 0;
}

// --------------------------------------------------------------------- execution model
// Each block runs on one OS thread; its GPU threads are fibers switched by a small
// context switch at barriers and wave operations only.  Blocks run in parallel on a pool
// of OS threads (each block's __shared__ arrays are thread_local statics of its OS thread).
extern "C" void td3emu_switch(void** save_sp, void* next_sp);
asm(R"(
.text
.weak td3emu_switch
.type td3emu_switch,@function
td3emu_switch:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
.size td3emu_switch, .-td3emu_switch
)");

extern "C" void __sanitizer_start_switch_fiber(void** fake_stack_save, const void* bottom,
                                               size_t size) __attribute__((weak));
extern "C" void __sanitizer_finish_switch_fiber(void* fake_stack_save, const void** bottom_old,
                                                size_t* size_old) __attribute__((weak));

namespace td3emu {

constexpr size_t kStack = 256 * 1024;

struct Fiber {
    void* sp = nullptr;
    char* stack = nullptr;
    dim3 tid, bid, bdim, gdim;
    bool done = false;
    bool waiting = false;
    unsigned wave = 0;
};

struct Group {  // a barrier: the block or one wave
    unsigned active = 0, arrived = 0;
    std::vector<Fiber*> waiters;
};

struct Sched {
    std::vector<Fiber> fibers;
    std::vector<char*> stacks;
    void* main_sp = nullptr;
    const void* main_bottom = nullptr;
    size_t main_size = 0;
    Fiber* cur = nullptr;
    Group block;
    std::vector<Group> waves;
    std::vector<std::vector<unsigned long long>> slots;
    std::function<void()>* body = nullptr;
};

inline thread_local Sched* sched = nullptr;

inline void to_scheduler() {
    Sched* s = sched;
    Fiber* f = s->cur;
    void* fake = nullptr;
    if (__sanitizer_start_switch_fiber) __sanitizer_start_switch_fiber(&fake, s->main_bottom, s->main_size);
    td3emu_switch(&f->sp, s->main_sp);
    if (__sanitizer_finish_switch_fiber)
        __sanitizer_finish_switch_fiber(fake, &s->main_bottom, &s->main_size);
}

inline void release(Group& g) {
    for (Fiber* w : g.waiters) w->waiting = false;
    g.waiters.clear();
    g.arrived = 0;
}

inline void arrive(Group& g) {
    Fiber* f = sched->cur;
    if (++g.arrived >= g.active) {
        release(g);
        return;
    }
    f->waiting = true;
    g.waiters.push_back(f);
    while (f->waiting) to_scheduler();
}

inline void drop(Group& g) {
    --g.active;
    if (g.arrived && g.arrived >= g.active) release(g);
}

extern "C" inline void td3emu_entry() {
    Sched* s = sched;
    Fiber* f = s->cur;
    if (__sanitizer_finish_switch_fiber)
        __sanitizer_finish_switch_fiber(nullptr, &s->main_bottom, &s->main_size);
    (*s->body)();
    f->done = true;
    drop(s->waves[f->wave]);
    drop(s->block);
    for (;;) to_scheduler();
}

inline void run_block(Sched& s, dim3 bid, dim3 bdim, dim3 gdim) {
    const unsigned n = bdim.x * bdim.y * bdim.z;
    const unsigned nw = (n + 63) / 64;
    s.fibers.assign(n, Fiber());
    while (s.stacks.size() < n) s.stacks.push_back((char*)aligned_alloc(64, kStack));
    s.block = Group();
    s.block.active = n;
    s.waves.assign(nw, Group());
    s.slots.assign(nw, std::vector<unsigned long long>(64, 0ull));
    for (unsigned w = 0; w < nw; ++w) s.waves[w].active = (w + 1) * 64 <= n ? 64 : n - w * 64;
    for (unsigned t = 0; t < n; ++t) {
        Fiber& f = s.fibers[t];
        f.stack = s.stacks[t];
        f.tid = dim3(t, 0, 0);
        f.bid = bid;
        f.bdim = bdim;
        f.gdim = gdim;
        f.wave = t >> 6;
        // initial frame: six callee-saved registers, then the entry as return address;
        // after the ret the stack pointer is 8 mod 16, as at a function entry
        uintptr_t top = ((uintptr_t)(f.stack + kStack)) & ~(uintptr_t)15;
        void** sp = (void**)(top - 8);
        *--sp = (void*)&td3emu_entry;
        for (int r = 0; r < 6; ++r) *--sp = nullptr;
        f.sp = sp;
    }
    unsigned remaining = n;
    while (remaining) {
        bool progressed = false;
        for (unsigned t = 0; t < n; ++t) {
            Fiber& f = s.fibers[t];
            if (f.done || f.waiting) continue;
            s.cur = &f;
            void* fake = nullptr;
            if (__sanitizer_start_switch_fiber) __sanitizer_start_switch_fiber(&fake, f.stack, kStack);
            td3emu_switch(&s.main_sp, f.sp);
            if (__sanitizer_finish_switch_fiber) __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
            progressed = true;
            if (f.done) --remaining;
        }
        if (!progressed && remaining) {
            fprintf(stderr, "td3emu: deadlock in a barrier or wave operation\n");
            abort();
        }
    }
}

template <typename F>
void launch(dim3 grid, dim3 block, F&& body) {
    const size_t nblocks = (size_t)grid.x * grid.y * grid.z;
    if (nblocks == 0 || block.x * block.y * block.z == 0) return;
    std::function<void()> fn = body;
    std::atomic<size_t> next{0};
    unsigned nthreads = std::thread::hardware_concurrency();
    if (nthreads == 0) nthreads = 4;
    if (nthreads > 8) nthreads = 8;
    if (nthreads > nblocks) nthreads = (unsigned)nblocks;
    auto worker = [&] {
        Sched s;
        s.body = &fn;
        Sched* prev = sched;
        sched = &s;
        for (size_t bi = next++; bi < nblocks; bi = next++) {
            const unsigned bx = (unsigned)(bi % grid.x);
            const unsigned by = (unsigned)((bi / grid.x) % grid.y);
            const unsigned bz = (unsigned)(bi / ((size_t)grid.x * grid.y));
            run_block(s, dim3(bx, by, bz), block, grid);
        }
        for (char* st : s.stacks) free(st);
        sched = prev;
    };
    std::vector<std::thread> pool;
    for (unsigned i = 1; i < nthreads; ++i) pool.emplace_back(worker);
    worker();
    for (auto& th : pool) th.join();
}

inline Fiber& me() { return *sched->cur; }
inline unsigned lane_id() { return me().tid.x & 63u; }
inline unsigned wave_id() { return me().wave; }
inline void sync_block() { arrive(sched->block); }
inline void sync_wave() { arrive(sched->waves[me().wave]); }

// All lanes of the wave publish a value; each gets the whole vector back.
inline std::vector<unsigned long long> exchange(unsigned long long v) {
    auto& slots = sched->slots[me().wave];
    sync_wave();
    slots[lane_id()] = v;
    sync_wave();
    std::vector<unsigned long long> out(slots);
    sync_wave();
    return out;
}

}  // namespace td3emu

#define threadIdx (td3emu::me().tid)
#define blockIdx (td3emu::me().bid)
#define blockDim (td3emu::me().bdim)
#define gridDim (td3emu::me().gdim)

#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...) \
    td3emu::launch(dim3(grid), dim3(block), [&] { kernel(__VA_ARGS__); })

inline void __syncthreads() { td3emu::sync_block(); }
inline void __builtin_amdgcn_wave_barrier() { td3emu::sync_wave(); }
inline void __builtin_amdgcn_s_barrier() { td3emu::sync_block(); }
#define __builtin_amdgcn_fence(...) std::atomic_thread_fence(std::memory_order_seq_cst)
inline int __builtin_amdgcn_readfirstlane(int v) {
    auto all = td3emu::exchange((unsigned long long)(unsigned)v);
    return (int)(unsigned)all[0];
}
inline int __builtin_amdgcn_readlane(int v, int lane) {
    auto all = td3emu::exchange((unsigned long long)(unsigned)v);
    return (int)(unsigned)all[lane];
}

inline unsigned long long __ballot(int pred) {
    auto all = td3emu::exchange(pred ? 1ull : 0ull);
    unsigned long long m = 0;
    for (unsigned l = 0; l < 64; ++l)
        if (all[l]) m |= 1ull << l;
    return m;
}

template <typename T>
inline unsigned long long td3emu_bits(T v) {
    unsigned long long b = 0;
    static_assert(sizeof(T) <= 8, "shuffle of a wide type");
    memcpy(&b, &v, sizeof(T));
    return b;
}
template <typename T>
inline T td3emu_from(unsigned long long b) {
    T v;
    memcpy(&v, &b, sizeof(T));
    return v;
}
template <typename T>
inline T __shfl(T v, int src, int width = 64) {
    auto all = td3emu::exchange(td3emu_bits(v));
    const unsigned base = td3emu::lane_id() & ~(unsigned)(width - 1);
    return td3emu_from<T>(all[base + ((unsigned)src & (unsigned)(width - 1))]);
}
template <typename T>
inline T __shfl_up(T v, unsigned delta, int width = 64) {
    auto all = td3emu::exchange(td3emu_bits(v));
    const unsigned lane = td3emu::lane_id();
    const unsigned pos = lane & (unsigned)(width - 1);
    return pos >= delta ? td3emu_from<T>(all[lane - delta]) : v;
}
template <typename T>
inline T __shfl_xor(T v, int mask, int width = 64) {
    auto all = td3emu::exchange(td3emu_bits(v));
    return td3emu_from<T>(all[td3emu::lane_id() ^ (unsigned)mask]);
}
// DPP moves within rows of 16 lanes: quad_perm (controls 0x00-0xff), row_mirror (0x140) and
// row_half_mirror (0x141), every lane enabled.
inline int __builtin_amdgcn_mov_dpp(int v, int ctrl, int, int, bool) {
    auto all = td3emu::exchange(td3emu_bits(v));
    const unsigned lane = td3emu::lane_id();
    unsigned src;
    if (ctrl >= 0 && ctrl <= 0xff)
        src = (lane & ~3u) | (((unsigned)ctrl >> (2 * (lane & 3))) & 3u);
    else if (ctrl == 0x140)
        src = (lane & ~15u) | (15u - (lane & 15u));
    else if (ctrl == 0x141)
        src = (lane & ~7u) | (7u - (lane & 7u));
    else
        abort();
    return td3emu_from<int>(all[src]);
}
inline int __float_as_int(float f) {
    int i;
    memcpy(&i, &f, 4);
    return i;
}
inline long long __double_as_longlong(double d) {
    long long i;
    memcpy(&i, &d, 8);
    return i;
}
inline double __longlong_as_double(long long i) {
    double d;
    memcpy(&d, &i, 8);
    return d;
}
inline float __int_as_float(int i) {
    float f;
    memcpy(&f, &i, 4);
    return f;
}

// Matrix cores: v_mfma_{f32,f64}_16x16x4 on one wavefront.  Lane l holds A[l % 16][l / 16]
// and B[l / 16][l % 16]; element r of lane l of the result is D[4 (l / 16) + r][l % 16] for
// f32 and D[l / 16 + 4 r][l % 16] for f64 (the layouts measured on gfx950 with one-hot
// operands).
// ext_vector_type(4) becomes a 32-byte GCC vector: four doubles, or eight floats of which
// the kernels use four.
#define ext_vector_type(n) vector_size(8 * (n))
typedef float td3emu_f32x4 __attribute__((vector_size(32)));
typedef double td3emu_f64x4 __attribute__((vector_size(32)));
template <typename V, typename T>
inline V td3emu_mfma16(T a, T b, V c) {
    auto A = td3emu::exchange(td3emu_bits(a));
    auto B = td3emu::exchange(td3emu_bits(b));
    const unsigned lane = td3emu::lane_id();
    const unsigned col = lane % 16;
    V d = c;
    for (unsigned r = 0; r < 4; ++r) {
        const unsigned row = sizeof(T) == 8 ? lane / 16 + 4 * r : 4 * (lane / 16) + r;
        T sum = c[r];
        for (unsigned kk = 0; kk < 4; ++kk)
            sum += td3emu_from<T>(A[kk * 16 + row]) * td3emu_from<T>(B[kk * 16 + col]);
        d[r] = sum;
    }
    return d;
}
inline td3emu_f32x4 __builtin_amdgcn_mfma_f32_16x16x4f32(float a, float b, td3emu_f32x4 c, int,
                                                         int, int) {
    return td3emu_mfma16<td3emu_f32x4, float>(a, b, c);
}
inline td3emu_f64x4 __builtin_amdgcn_mfma_f64_16x16x4f64(double a, double b, td3emu_f64x4 c, int,
                                                         int, int) {
    return td3emu_mfma16<td3emu_f64x4, double>(a, b, c);
}

inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }
inline int __ffsll(unsigned long long v) { return __builtin_ffsll((long long)v); }
inline int __clzll(long long v) { return v ? __builtin_clzll((unsigned long long)v) : 64; }

// atomics (host atomics on plain memory)
template <typename T>
inline T atomicAdd(T* p, T v) {
    return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
}
template <typename T>
inline T atomicOr(T* p, T v) {
    return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST);
}
template <class T>
inline T atomicMax(T* p, T v) {
    T old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (old < v && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST,
                                                   __ATOMIC_SEQ_CST)) {
    }
    return old;
}
inline unsigned long long atomicCAS(unsigned long long* p, unsigned long long cmp,
                                    unsigned long long v) {
    __atomic_compare_exchange_n(p, &cmp, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
    return cmp;
}
#define __ATOMIC_RELAXED_HIP __ATOMIC_RELAXED
#define __HIP_MEMORY_SCOPE_AGENT 0
template <typename T>
inline T __hip_atomic_load(T* p, int order, int) {
    return __atomic_load_n(p, __ATOMIC_SEQ_CST);
}

// math
inline double __dadd_rn(double a, double b) { return a + b; }
inline double __dmul_rn(double a, double b) { return a * b; }
inline double __dsqrt_rn(double a) { return std::sqrt(a); }
inline float __logf(float a) { return std::log(a); }
inline float __frcp_rn(float a) { return 1.0f / a; }
inline float __frsqrt_rn(float a) { return 1.0f / std::sqrt(a); }
// the hardware's approximate transcendentals (v_log_f32 is log2)
inline float __builtin_amdgcn_logf(float a) { return std::log2(a); }
inline float __builtin_amdgcn_rcpf(float a) { return 1.0f / a; }
inline double __builtin_amdgcn_rcp(double a) { return 1.0 / a; }
inline float __builtin_amdgcn_rsqf(float a) { return 1.0f / std::sqrt(a); }
inline double rsqrt(double a) { return 1.0 / std::sqrt(a); }
using std::atan2;
using std::fabs;
using std::log;
using std::sqrt;
