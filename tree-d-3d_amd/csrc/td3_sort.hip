// LSD radix sort and exclusive scan used by the tree build (see td3_sort.h).

#include "td3_sort.h"

namespace td3sort {
namespace {

__global__ __launch_bounds__(kThreads) void hist_kernel(const u64* __restrict__ keys, i64 n,
                                                        int bit, u32 mask, i64 ntiles,
                                                        u32* __restrict__ hist) {
    __shared__ u32 s_hist[kRadix];
    s_hist[threadIdx.x] = 0;
    __syncthreads();
    const i64 base = (i64)blockIdx.x * kTile;
#pragma unroll 4
    for (int r = 0; r < kItems; ++r) {
        const i64 idx = base + r * kThreads + threadIdx.x;
        if (idx < n) {
            const u32 d = (u32)(keys[idx] >> bit) & mask;
            atomicAdd(&s_hist[d], 1u);
        }
    }
    __syncthreads();
    hist[(i64)threadIdx.x * ntiles + blockIdx.x] = s_hist[threadIdx.x];
}

// Single-workgroup exclusive scan of a short array (the per-tile sums): each of 1024
// threads owns a contiguous segment.  Writes the grand total to *total when non-null.
__global__ __launch_bounds__(1024) void scan_small_kernel(u32* data, i64 n,
                                                          u32* __restrict__ total) {
    __shared__ u32 s_sum[1024];
    const i64 per = (n + 1023) / 1024;
    const i64 lo = (i64)threadIdx.x * per;
    const i64 hi = lo + per < n ? lo + per : n;
    u32 sum = 0;
    for (i64 k = lo; k < hi; ++k) sum += data[k];
    s_sum[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const u32 v = threadIdx.x >= (unsigned)off ? s_sum[threadIdx.x - off] : 0u;
        __syncthreads();
        s_sum[threadIdx.x] += v;
        __syncthreads();
    }
    u32 run = threadIdx.x > 0 ? s_sum[threadIdx.x - 1] : 0u;
    for (i64 k = lo; k < hi; ++k) {
        const u32 v = data[k];
        data[k] = run;
        run += v;
    }
    if (total != nullptr && threadIdx.x == 1023) *total = s_sum[1023];
}

// Sum of each tile of kTile counts (coalesced loads).
__global__ __launch_bounds__(kThreads) void tile_sum_kernel(const u32* __restrict__ in, i64 n,
                                                            u32* __restrict__ part) {
    __shared__ u32 s_wave[kThreads / TD3_WAVE];
    const i64 base = (i64)blockIdx.x * kTile;
    u32 sum = 0;
#pragma unroll 4
    for (int r = 0; r < kItems; ++r) {
        const i64 idx = base + r * kThreads + threadIdx.x;
        if (idx < n) sum += in[idx];
    }
    for (int off = TD3_WAVE / 2; off > 0; off >>= 1) sum += __shfl_xor(sum, off, TD3_WAVE);
    if (threadIdx.x % TD3_WAVE == 0) s_wave[threadIdx.x / TD3_WAVE] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        u32 t = 0;
        for (int w = 0; w < kThreads / TD3_WAVE; ++w) t += s_wave[w];
        part[blockIdx.x] = t;
    }
}

// Exclusive scan of each tile, offset by the scanned tile sums.  Each element is read and
// written by the same thread, so `out` may alias `in`.
// With part null (one tile) the carry starts at zero and the total goes to *total.
__global__ __launch_bounds__(kThreads) void tile_scan_kernel(const u32* in, u32* out, i64 n,
                                                             const u32* __restrict__ part,
                                                             u32* __restrict__ total) {
    __shared__ u32 s_wave[kThreads / TD3_WAVE];
    const int lane = threadIdx.x % TD3_WAVE;
    const int wave = threadIdx.x / TD3_WAVE;
    const i64 base = (i64)blockIdx.x * kTile;
    u32 carry = part ? part[blockIdx.x] : 0u;
    for (int r = 0; r < kItems; ++r) {
        const i64 idx = base + r * kThreads + threadIdx.x;
        const u32 v = idx < n ? in[idx] : 0u;
        u32 incl = v;
#pragma unroll
        for (int off = 1; off < TD3_WAVE; off <<= 1) {
            const u32 t = __shfl_up(incl, off, TD3_WAVE);
            if (lane >= off) incl += t;
        }
        if (lane == TD3_WAVE - 1) s_wave[wave] = incl;
        __syncthreads();
        u32 before = 0, all = 0;
#pragma unroll
        for (int w = 0; w < kThreads / TD3_WAVE; ++w) {
            const u32 t = s_wave[w];
            before += w < wave ? t : 0u;
            all += t;
        }
        if (idx < n) out[idx] = carry + before + incl - v;
        carry += all;
        __syncthreads();
    }
    if (total != nullptr && threadIdx.x == 0) *total = carry;
}

template <bool AUX, bool IDENT>
__global__ __launch_bounds__(kThreads) void scatter_kernel(
    const u64* __restrict__ keys, const u32* __restrict__ vals, const u64* __restrict__ aux,
    u64* __restrict__ okeys, u32* __restrict__ ovals, u64* __restrict__ oaux, i64 n, int bit,
    u32 mask, const u32* __restrict__ offsets, i64 ntiles) {
    __shared__ u32 s_base[kRadix];
    __shared__ u32 s_wave[kThreads / TD3_WAVE][kRadix];
    const int tid = threadIdx.x;
    const int wave = tid / TD3_WAVE;
    const int lane = tid % TD3_WAVE;
    const u64 lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    s_base[tid] = offsets[(i64)tid * ntiles + blockIdx.x];
    const i64 base = (i64)blockIdx.x * kTile;
    for (int r = 0; r < kItems; ++r) {
        const i64 idx = base + r * kThreads + tid;
        const bool valid = idx < n;
        const u64 key = valid ? keys[idx] : 0ull;
        const u32 d = (u32)(key >> bit) & mask;
#pragma unroll
        for (int k = 0; k < kRadix / TD3_WAVE; ++k) s_wave[wave][lane * (kRadix / TD3_WAVE) + k] = 0;
        __syncthreads();
        u64 peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool set = (d >> b) & 1u;
            const u64 m = __ballot(valid && set);
            peers &= set ? m : ~m;
        }
        const u32 rank = (u32)__popcll(peers & lt_mask);
        if (valid && rank == 0) s_wave[wave][d] = (u32)__popcll(peers);
        __syncthreads();
        {
            u32 run = s_base[tid];
#pragma unroll
            for (int w = 0; w < kThreads / TD3_WAVE; ++w) {
                const u32 c = s_wave[w][tid];
                s_wave[w][tid] = run;
                run += c;
            }
            s_base[tid] = run;
        }
        __syncthreads();
        if (valid) {
            const u32 pos = s_wave[wave][d] + rank;
            okeys[pos] = key;
            ovals[pos] = IDENT ? (u32)idx : vals[idx];
            if (AUX) oaux[pos] = aux[idx];
        }
        __syncthreads();
    }
}

}  // namespace

int exclusive_scan_u32(const u32* in, u32* out, i64 n, u32* total, u32* scratch,
                       hipStream_t stream) {
    if (n <= 0) {
        if (total) TD3_HIP(hipMemsetAsync(total, 0, sizeof(u32), stream));
        return TD3_OK;
    }
    const i64 nt = tiles(n);
    if (nt == 1) {
        // one tile: a single launch
        TD3_TIMED(stream, "scan", hipLaunchKernelGGL(tile_scan_kernel, dim3(1), dim3(kThreads), 0, stream, in, out, n,
                                                     (const u32*)nullptr, total));
        TD3_LAUNCHED();
        return TD3_OK;
    }
    u32* part = scratch;
    TD3_TIMED(stream, "scan", hipLaunchKernelGGL(tile_sum_kernel, dim3((unsigned)nt), dim3(kThreads), 0, stream, in, n, part));
    TD3_LAUNCHED();
    TD3_TIMED(stream, "scan", hipLaunchKernelGGL(scan_small_kernel, dim3(1), dim3(1024), 0, stream, part, nt, total));
    TD3_LAUNCHED();
    TD3_TIMED(stream, "scan", hipLaunchKernelGGL(tile_scan_kernel, dim3((unsigned)nt), dim3(kThreads), 0, stream, in, out, n, part,
                                                 (u32*)nullptr));
    TD3_LAUNCHED();
    return TD3_OK;
}

int sort_pairs(u64* keys, u32* vals, u64* aux, u64* keys_alt, u32* vals_alt, u64* aux_alt,
               i64 n, int bit_lo, int bit_hi, bool vals_identity, u32* scratch,
               hipStream_t stream) {
    if (n <= 0 || bit_hi <= bit_lo) return TD3_OK;
    const i64 nt = tiles(n);
    u32* hist = scratch;
    u64* src_k = keys;
    u32* src_v = vals;
    u64* src_a = aux;
    u64* dst_k = keys_alt;
    u32* dst_v = vals_alt;
    u64* dst_a = aux_alt;
    bool ident = vals_identity;
    int passes = 0;
    for (int bit = bit_lo; bit < bit_hi; bit += 8) {
        const int width = (bit_hi - bit) < 8 ? (bit_hi - bit) : 8;
        const u32 mask = (1u << width) - 1u;
        TD3_TIMED(stream, "sort_hist", hipLaunchKernelGGL(hist_kernel, dim3((unsigned)nt), dim3(kThreads), 0, stream, src_k, n, bit,
                           mask, nt, hist));
        TD3_LAUNCHED();
        if (int rc = exclusive_scan_u32(hist, hist, nt * kRadix, nullptr, hist + nt * kRadix,
                                        stream))
            return rc;
        if (aux != nullptr) {
            if (ident)
                TD3_TIMED(stream, "sort_scatter", hipLaunchKernelGGL((scatter_kernel<true, true>), dim3((unsigned)nt), dim3(kThreads), 0,
                                   stream, src_k, src_v, src_a, dst_k, dst_v, dst_a, n, bit, mask,
                                   hist, nt));
            else
                TD3_TIMED(stream, "sort_scatter", hipLaunchKernelGGL((scatter_kernel<true, false>), dim3((unsigned)nt), dim3(kThreads), 0,
                                   stream, src_k, src_v, src_a, dst_k, dst_v, dst_a, n, bit, mask,
                                   hist, nt));
        } else {
            if (ident)
                TD3_TIMED(stream, "sort_scatter", hipLaunchKernelGGL((scatter_kernel<false, true>), dim3((unsigned)nt), dim3(kThreads), 0,
                                   stream, src_k, src_v, src_a, dst_k, dst_v, dst_a, n, bit, mask,
                                   hist, nt));
            else
                TD3_TIMED(stream, "sort_scatter", hipLaunchKernelGGL((scatter_kernel<false, false>), dim3((unsigned)nt), dim3(kThreads),
                                   0, stream, src_k, src_v, src_a, dst_k, dst_v, dst_a, n, bit, mask,
                                   hist, nt));
        }
        TD3_LAUNCHED();
        ident = false;
        u64* tk = src_k;
        src_k = dst_k;
        dst_k = tk;
        u32* tv = src_v;
        src_v = dst_v;
        dst_v = tv;
        u64* ta = src_a;
        src_a = dst_a;
        dst_a = ta;
        ++passes;
    }
    if (passes & 1) {
        // The result sits in the alt buffers: copy it back.
        TD3_HIP(hipMemcpyAsync(keys, keys_alt, sizeof(u64) * (size_t)n, hipMemcpyDeviceToDevice, stream));
        TD3_HIP(hipMemcpyAsync(vals, vals_alt, sizeof(u32) * (size_t)n, hipMemcpyDeviceToDevice, stream));
        if (aux != nullptr)
            TD3_HIP(hipMemcpyAsync(aux, aux_alt, sizeof(u64) * (size_t)n, hipMemcpyDeviceToDevice, stream));
    }
    return TD3_OK;
}

}  // namespace td3sort
