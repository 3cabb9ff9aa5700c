// Dual-tree traversal and the interaction lists of the FMM.
//
// Replaces flat_tree.py:1318-1358 (_traverse: one fmm_traverse launch per generation,
// each followed by a host read of the counters) and kernels/fmm_traverse/fmm_traverse_cuda.cu:14-97.
// The rules are those of kernels/fmm_traverse/fmm_traverse_pyccel.py:15-106, evaluated in
// FP64 without contraction so the pair sets equal the reference's:
//   empty cell -> dropped; (a, a) leaf -> P2P; (a, a) internal -> the 10 child pairs c <= d;
//   sqrt(2) (h_a + h_b) < theta d -> M2L; two leaves -> P2P; else split a if b is a leaf or
//   (a is not a leaf and h_a >= h_b), else split b.
// Every generation is one fixed-size grid that reads its frontier size from device
// memory, so the whole traversal is enqueued without a host sync (2 depth + 1 launches).
// Appends use one atomic per wavefront.  The lists are then turned into rows by node
// (td3_fmm_lists) and every row is sorted, so the later sums are deterministic.

#include "td3_common.h"

namespace {

constexpr int kTravThreads = 256;
// grid of every generation (grid-stride over the frontier; any size is correct)
#ifndef TD3_TRAVERSE_BLOCKS
#define TD3_TRAVERSE_BLOCKS 2048
#endif
constexpr unsigned kTravBlocks = TD3_TRAVERSE_BLOCKS;

// Each block takes kTravChunk consecutive frontier pairs per iteration and reserves the
// room for its M2L, P2P and child pairs with one atomic per counter: appends from all
// eight XCDs meet on three counters, so fewer, larger reservations matter.
constexpr int kTravItems = 4;
constexpr u32 kTravChunk = kTravThreads * kTravItems;
constexpr int kTravWaves = kTravThreads / 64;

__device__ __forceinline__ u32 wave_inclusive(u32 v, int lane) {
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const u32 t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

// DUAL: a is a cell of the target tree (ca, ha, ia) and b of the source tree (cb, hb, ib),
// with the rules of kernels/dual_traverse/dual_traverse_pyccel.py:14-86 (no self pairs,
// one-way lists); otherwise both are cells of one tree.
// ctr: [0] m2l count, [1] p2p count, [2] overflow, [3 + g] frontier size of generation g.
template <bool DUAL>
__global__ __launch_bounds__(kTravThreads) void traverse_kernel(
    const double* __restrict__ ca, const double* __restrict__ ha, const int4* __restrict__ ia,
    const double* __restrict__ cb, const double* __restrict__ hb, const int4* __restrict__ ib,
    double theta, int g, const int2* __restrict__ frontier,
    u32 cap_front, int2* __restrict__ next, int2* __restrict__ m2l, u32 cap_m2l,
    int2* __restrict__ p2p, u32 cap_p2p, u32* __restrict__ ctr) {
    __shared__ u32 s_tot[2][kTravWaves];
    __shared__ u32 s_base[3];
    __shared__ u32 s_front;
    // one read for the whole block, so every thread takes the same path to the barriers:
    // the frontier size, or zero when a list or frontier overflowed earlier (the caller
    // then retries with larger buffers)
    if (threadIdx.x == 0) {
        const u32 stop = __hip_atomic_load(&ctr[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const u32 nf = __hip_atomic_load(&ctr[3 + g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_front = stop ? 0u : (nf < cap_front ? nf : cap_front);
    }
    __syncthreads();
    const u32 n_front = s_front;
    const int lane = td3_lane();
    const int wave = threadIdx.x / 64;
    for (u32 base = blockIdx.x * kTravChunk; base < n_front; base += gridDim.x * kTravChunk) {
        int a[kTravItems], b[kTravItems], kind[kTravItems], child[kTravItems];
        u32 n_pair = 0, n_next = 0;  // m2l in the low, p2p in the high 16 bits
#pragma unroll
        for (int r = 0; r < kTravItems; ++r) {
            const u32 k = base + r * kTravThreads + threadIdx.x;
            a[r] = b[r] = child[r] = 0;
            kind[r] = 0;  // 0 none, 1 m2l, 2 p2p, 3 split a (4), 4 split b (4), 5 self split (10)
            if (k < n_front) {
                const int2 pr = frontier[k];
                a[r] = pr.x;
                b[r] = pr.y;
                const int4 fa = ia[pr.x];
                const int4 fb = ib[pr.y];
                if (fa.y != fa.x && fb.y != fb.x) {
                    const bool leaf_a = fa.z < 0, leaf_b = fb.z < 0;
                    if (!DUAL && pr.x == pr.y) {
                        kind[r] = leaf_a ? 2 : 5;
                    } else {
                        const int i = pr.x, j = pr.y;
                        const double dx = __dadd_rn(ca[2 * i], -cb[2 * j]);
                        const double dy = __dadd_rn(ca[2 * i + 1], -cb[2 * j + 1]);
                        const double distance =
                            __dsqrt_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)));
                        const double radii = __dmul_rn(1.4142135623730951, __dadd_rn(ha[i], hb[j]));
                        if (radii < __dmul_rn(theta, distance))
                            kind[r] = 1;
                        else if (leaf_a && leaf_b)
                            kind[r] = 2;
                        else if (leaf_b || (!leaf_a && ha[i] >= hb[j]))
                            kind[r] = 3;
                        else
                            kind[r] = 4;
                    }
                    child[r] = kind[r] == 4 ? fb.z : fa.z;
                }
            }
            n_pair += kind[r] == 1 ? 1u : kind[r] == 2 ? 0x10000u : 0u;
            n_next += kind[r] == 5 ? 10u : kind[r] >= 3 ? 4u : 0u;
        }
        const u32 inc_pair = wave_inclusive(n_pair, lane);
        const u32 inc_next = wave_inclusive(n_next, lane);
        if (lane == 63) {
            s_tot[0][wave] = inc_pair;
            s_tot[1][wave] = inc_next;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            u32 tp = 0, tn = 0;
#pragma unroll
            for (int w = 0; w < kTravWaves; ++w) {
                tp += s_tot[0][w];
                tn += s_tot[1][w];
            }
            s_base[0] = (tp & 0xffffu) ? atomicAdd(&ctr[0], tp & 0xffffu) : 0u;
            s_base[1] = (tp >> 16) ? atomicAdd(&ctr[1], tp >> 16) : 0u;
            s_base[2] = tn ? atomicAdd(&ctr[3 + g + 1], tn) : 0u;
        }
        __syncthreads();
        u32 before_pair = 0, before_next = 0;
#pragma unroll
        for (int w = 0; w < kTravWaves; ++w) {
            before_pair += w < wave ? s_tot[0][w] : 0u;
            before_next += w < wave ? s_tot[1][w] : 0u;
        }
        const u32 excl_pair = before_pair + inc_pair - n_pair;
        u32 pos_m2l = s_base[0] + (excl_pair & 0xffffu);
        u32 pos_p2p = s_base[1] + (excl_pair >> 16);
        u32 pos_next = s_base[2] + before_next + inc_next - n_next;
        __syncthreads();  // s_tot and s_base are rewritten by the next iteration
#pragma unroll
        for (int r = 0; r < kTravItems; ++r) {
            if (kind[r] == 1) {
                if (pos_m2l < cap_m2l)
                    m2l[pos_m2l] = make_int2(a[r], b[r]);
                else
                    ctr[2] = 1u;
                ++pos_m2l;
            } else if (kind[r] == 2) {
                if (pos_p2p < cap_p2p)
                    p2p[pos_p2p] = make_int2(a[r], b[r]);
                else
                    ctr[2] = 1u;
                ++pos_p2p;
            } else if (kind[r] >= 3) {
                const u32 cnt = kind[r] == 5 ? 10u : 4u;
                if (pos_next + cnt > cap_front) {
                    ctr[2] = 1u;
                } else if (kind[r] == 5) {
                    const int first = child[r];
                    u32 o = pos_next;
                    for (int c = 0; c < 4; ++c)
                        for (int d = c; d < 4; ++d) next[o++] = make_int2(first + c, first + d);
                } else if (kind[r] == 3) {
                    for (int c = 0; c < 4; ++c) next[pos_next + c] = make_int2(child[r] + c, b[r]);
                } else {
                    for (int c = 0; c < 4; ++c) next[pos_next + c] = make_int2(a[r], child[r] + c);
                }
                pos_next += cnt;
            }
        }
    }
}

__global__ void traverse_init_kernel(int2* frontier, u32* ctr) {
    frontier[0] = make_int2(0, 0);
    ctr[3] = 1u;
}

// Row lengths of the directed lists.  mode 0: both ways; 1: both ways, a self pair once;
// 2: one way (b under a only).
__device__ __forceinline__ bool both_ways(int mode, int2 p) {
    return mode == 0 || (mode == 1 && p.x != p.y);
}

__global__ void degree_kernel(const int2* __restrict__ pairs, u32 count, int mode,
                              u32* __restrict__ deg) {
    const u32 k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const int2 p = pairs[k];
    atomicAdd(&deg[p.x], 1u);
    if (both_ways(mode, p)) atomicAdd(&deg[p.y], 1u);
}

__global__ void fill_kernel(const int2* __restrict__ pairs, u32 count, int mode,
                            u32* __restrict__ cursor, int* __restrict__ col,
                            int* __restrict__ dst) {
    const u32 k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const int2 p = pairs[k];
    u32 o = atomicAdd(&cursor[p.x], 1u);
    col[o] = p.y;
    if (dst) dst[o] = p.x;
    if (both_ways(mode, p)) {
        o = atomicAdd(&cursor[p.y], 1u);
        col[o] = p.x;
        if (dst) dst[o] = p.y;
    }
}

// Sorts every row of col ascending: one wavefront per row, bitonic in LDS.
constexpr int kSortMax = 4096;
__global__ __launch_bounds__(256) void row_sort_kernel(const u32* __restrict__ row_start, i64 rows,
                                                       int* __restrict__ col, u32* __restrict__ err) {
    __shared__ int s_buf[4][kSortMax];
    const int wave = threadIdx.x / 64, lane = td3_lane();
    int* buf = s_buf[wave];
    for (i64 r = (i64)blockIdx.x * 4 + wave; r < rows; r += (i64)gridDim.x * 4) {
        const u32 s = row_start[r], e = row_start[r + 1];
        const int len = (int)(e - s);
        if (len <= 1) continue;
        if (len <= 64) {
            // bitonic network across the lanes, one entry per lane
            int v = lane < len ? col[s + lane] : 0x7fffffff;
#pragma unroll
            for (int size = 2; size <= 64; size <<= 1) {
#pragma unroll
                for (int stride = size >> 1; stride > 0; stride >>= 1) {
                    const int o = __shfl_xor(v, stride, 64);
                    const bool up = (lane & size) == 0;
                    const bool lower = (lane & stride) == 0;
                    v = (lower == up) ? (v < o ? v : o) : (v > o ? v : o);
                }
            }
            if (lane < len) col[s + lane] = v;
            continue;
        }
        if (len > kSortMax) {
            if (lane == 0) atomicOr(err, 1u);
            continue;
        }
        int pw = 2;
        while (pw < len) pw <<= 1;
        for (int i = lane; i < pw; i += 64) buf[i] = i < len ? col[s + i] : 0x7fffffff;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        for (int size = 2; size <= pw; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int i = lane; i < pw / 2; i += 64) {
                    const int lo = 2 * i - (i & (stride - 1));
                    const int hi = lo + stride;
                    const bool up = (lo & size) == 0;
                    const int x = buf[lo], y = buf[hi];
                    if ((x > y) == up) {
                        buf[lo] = y;
                        buf[hi] = x;
                    }
                }
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            }
        }
        for (int i = lane; i < len; i += 64) col[s + i] = buf[i];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
}

// dst of each directed pair after the row sort (rows keep their node).
__global__ void row_dst_kernel(const u32* __restrict__ row_start, i64 rows, int* __restrict__ dst) {
    const i64 r = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    for (u32 k = row_start[r]; k < row_start[r + 1]; ++k) dst[k] = (int)r;
}

}  // namespace

#include "td3_sort.h"

extern "C" size_t td3_traverse_workspace_bytes(int64_t cap_front) {
    return (size_t)(2 * cap_front) * sizeof(int2) + 4096;
}

extern "C" int td3_traverse(const double* node_center, const double* node_half,
                            const void* node_info, int64_t node_count, int64_t tree_depth,
                            double theta, int64_t cap_m2l, int64_t cap_p2p, int64_t cap_front,
                            int32_t* m2l, int32_t* p2p, void* workspace, size_t workspace_bytes,
                            int64_t* out_counts, void* stream) {
    TD3_ARG(node_count >= 1 && node_count < 0x7fffffff, "node_count");
    TD3_ARG(tree_depth >= 0 && tree_depth <= 64, "tree_depth");
    TD3_ARG(cap_front >= 16 && cap_front < 0x7fffffff && cap_m2l < 0xffffffffll &&
                cap_p2p < 0xffffffffll, "capacities");
    TD3_ARG(workspace_bytes >= td3_traverse_workspace_bytes(cap_front), "workspace too small");
    TD3_ARG(out_counts != nullptr, "out_counts");
    hipStream_t s = (hipStream_t)stream;
    int2* fa = (int2*)workspace;
    int2* fb = fa + cap_front;
    u32* ctr = (u32*)(fb + cap_front);
    const int gens = (int)(2 * tree_depth + 1);
    TD3_HIP(hipMemsetAsync(ctr, 0, sizeof(u32) * (size_t)(gens + 8), s));
    hipLaunchKernelGGL(traverse_init_kernel, dim3(1), dim3(1), 0, s, fa, ctr);
    TD3_LAUNCHED();
    for (int g = 0; g < gens; ++g) {
        TD3_TIMED(s, "traverse", hipLaunchKernelGGL(traverse_kernel<false>, dim3(kTravBlocks), dim3(kTravThreads), 0, s, node_center,
                           node_half, (const int4*)node_info, node_center, node_half,
                           (const int4*)node_info, theta, g, (g & 1) ? fb : fa,
                           (u32)cap_front, (g & 1) ? fa : fb, (int2*)m2l, (u32)cap_m2l, (int2*)p2p,
                           (u32)cap_p2p, ctr));
        TD3_LAUNCHED();
    }
    u32 host[4 + 130];
    TD3_HIP(hipMemcpyAsync(host, ctr, sizeof(u32) * (size_t)(gens + 4), hipMemcpyDeviceToHost, s));
    TD3_HIP(hipStreamSynchronize(s));
    u32 max_front = 0;
    for (int g = 0; g <= gens; ++g) max_front = host[3 + g] > max_front ? host[3 + g] : max_front;
    out_counts[0] = host[0];
    out_counts[1] = host[1];
    out_counts[2] = max_front;
    out_counts[3] = host[2] ? 1 : 0;
    if (host[2] || host[3 + gens] != 0) {
        if (!host[2]) {
            td3_set_error("traversal did not finish in %d generations", gens);
            return TD3_ERR_ARG;
        }
        td3_set_error("traversal lists overflowed (m2l %u, p2p %u, frontier %u)", host[0], host[1],
                      max_front);
        return TD3_ERR_CAPACITY;
    }
    return TD3_OK;
}

extern "C" int td3_dual_traverse(const double* t_center, const double* t_half, const void* t_info,
                                 int64_t t_count, int64_t t_depth, const double* s_center,
                                 const double* s_half, const void* s_info, int64_t s_count,
                                 int64_t s_depth, double theta, int64_t cap_m2l, int64_t cap_p2p,
                                 int64_t cap_front, int32_t* m2l, int32_t* p2p, void* workspace,
                                 size_t workspace_bytes, int64_t* out_counts, void* stream) {
    TD3_ARG(t_count >= 1 && t_count < 0x7fffffff && s_count >= 1 && s_count < 0x7fffffff, "node counts");
    TD3_ARG(t_depth >= 0 && t_depth <= 64 && s_depth >= 0 && s_depth <= 64, "depths");
    TD3_ARG(cap_front >= 16 && cap_front < 0x7fffffff && cap_m2l < 0xffffffffll &&
                cap_p2p < 0xffffffffll, "capacities");
    TD3_ARG(workspace_bytes >= td3_traverse_workspace_bytes(cap_front), "workspace too small");
    TD3_ARG(out_counts != nullptr, "out_counts");
    hipStream_t s = (hipStream_t)stream;
    int2* fa = (int2*)workspace;
    int2* fb = fa + cap_front;
    u32* ctr = (u32*)(fb + cap_front);
    const int gens = (int)(t_depth + s_depth + 1);
    TD3_HIP(hipMemsetAsync(ctr, 0, sizeof(u32) * (size_t)(gens + 8), s));
    hipLaunchKernelGGL(traverse_init_kernel, dim3(1), dim3(1), 0, s, fa, ctr);
    TD3_LAUNCHED();
    for (int g = 0; g < gens; ++g) {
        TD3_TIMED(s, "traverse", hipLaunchKernelGGL(traverse_kernel<true>, dim3(kTravBlocks), dim3(kTravThreads), 0, s,
                           t_center, t_half, (const int4*)t_info, s_center, s_half,
                           (const int4*)s_info, theta, g, (g & 1) ? fb : fa, (u32)cap_front,
                           (g & 1) ? fa : fb, (int2*)m2l, (u32)cap_m2l, (int2*)p2p, (u32)cap_p2p, ctr));
        TD3_LAUNCHED();
    }
    u32 host[4 + 140];
    TD3_HIP(hipMemcpyAsync(host, ctr, sizeof(u32) * (size_t)(gens + 4), hipMemcpyDeviceToHost, s));
    TD3_HIP(hipStreamSynchronize(s));
    u32 max_front = 0;
    for (int g = 0; g <= gens; ++g) max_front = host[3 + g] > max_front ? host[3 + g] : max_front;
    out_counts[0] = host[0];
    out_counts[1] = host[1];
    out_counts[2] = max_front;
    out_counts[3] = host[2] ? 1 : 0;
    if (host[2]) {
        td3_set_error("dual traversal lists overflowed (m2l %u, p2p %u, frontier %u)", host[0],
                      host[1], max_front);
        return TD3_ERR_CAPACITY;
    }
    if (host[3 + gens] != 0) {
        td3_set_error("dual traversal did not finish in %d generations", gens);
        return TD3_ERR_ARG;
    }
    return TD3_OK;
}

extern "C" size_t td3_lists_workspace_bytes(int64_t node_count) {
    return (size_t)(node_count + 1) * sizeof(u32) * 2 + td3sort::scratch_bytes(node_count + 1) + 1024;
}

extern "C" int td3_lists(const int32_t* pairs, int64_t count, int mode, int64_t node_count,
                         uint32_t* row_start, int32_t* col, int32_t* dst, void* workspace,
                         size_t workspace_bytes, void* stream) {
    TD3_ARG(count >= 0 && node_count >= 1, "sizes");
    TD3_ARG(mode >= 0 && mode <= 2, "mode");
    TD3_ARG(workspace_bytes >= td3_lists_workspace_bytes(node_count), "workspace too small");
    hipStream_t s = (hipStream_t)stream;
    u32* deg = (u32*)workspace;
    u32* cursor = deg + (node_count + 1);
    u32* scratch = (u32*)(((uintptr_t)(cursor + node_count + 1) + 255) & ~(uintptr_t)255);
    TD3_HIP(hipMemsetAsync(deg, 0, sizeof(u32) * (size_t)(node_count + 1), s));
    if (count > 0) {
        TD3_TIMED(s, "lists_degree", hipLaunchKernelGGL(degree_kernel, dim3(td3_blocks(count, 256)), dim3(256), 0, s, (const int2*)pairs,
                           (u32)count, mode, deg));
        TD3_LAUNCHED();
    }
    if (int rc = td3sort::exclusive_scan_u32(deg, row_start, node_count + 1, nullptr, scratch, s)) return rc;
    if (count > 0) {
        TD3_HIP(hipMemcpyAsync(cursor, row_start, sizeof(u32) * (size_t)node_count, hipMemcpyDeviceToDevice, s));
        TD3_TIMED(s, "lists_fill", hipLaunchKernelGGL(fill_kernel, dim3(td3_blocks(count, 256)), dim3(256), 0, s, (const int2*)pairs,
                           (u32)count, mode, cursor, col, (int*)nullptr));
        TD3_LAUNCHED();
        u32* err = deg;  // reuse: zero it first
        TD3_HIP(hipMemsetAsync(err, 0, sizeof(u32), s));
        unsigned blocks = td3_blocks(node_count, 4);
        if (blocks > 8192) blocks = 8192;
        TD3_TIMED(s, "row_sort", hipLaunchKernelGGL(row_sort_kernel, dim3(blocks), dim3(256), 0, s, row_start, node_count, col, err));
        TD3_LAUNCHED();
        if (dst) {
            hipLaunchKernelGGL(row_dst_kernel, dim3(td3_blocks(node_count, 256)), dim3(256), 0, s, row_start,
                               node_count, dst);
            TD3_LAUNCHED();
        }
        u32 herr = 0;
        TD3_HIP(hipMemcpyAsync(&herr, err, sizeof(u32), hipMemcpyDeviceToHost, s));
        TD3_HIP(hipStreamSynchronize(s));
        if (herr) {
            td3_set_error("an interaction row holds more than %d cells", kSortMax);
            return TD3_ERR_ARG;
        }
    }
    return TD3_OK;
}
