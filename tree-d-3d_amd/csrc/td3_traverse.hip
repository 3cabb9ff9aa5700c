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

// The directed entries of a pair list as sort keys (row node << sb) | column.  mode 0:
// both ways; 1: both ways, a self pair once (its second slot a padding key, all ones, which
// sorts after every entry); 2: one way (b under a only).
__global__ void list_keys_kernel(const int2* __restrict__ pairs, u32 count, int mode, int sb,
                                 u64* __restrict__ keys) {
    const u32 k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const int2 p = pairs[k];
    const u64 ab = ((u64)(u32)p.x << sb) | (u32)p.y;
    if (mode == 2) {
        keys[k] = ab;
        return;
    }
    keys[2 * (u64)k] = ab;
    keys[2 * (u64)k + 1] = (mode == 0 || p.x != p.y) ? (((u64)(u32)p.y << sb) | (u32)p.x) : ~0ull;
}

// Entry k's column (and row node) from the sorted keys (padding keys, past the entries, are
// skipped).
__global__ void list_cols_kernel(const u64* __restrict__ keys, i64 n, int sb, i64 node_count,
                                 int* __restrict__ col, int* __restrict__ dst) {
    const i64 k = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const u64 key = keys[k];
    if ((key >> sb) >= (u64)node_count) return;
    col[k] = (int)(key & ((1ull << sb) - 1ull));
    if (dst) dst[k] = (int)(key >> sb);
}

// row_start[r], r <= node_count: the first sorted key whose row node is r or more (a
// padding key counts as row node_count), by binary search.
__global__ void list_starts_kernel(const u64* __restrict__ keys, i64 n, int sb, i64 node_count,
                                   u32* __restrict__ row_start) {
    const i64 r = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (r > node_count) return;
    i64 lo = 0, hi = n;
    while (lo < hi) {
        const i64 mid = (lo + hi) >> 1;
        const u64 d = keys[mid] >> sb;
        if (d < (u64)r) lo = mid + 1;
        else hi = mid;
    }
    row_start[r] = (u32)lo;
}

__global__ void list_empty_kernel(i64 node_count, u32* __restrict__ row_start) {
    const i64 r = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (r <= node_count) row_start[r] = 0u;
}

static inline int bits_for(i64 v) {  // bits of v (at least 1)
    int b = 1;
    while (b < 63 && (v >> b) != 0) ++b;
    return b;
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

static inline i64 list_entries(i64 count, int mode) { return mode == 2 ? count : 2 * count; }

static inline size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

extern "C" size_t td3_lists_workspace_bytes(int64_t node_count, int64_t count, int mode) {
    const i64 n = list_entries(count < 0 ? 0 : count, mode) + 1;
    return 2 * align256((size_t)n * sizeof(u64)) + 2 * align256((size_t)n * sizeof(u32)) +
           align256(td3sort::scratch_bytes(n)) + 256;
}

extern "C" int td3_lists(const int32_t* pairs, int64_t count, int mode, int64_t node_count,
                         int64_t col_count, uint32_t* row_start, int32_t* col, int32_t* dst,
                         void* workspace, size_t workspace_bytes, void* stream) {
    TD3_ARG(count >= 0 && node_count >= 1 && col_count >= 1, "sizes");
    TD3_ARG(mode >= 0 && mode <= 2, "mode");
    TD3_ARG(mode == 2 || col_count == node_count, "col_count (the same tree for modes 0, 1)");
    TD3_ARG(workspace_bytes >= td3_lists_workspace_bytes(node_count, count, mode),
            "workspace too small");
    hipStream_t s = (hipStream_t)stream;
    const i64 n = list_entries(count, mode);
    if (n == 0) {
        hipLaunchKernelGGL(list_empty_kernel, dim3(td3_blocks(node_count + 1, 256)), dim3(256), 0,
                           s, node_count, row_start);
        TD3_LAUNCHED();
        return TD3_OK;
    }
    // keys: row node above column, node_count < 2^db so that a padding key (all ones) sorts
    // after every entry
    const int sb = bits_for(col_count - 1), db = bits_for(node_count);
    TD3_ARG(sb + db <= 64, "sizes (keys of more than 64 bits)");
    char* base = (char*)(((uintptr_t)workspace + 255) & ~(uintptr_t)255);
    u64* keys = (u64*)base;
    u64* keys_alt = (u64*)(base + align256((size_t)n * sizeof(u64)));
    u32* vals = (u32*)(base + 2 * align256((size_t)n * sizeof(u64)));
    u32* vals_alt = (u32*)((char*)vals + align256((size_t)n * sizeof(u32)));
    u32* scratch = (u32*)((char*)vals_alt + align256((size_t)n * sizeof(u32)));
    TD3_TIMED(s, "lists_keys", hipLaunchKernelGGL(list_keys_kernel, dim3(td3_blocks(count, 256)),
              dim3(256), 0, s, (const int2*)pairs, (u32)count, mode, sb, keys));
    TD3_LAUNCHED();
    if (int rc = td3sort::sort_pairs(keys, vals, nullptr, keys_alt, vals_alt, nullptr, n, 0,
                                     sb + db, true, scratch, s))
        return rc;
    TD3_TIMED(s, "lists_rows", hipLaunchKernelGGL(list_cols_kernel, dim3(td3_blocks(n, 256)),
              dim3(256), 0, s, keys, n, sb, node_count, col, dst));
    TD3_LAUNCHED();
    TD3_TIMED(s, "lists_rows", hipLaunchKernelGGL(list_starts_kernel,
              dim3(td3_blocks(node_count + 1, 256)), dim3(256), 0, s, keys, n, sb, node_count,
              row_start));
    TD3_LAUNCHED();
    return TD3_OK;
}
