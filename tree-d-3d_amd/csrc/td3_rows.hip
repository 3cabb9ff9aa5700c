// Interaction rows by node, straight from the traversal rules.
//
// The reference's traversal (flat_tree.py:1318-1358, kernels/fmm_traverse/
// fmm_traverse_pyccel.py:15-106) walks unordered cell pairs from (root, root) and emits
// M2L and P2P pairs; the rows by node come after (kernels/pair_csr/pair_csr_pyccel.py).
// Here every node a gathers its own rows: the pairs {a, b} the traversal visits are, for a
// fixed a, those inherited from a's parent (pairs the parent split on its own side, and
// the sibling pairs of a self pair) closed under splitting the other side.  So, level by
// level, one wavefront per node takes its inherited entries (one lane each) and walks
// them depth first:
//   empty cell -> dropped; self pair: leaf -> P2P, else the sibling pairs go to the
//   children; MAC (sqrt(2)(h_a + h_b) < theta d) -> M2L; two leaves -> P2P; else the rule
//   splits the first cell if the second is a leaf or (the first is not a leaf and
//   h_first >= h_second), else the second: splitting a hands (b, side) to a's four
//   children, splitting b continues the walk with b's children.
// The side (a first or second) follows each pair from its creation, as in the reference,
// because the tie h_first == h_second splits the first cell.  Each node's rows are the
// concatenation, in lane order, of its lanes' walks: deterministic without a sort, and
// without atomics.  The pair sets are those of td3_traverse (both rows of a mutual pair
// come from the same decision, evaluated with the same FP64 operations).
//
// Per level three steps: count (each lane walks and counts M2L, P2P and hand-downs per
// child), exclusive scans (row starts of this level, inherited-list starts of the next),
// write (each lane walks again and writes at offsets from the stored counts).
// A walk needs no stack: it moves through the other cell's subtree by child and parent
// links, so any depth works.

#include "td3_common.h"
#include "td3_sort.h"

namespace {

constexpr int kRowsWaves = 4;
constexpr u32 kSecond = 0x80000000u;

struct RowsTree {
    const double* center;
    const double* half;
    const int4* info;  // {start, end, child, depth}
    const int* parent;
};

// The decision for the pair {a, x} (a is `a_second` ? second : first).  Returns 0 drop,
// 1 M2L, 2 P2P, 3 split a, 4 split x.
__device__ __forceinline__ int decide(const RowsTree& t, int a, int4 fa, int x, bool a_second,
                                      double theta) {
    const int4 fx = t.info[x];
    if (fx.y == fx.x || fa.y == fa.x) return 0;
    const bool leaf_a = fa.z < 0, leaf_x = fx.z < 0;
    const int f = a_second ? x : a, s = a_second ? a : x;
    const double dx = __dadd_rn(t.center[2 * f], -t.center[2 * s]);
    const double dy = __dadd_rn(t.center[2 * f + 1], -t.center[2 * s + 1]);
    const double distance = __dsqrt_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)));
    const double radii = __dmul_rn(1.4142135623730951, __dadd_rn(t.half[f], t.half[s]));
    if (radii < __dmul_rn(theta, distance)) return 1;
    if (leaf_a && leaf_x) return 2;
    const bool leaf_f = a_second ? leaf_x : leaf_a, leaf_s = a_second ? leaf_a : leaf_x;
    const bool split_first = leaf_s || (!leaf_f && t.half[f] >= t.half[s]);
    const bool split_a = split_first != a_second;
    return split_a ? 3 : 4;
}

struct Counts {
    u32 m2l, p2p, inh[4];
};

__device__ __forceinline__ u32 entry_with(int x, bool second) {
    return (u32)x | (second ? kSecond : 0u);
}

// Walks one inherited entry of node a.  WRITE: emit at the given offsets (advanced).
template <bool WRITE>
__device__ void walk_entry(const RowsTree& t, int a, int4 fa, u32 entry, double theta,
                           Counts* c, int* m2l_col, int* p2p_col, u32* inh_out,
                           const u32* child_base) {
    if ((int)(entry & ~kSecond) == a) {
        // self pair {a, a}
        if (fa.y == fa.x) return;
        if (fa.z < 0) {
            if (WRITE) p2p_col[c->p2p] = a;
            ++c->p2p;
            return;
        }
        // the ten child pairs (c_i, c_j), i <= j: child k gets every sibling, itself as a
        // self pair, c_j (j > k) with k first, c_i (i < k) with k second
        for (int k = 0; k < 4; ++k) {
            if (WRITE) {
                u32* out = inh_out + child_base[k] + c->inh[k];
                for (int j = 0; j < 4; ++j)
                    out[j] = (u32)(fa.z + j) | (j < k ? kSecond : 0u);
            }
            c->inh[k] += 4;
        }
        return;
    }
    // depth first over the subtree of the other cell x0, without a stack: children in
    // order, back up through the parents once a fourth child is done
    const int x0 = (int)(entry & ~kSecond);
    const bool second = (entry & kSecond) != 0;
    int x = x0;
    while (true) {
        const int kind = decide(t, a, fa, x, second, theta);
        if (kind == 4) {
            x = t.info[x].z;
            continue;
        }
        if (kind == 1) {
            if (WRITE) m2l_col[c->m2l] = x;
            ++c->m2l;
        } else if (kind == 2) {
            if (WRITE) p2p_col[c->p2p] = x;
            ++c->p2p;
        } else if (kind == 3) {
            for (int k = 0; k < 4; ++k) {
                if (WRITE) inh_out[child_base[k] + c->inh[k]] = entry_with(x, second);
                ++c->inh[k];
            }
        }
        while (x != x0 && x - t.info[t.parent[x]].z == 3) x = t.parent[x];
        if (x == x0) break;
        ++x;
    }
}

__device__ __forceinline__ u32 wave_sum_u32(u32 v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ u32 wave_excl(u32 v, int lane, u32* total) {
    u32 incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const u32 y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
    }
    *total = __shfl(incl, 63, 64);
    return incl - v;
}

// Count step: per node of the level, its M2L and P2P row lengths and the number of
// entries each child inherits.
__global__ __launch_bounds__(64 * kRowsWaves) void rows_count_kernel(
    RowsTree t, i64 lo, i64 hi, const u32* __restrict__ inh, const u32* __restrict__ inh_start,
    double theta, u32* __restrict__ m2l_n, u32* __restrict__ p2p_n, u32* __restrict__ inh_n,
    i64 next_lo, u32 cap_inh, u32* __restrict__ ecount, u32* __restrict__ flag) {
    const int wave = threadIdx.x / 64, lane = td3_lane();
    const i64 a = lo + (i64)blockIdx.x * kRowsWaves + wave;
    if (a >= hi) return;  // uniform over the wavefront
    // after an overflow at an earlier level the lists are incomplete: stop (the caller
    // retries with more room, or takes the pair-list path)
    if (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    const int4 fa = t.info[a];
    // an overflowed list (flagged by the write step) is read only up to the capacity
    const u32 s0 = inh_start[a - lo], s1 = inh_start[a - lo + 1];
    const u32 e0 = s0 < cap_inh ? s0 : cap_inh, e1 = s1 < cap_inh ? s1 : cap_inh;
    Counts tot = {0, 0, {0, 0, 0, 0}};
    for (u32 b = e0; b < e1; b += 64) {
        Counts c = {0, 0, {0, 0, 0, 0}};
        if (b + lane < e1) {
            walk_entry<false>(t, (int)a, fa, inh[b + lane], theta, &c, nullptr, nullptr,
                              nullptr, nullptr);
            u32* ec = ecount + 6 * (size_t)(b + lane);
            ec[0] = c.m2l;
            ec[1] = c.p2p;
            for (int k = 0; k < 4; ++k) ec[2 + k] = c.inh[k];
        }
        tot.m2l += wave_sum_u32(c.m2l);
        tot.p2p += wave_sum_u32(c.p2p);
#pragma unroll
        for (int k = 0; k < 4; ++k) tot.inh[k] += wave_sum_u32(c.inh[k]);
    }
    if (lane == 0) {
        m2l_n[a - lo] = tot.m2l;
        p2p_n[a - lo] = tot.p2p;
        if (fa.z >= 0)
            for (int k = 0; k < 4; ++k) inh_n[fa.z + k - next_lo] = tot.inh[k];
    }
}

// Write step: every lane takes its entry's counts from the count step (its offsets inside
// the node's rows and its children's lists) and walks the entry again, writing.
__global__ __launch_bounds__(64 * kRowsWaves) void rows_write_kernel(
    RowsTree t, i64 lo, i64 hi, const u32* __restrict__ inh, const u32* __restrict__ inh_start,
    double theta, const u32* __restrict__ m2l_scan, const u32* __restrict__ p2p_scan,
    const u32* __restrict__ base, int level, u32* __restrict__ m2l_row, int* __restrict__ m2l_col,
    u32* __restrict__ p2p_row, int* __restrict__ p2p_col, u32* __restrict__ inh_out,
    const u32* __restrict__ inh_out_start, i64 next_lo, u32 cap_m2l, u32 cap_p2p,
    u32 cap_inh, const u32* __restrict__ ecount, u32* __restrict__ flag) {
    const int wave = threadIdx.x / 64, lane = td3_lane();
    const i64 a = lo + (i64)blockIdx.x * kRowsWaves + wave;
    if (a >= hi) return;  // uniform over the wavefront
    if (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    const int4 fa = t.info[a];
    const u32 s0 = inh_start[a - lo], s1 = inh_start[a - lo + 1];
    const u32 e0 = s0 < cap_inh ? s0 : cap_inh, e1 = s1 < cap_inh ? s1 : cap_inh;
    u32 m2l_at = base[2 * level] + m2l_scan[a - lo];
    u32 p2p_at = base[2 * level + 1] + p2p_scan[a - lo];
    if (lane == 0) {
        m2l_row[a] = m2l_at;
        p2p_row[a] = p2p_at;
    }
    u32 child_at[4] = {0, 0, 0, 0};
    if (fa.z >= 0)
        for (int k = 0; k < 4; ++k) child_at[k] = inh_out_start[fa.z + k - next_lo];
    for (u32 b = e0; b < e1; b += 64) {
        const bool mine = b + lane < e1;
        const u32 entry = mine ? inh[b + lane] : 0u;
        Counts c = {0, 0, {0, 0, 0, 0}};
        if (mine) {
            const u32* ec = ecount + 6 * (size_t)(b + lane);
            c.m2l = ec[0];
            c.p2p = ec[1];
            for (int k = 0; k < 4; ++k) c.inh[k] = ec[2 + k];
        }
        u32 tm, tp, ti[4];
        const u32 om = wave_excl(c.m2l, lane, &tm);
        const u32 op = wave_excl(c.p2p, lane, &tp);
        u32 oi[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) oi[k] = wave_excl(c.inh[k], lane, &ti[k]);
        const bool fits = m2l_at + tm <= cap_m2l && p2p_at + tp <= cap_p2p &&
                          (fa.z < 0 || (child_at[0] + ti[0] <= cap_inh && child_at[1] + ti[1] <= cap_inh &&
                                        child_at[2] + ti[2] <= cap_inh && child_at[3] + ti[3] <= cap_inh));
        if (!fits) {
            if (lane == 0) atomicOr(flag, 2u);
            return;
        }
        if (mine) {
            Counts w = {m2l_at + om, p2p_at + op, {oi[0], oi[1], oi[2], oi[3]}};
            walk_entry<true>(t, (int)a, fa, entry, theta, &w, m2l_col, p2p_col, inh_out,
                             child_at);
        }
        m2l_at += tm;
        p2p_at += tp;
#pragma unroll
        for (int k = 0; k < 4; ++k) child_at[k] += ti[k];
    }
}

__global__ void rows_parent_kernel(const int4* __restrict__ info, i64 count, int* __restrict__ parent) {
    const i64 node = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= count) return;
    if (node == 0) parent[0] = -1;
    const int c = info[node].z;
    if (c >= 0)
        for (int k = 0; k < 4; ++k) parent[c + k] = (int)node;
}

// base[2 (d + 1) + j] = base[2 d + j] + total_j (row starts of the next level).
__global__ void rows_base_kernel(u32* base, int level, const u32* totals) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        base[2 * level + 2] = base[2 * level] + totals[0];
        base[2 * level + 3] = base[2 * level + 1] + totals[1];
    }
}

__global__ void rows_root_kernel(u32* inh, u32* inh_start, u32* base) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        inh[0] = 0u;  // the pair (root, root)
        inh_start[0] = 0u;
        inh_start[1] = 1u;
        base[0] = 0u;
        base[1] = 0u;
    }
}

// dst of every M2L entry (the row's node), for the lane-per-pair FP64 M2L
__global__ void rows_dst_kernel(const u32* __restrict__ row, i64 count, int* __restrict__ dst) {
    const i64 r = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= count) return;
    for (u32 k = row[r]; k < row[r + 1]; ++k) dst[k] = (int)r;
}

__global__ void rows_end_kernel(const u32* base, int levels, i64 count, u32* m2l_row,
                                u32* p2p_row) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        m2l_row[count] = base[2 * levels];
        p2p_row[count] = base[2 * levels + 1];
    }
}

struct RowsWs {
    u32 *inh_a, *inh_b, *start_a, *start_b, *m2l_n, *p2p_n, *m2l_scan, *p2p_scan, *totals,
        *base, *flag, *scratch, *ecount;
    int* parent;
};

size_t rows_carve(RowsWs* w, void* p, i64 count, i64 width, i64 cap_inh, int levels) {
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return p ? (void*)((char*)p + o) : nullptr;
    };
    w->inh_a = (u32*)take(sizeof(u32) * (size_t)cap_inh);
    w->inh_b = (u32*)take(sizeof(u32) * (size_t)cap_inh);
    w->start_a = (u32*)take(sizeof(u32) * (size_t)(width + 1));
    w->start_b = (u32*)take(sizeof(u32) * (size_t)(width + 1));
    w->m2l_n = (u32*)take(sizeof(u32) * (size_t)(width + 1));
    w->p2p_n = (u32*)take(sizeof(u32) * (size_t)(width + 1));
    w->m2l_scan = (u32*)take(sizeof(u32) * (size_t)(width + 1));
    w->p2p_scan = (u32*)take(sizeof(u32) * (size_t)(width + 1));
    w->totals = (u32*)take(sizeof(u32) * 4);
    w->base = (u32*)take(sizeof(u32) * (size_t)(2 * levels + 4));
    w->flag = (u32*)take(sizeof(u32) * 4);
    w->scratch = (u32*)take(td3sort::scratch_bytes(width + 1));
    w->parent = (int*)take(sizeof(int) * (size_t)count);
    w->ecount = (u32*)take(sizeof(u32) * 6 * (size_t)cap_inh);
    return off;
}

i64 widest_level(const int64_t* levels, int n_levels) {
    i64 w = 1;
    for (int d = 0; d < n_levels; ++d) w = levels[d + 1] - levels[d] > w ? levels[d + 1] - levels[d] : w;
    return w;
}

}  // namespace

extern "C" size_t td3_rows_workspace_bytes(const int64_t* levels_host, int64_t n_levels,
                                           int64_t cap_inherit) {
    RowsWs w;
    const i64 width = widest_level(levels_host, (int)n_levels);
    return rows_carve(&w, nullptr, levels_host[n_levels], width, cap_inherit, (int)n_levels);
}

extern "C" int td3_rows(const double* node_center, const double* node_half, const void* node_info,
                        const int64_t* levels_host, int64_t n_levels, double theta,
                        int64_t cap_m2l, int64_t cap_p2p, int64_t cap_inherit,
                        uint32_t* m2l_row, int32_t* m2l_col, int32_t* m2l_dst, uint32_t* p2p_row,
                        int32_t* p2p_col, void* workspace, size_t workspace_bytes,
                        int64_t* out_counts, void* stream) {
    TD3_ARG(levels_host != nullptr && n_levels >= 1 && n_levels <= 64, "levels");
    const i64 count = levels_host[n_levels];
    TD3_ARG(count >= 1 && count < 0x7fffffff, "node count");
    TD3_ARG(cap_m2l >= 1 && cap_m2l < 0xffffffffll && cap_p2p >= 1 && cap_p2p < 0xffffffffll &&
                cap_inherit >= 16 && cap_inherit < 0xffffffffll, "capacities");
    TD3_ARG(out_counts != nullptr, "out_counts");
    TD3_ARG(workspace_bytes >= td3_rows_workspace_bytes(levels_host, n_levels, cap_inherit),
            "workspace too small");
    hipStream_t s = (hipStream_t)stream;
    RowsWs w;
    const i64 width = widest_level(levels_host, (int)n_levels);
    rows_carve(&w, workspace, count, width, cap_inherit, (int)n_levels);
    RowsTree t;
    t.center = node_center;
    t.half = node_half;
    t.info = (const int4*)node_info;
    t.parent = w.parent;
    TD3_HIP(hipMemsetAsync(w.flag, 0, sizeof(u32) * 4, s));
    hipLaunchKernelGGL(rows_parent_kernel, dim3(td3_blocks(count, 256)), dim3(256), 0, s,
                       (const int4*)node_info, count, w.parent);
    TD3_LAUNCHED();
    hipLaunchKernelGGL(rows_root_kernel, dim3(1), dim3(1), 0, s, w.inh_a, w.start_a, w.base);
    TD3_LAUNCHED();
    u32 *inh = w.inh_a, *inh_next = w.inh_b, *start = w.start_a, *start_next = w.start_b;
    for (int d = 0; d < n_levels; ++d) {
        const i64 lo = levels_host[d], hi = levels_host[d + 1];
        const i64 next_lo = hi, next_hi = d + 1 < n_levels ? levels_host[d + 2] : hi;
        const i64 nodes = hi - lo, next = next_hi - next_lo;
        const unsigned blocks = (unsigned)((nodes + kRowsWaves - 1) / kRowsWaves);
        if (next > 0) TD3_HIP(hipMemsetAsync(start_next, 0, sizeof(u32) * (size_t)(next + 1), s));
        TD3_TIMED(s, "rows_count", hipLaunchKernelGGL(rows_count_kernel, dim3(blocks), dim3(64 * kRowsWaves), 0, s, t, lo, hi,
                           inh, start, theta, w.m2l_n, w.p2p_n, start_next, next_lo,
                           (u32)cap_inherit, w.ecount, w.flag));
        TD3_LAUNCHED();
        if (int rc = td3sort::exclusive_scan_u32(w.m2l_n, w.m2l_scan, nodes, w.totals, w.scratch, s)) return rc;
        if (int rc = td3sort::exclusive_scan_u32(w.p2p_n, w.p2p_scan, nodes, w.totals + 1, w.scratch, s)) return rc;
        if (next > 0) {
            // inherited-list starts of the next level, and the total at [next]
            if (int rc = td3sort::exclusive_scan_u32(start_next, start_next, next + 1, nullptr, w.scratch, s)) return rc;
        }
        TD3_TIMED(s, "rows_write", hipLaunchKernelGGL(rows_write_kernel, dim3(blocks), dim3(64 * kRowsWaves), 0, s, t, lo, hi,
                           inh, start, theta, w.m2l_scan, w.p2p_scan, w.base, d, m2l_row,
                           (int*)m2l_col, p2p_row, (int*)p2p_col, inh_next, start_next, next_lo,
                           (u32)cap_m2l, (u32)cap_p2p, (u32)cap_inherit, w.ecount, w.flag));
        TD3_LAUNCHED();
        hipLaunchKernelGGL(rows_base_kernel, dim3(1), dim3(1), 0, s, w.base, d, w.totals);
        TD3_LAUNCHED();
        u32* tmp = inh;
        inh = inh_next;
        inh_next = tmp;
        tmp = start;
        start = start_next;
        start_next = tmp;
    }
    hipLaunchKernelGGL(rows_end_kernel, dim3(1), dim3(1), 0, s, w.base, (int)n_levels, count,
                       m2l_row, p2p_row);
    TD3_LAUNCHED();
    u32 host[4];
    TD3_HIP(hipMemcpyAsync(host, w.flag, sizeof(u32) * 2, hipMemcpyDeviceToHost, s));
    TD3_HIP(hipMemcpyAsync(host + 2, w.base + 2 * n_levels, sizeof(u32) * 2, hipMemcpyDeviceToHost, s));
    TD3_HIP(hipStreamSynchronize(s));
    out_counts[0] = host[2];
    out_counts[1] = host[3];
    out_counts[2] = 0;
    out_counts[3] = host[0] & 2u ? 1 : 0;
    if (host[0] & 2u) {
        td3_set_error("interaction rows overflowed (m2l %u, p2p %u)", host[2], host[3]);
        return TD3_ERR_CAPACITY;
    }
    if (m2l_dst != nullptr && host[2] > 0) {
        hipLaunchKernelGGL(rows_dst_kernel, dim3(td3_blocks(count, 256)), dim3(256), 0, s, m2l_row,
                           count, (int*)m2l_dst);
        TD3_LAUNCHED();
    }
    return TD3_OK;
}
