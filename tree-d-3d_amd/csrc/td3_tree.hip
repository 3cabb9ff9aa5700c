// Bit-exact flat quadtree build.
//
// Replaces the reference's per-level build (flat_tree.py:532-625: quadtree_quadrants +
// a (N, 4) cumsum + quadtree_scatter per level, with a host sync per level) and its
// serial CPU build (kernels/quadtree_build/quadtree_build_pyccel.py:12-137), producing
// identical arrays.
//
// 1. keys: every particle descends from the root, taking at level L the digit
//    q = [x > cx] + 2 [y < cy] of the FP64 comparison against the centre of its
//    depth-(L-1) cell, then moving to the child centre c +- h/2 - the same FP64
//    operations the reference does per level, so the digit path is exact.  Digits are
//    packed most significant first: levels 1-32 in `hi`, 33-64 in `lo`.
// 2. a stable LSD radix sort by key (values = input index): every cell is a contiguous
//    run, cells in breadth-first order within each depth.
// 3. topology without a level loop: with L_i the common-prefix depth of sorted keys
//    i and i+1 and M_i that of keys i and i+cap, particle i starts a splitting node at
//    every depth d with L_{i-1} < d <= min(M_i, max_depth-1).  Those (d, i) entries,
//    bucketed by d, number the splitting nodes of each depth in order; node k's four
//    children are levels[d+1] + 4 rank + q.  One thread per splitting node writes its
//    children (centres by the same exact descent, ranges by binary search on the
//    digit of level d+1, child index by binary search in the next bucket).
// 4. particles keep input order inside each leaf: a stable radix sort of the input
//    indices by the start of their leaf gives `perm`.
// One host read-back (the splitting-node count per depth) sizes the node arrays.
//
// A forest of S independent systems (an ensemble) is built by the same pass: a system's
// index heads its particles' keys (sbits bits, so levels 1..l1 = (64 - sbits) / 2 sit in
// `hi` after it), particles of different systems share no prefix, and the S roots are
// nodes 0..S-1; below them the numbering is breadth first over all systems at once.
// With S = 1 and sbits = 0 this is exactly the single tree.

#include "td3_sort.h"

namespace {

constexpr int kMaxLevels = 64;
constexpr unsigned kCountBlocks = 512;  // count_kernel's grid at most

// Key layout: the system index in the top sbits bits of hi, then levels 1..l1 (2 bits
// each) in hi, levels l1+1.. in lo from its top.
struct Keys {
    int sbits, l1;
};

__device__ __forceinline__ int lcp_levels(u64 ahi, u64 alo, u64 bhi, u64 blo, int nkey, Keys kl) {
    u64 x = ahi ^ bhi;
    if (kl.sbits && (x >> (64 - kl.sbits))) return -1;  // different systems: no common cell
    x <<= kl.sbits;
    int l;
    if (x)
        l = __clzll(x) >> 1;
    else if ((x = alo ^ blo))
        l = kl.l1 + (__clzll(x) >> 1);
    else
        l = 64;
    return l < nkey ? l : nkey;
}

__device__ __forceinline__ int digit_at(u64 hi, u64 lo, int level, Keys kl) {
    return level <= kl.l1 ? (int)((hi >> (64 - kl.sbits - 2 * level)) & 3ull)
                          : (int)((lo >> (64 - 2 * (level - kl.l1))) & 3ull);
}

__device__ __forceinline__ int system_of(u64 hi, Keys kl) {
    return kl.sbits ? (int)(hi >> (64 - kl.sbits)) : 0;
}

// Root (centre, half) of a system: roots (S, 3) when given, else the single root.
// y_down: a point with y == cy goes to the lower children (the static cells of
// smooth.py:109-110) instead of the upper ones (quadtree_quadrants_pyccel.py:38-41).
// check: the keys kernel reports positions that are not finite (bit 0) or outside the
// closed root square (bit 1), the reference's checks (flat_tree.py:785-794).
struct Root {
    const double* roots;
    double cx, cy, h;
    int y_down;
    int check;
};

__device__ __forceinline__ void root_of(Root rt, int sys, double* cx, double* cy, double* h) {
    if (rt.roots) {
        *cx = rt.roots[3 * sys];
        *cy = rt.roots[3 * sys + 1];
        *h = rt.roots[3 * sys + 2];
    } else {
        *cx = rt.cx;
        *cy = rt.cy;
        *h = rt.h;
    }
}

__global__ void keys_kernel(const double* __restrict__ pos, i64 n, Root rt,
                            const int* __restrict__ psys, Keys kl, int nlev,
                            u64* __restrict__ key_hi, u64* __restrict__ key_lo,
                            u32* __restrict__ report) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = pos[i];
    const double y = pos[n + i];
    const int sys = psys ? psys[i] : 0;
    double cx, cy, h;
    root_of(rt, sys, &cx, &cy, &h);
    if (rt.check) {
        if (!(x - x == 0.0) || !(y - y == 0.0))
            atomicOr(report, 1u);
        else if (x < cx - h || x > cx + h || y < cy - h || y > cy + h)
            atomicOr(report, 2u);
    }
    u64 hi = kl.sbits ? (u64)sys << (64 - kl.sbits) : 0ull, lo = 0;
    for (int level = 1; level <= nlev; ++level) {
        const int q = (x > cx ? 1 : 0) + ((y < cy || (rt.y_down && y == cy)) ? 2 : 0);
        h = 0.5 * h;
        cx = (q & 1) ? cx + h : cx - h;
        cy = (q & 2) ? cy - h : cy + h;
        if (level <= kl.l1)
            hi |= (u64)q << (64 - kl.sbits - 2 * level);
        else
            lo |= (u64)q << (64 - 2 * (level - kl.l1));
    }
    key_hi[i] = hi;
    if (key_lo != nullptr) key_lo[i] = lo;
}

// Splitting-node starts per sorted particle, and their count per depth.
__global__ __launch_bounds__(256) void count_kernel(const u64* __restrict__ khi,
                                                    const u64* __restrict__ klo, i64 n, i64 cap,
                                                    int max_split_depth, int nkey, Keys kl,
                                                    u32* __restrict__ cnt,
                                                    u32* __restrict__ depth_count,
                                                    u32* __restrict__ flag) {
    __shared__ u32 s_depth[kMaxLevels + 1];
    __shared__ u32 s_deeper;
    if (threadIdx.x <= kMaxLevels) s_depth[threadIdx.x] = 0;
    if (threadIdx.x == 0) s_deeper = 0;
    __syncthreads();
    bool deeper = false;
    // a grid of at most kCountBlocks blocks strides over the particles: the blocks' totals
    // go to a few depth counters with one atomic each, and same-address atomics serialise
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (i64)gridDim.x * blockDim.x) {
        const u64 ahi = khi[i], alo = klo ? klo[i] : 0ull;
        int prev = -1;
        if (i > 0) prev = lcp_levels(khi[i - 1], klo ? klo[i - 1] : 0ull, ahi, alo, nkey, kl);
        int m = -1;
        if (i + cap < n) m = lcp_levels(ahi, alo, khi[i + cap], klo ? klo[i + cap] : 0ull, nkey, kl);
        const int top = m < max_split_depth ? m : max_split_depth;
        const int c = top > prev ? top - prev : 0;
        cnt[i] = (u32)c;
        for (int d = prev + 1; d <= top; ++d) atomicAdd(&s_depth[d], 1u);
        // a common prefix as long as the key while cells that deep may still split:
        // the key must resolve more levels
        if (m >= nkey && nkey <= max_split_depth) deeper = true;
    }
    if (deeper) s_deeper = 1u;
    __syncthreads();
    if (threadIdx.x <= kMaxLevels && s_depth[threadIdx.x])
        atomicAdd(&depth_count[threadIdx.x], s_depth[threadIdx.x]);
    if (threadIdx.x == 0 && s_deeper) atomicOr(flag, 1u);
}

// One (depth, particle) entry per splitting node, in particle order.
__global__ void entries_kernel(const u64* __restrict__ khi, const u64* __restrict__ klo, i64 n,
                               int nkey, Keys kl, const u32* __restrict__ cnt, const u32* __restrict__ off,
                               i64 ne_cap, u64* __restrict__ ekey, u32* __restrict__ eval) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const u32 c = cnt[i];
    if (!c) return;
    int prev = -1;
    if (i > 0) prev = lcp_levels(khi[i - 1], klo ? klo[i - 1] : 0ull, khi[i], klo ? klo[i] : 0ull, nkey, kl);
    const u32 o = off[i];
    for (u32 j = 0; j < c; ++j) {
        if ((i64)o + j >= ne_cap) return;  // more splitting nodes than the capacity holds
        ekey[o + j] = (u64)(prev + 1 + (int)j);
        eval[o + j] = (u32)i;
    }
}

// Level offsets (int64, reference layout) and per-depth bucket starts from the counts of
// splitting nodes per depth; the S roots are nodes 0..S-1 (depth 0).
__global__ void levels_kernel(const u32* __restrict__ depth_count, i64 roots_n,
                              i64* __restrict__ levels, u32* __restrict__ bstart) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    levels[0] = 0;
    levels[1] = roots_n;
    i64 lv = roots_n;
    u32 b = 0;
    for (int d = 0; d <= kMaxLevels; ++d) {
        bstart[d] = b;
        const u32 c = depth_count[d];
        b += c;
        if (c) {
            lv += 4 * (i64)c;
            levels[d + 2] = lv;
        }
    }
    bstart[kMaxLevels + 1] = b;
}

// The build without a host read-back (td3_forest_build_nosync): node count, level count
// and level starts from the per-depth counts on the device.  sizes = {node count, levels,
// needed capacity when it overflows (else 0), largest need of any overflow so far, 1 when
// a cell would split below the key's levels (else 0), the same for any call so far}; the
// "so far" entries are kept until the caller clears them.  Every level start past the
// deepest level is the node count, so those levels are empty.  When the nodes do not fit,
// or the key is too short, the forest is the roots alone, each a leaf.
__global__ void levels_nosync_kernel(const u32* __restrict__ depth_count, i64 roots_n,
                                     i64 capacity, const u32* __restrict__ short_key,
                                     i64* __restrict__ levels, u32* __restrict__ bstart,
                                     i64* __restrict__ sizes) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    i64 count = roots_n;
    int n_levels = 1;
    for (int d = 0; d <= kMaxLevels; ++d) {
        const u32 c = depth_count[d];
        if (c) {
            count += 4 * (i64)c;
            n_levels = d + 2;
        }
    }
    const bool too_short = *short_key != 0;
    const bool over = count > capacity || too_short;
    levels[0] = 0;
    levels[1] = roots_n;
    i64 lv = roots_n;
    u32 b = 0;
    for (int d = 0; d <= kMaxLevels; ++d) {
        bstart[d] = b;
        const u32 c = depth_count[d];
        b += c;
        if (!over) lv += 4 * (i64)c;
        levels[d + 2] = lv;
    }
    bstart[kMaxLevels + 1] = b;
    sizes[0] = over ? roots_n : count;
    sizes[1] = over ? 1 : n_levels;
    const i64 need = !too_short && count > capacity ? count : 0;
    sizes[2] = over ? (need ? need : -1) : 0;
    if (need > sizes[3]) sizes[3] = need;
    sizes[4] = too_short ? 1 : 0;
    if (too_short) sizes[5] = 1;
}

__global__ void zero_u32_kernel(u32* __restrict__ p, int n) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) p[i] = 0u;
}

// Entries [s_total, ne) get the depth key 255, which sorts after every real depth.
__global__ void pad_entries_kernel(u64* __restrict__ ekey, u32* __restrict__ eval, i64 ne,
                                   const u32* __restrict__ s_total) {
    const i64 k = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ne || k < (i64)*s_total) return;
    ekey[k] = 255ull;
    eval[k] = 0u;
}

// The root of every system: its particle range [sys_start[s], sys_start[s + 1]) (the whole
// set when sys_start is null) and, when it splits, its children from the rank of its
// depth-0 entry.
__global__ void roots_kernel(i64 roots_n, const i64* __restrict__ sys_start, i64 n, i64 cap,
                             int max_depth, Root rt, const u32* __restrict__ epos,
                             const u32* __restrict__ depth_count, const i64* __restrict__ levels,
                             double* __restrict__ node_center, double* __restrict__ node_half,
                             i64* __restrict__ node_start, i64* __restrict__ node_end,
                             i64* __restrict__ node_child, i64* __restrict__ node_depth,
                             int4* __restrict__ node_info, const i64* __restrict__ sizes) {
    const i64 sys = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (sys >= roots_n) return;
    const i64 start = sys_start ? sys_start[sys] : 0;
    const i64 end = sys_start ? sys_start[sys + 1] : n;
    const bool over = sizes != nullptr && sizes[2] != 0;
    const bool split = end - start > cap && max_depth > 0 && !over;
    i64 child = -1;
    if (split) {
        // depth-0 entries are the first bucket, in particle order
        i64 a = 0, z = depth_count[0];
        while (a < z) {
            const i64 mid = (a + z) >> 1;
            if ((i64)epos[mid] < start)
                a = mid + 1;
            else
                z = mid;
        }
        child = levels[1] + 4 * a;
    }
    double cx, cy, h;
    root_of(rt, (int)sys, &cx, &cy, &h);
    node_center[2 * sys] = cx;
    node_center[2 * sys + 1] = cy;
    node_half[sys] = h;
    node_start[sys] = start;
    node_end[sys] = end;
    node_child[sys] = child;
    node_depth[sys] = 0;
    if (node_info) node_info[sys] = make_int4((int)start, (int)end, (int)child, 0);
}

// One thread per splitting node: writes its four children.
__global__ void nodes_kernel(const u64* __restrict__ khi, const u64* __restrict__ klo, i64 n,
                             i64 cap, int max_depth, int nkey, Keys kl, Root rt,
                             const u64* __restrict__ edepth, const u32* __restrict__ epos, i64 s_total,
                             const u32* __restrict__ bstart, const u32* __restrict__ depth_count,
                             const i64* __restrict__ levels, double* __restrict__ node_center,
                             double* __restrict__ node_half, i64* __restrict__ node_start,
                             i64* __restrict__ node_end, i64* __restrict__ node_child,
                             i64* __restrict__ node_depth, int4* __restrict__ node_info,
                             const u32* __restrict__ s_limit, const i64* __restrict__ sizes) {
    const i64 g = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    // without a read-back: the entry count on the device, bounded by the entry capacity
    // (s_total), and nothing when the nodes overflow
    if (s_limit != nullptr && (sizes[2] != 0 || g >= (i64)*s_limit)) return;
    if (g >= s_total) return;
    const int d = (int)edepth[g];
    const i64 i = (i64)epos[g];
    const i64 r = g - (i64)bstart[d];
    const u64 ihi = khi[i], ilo = klo ? klo[i] : 0ull;
    // Centre and half-width of the depth-d cell of particle i: the exact descent.
    double cx, cy, h;
    root_of(rt, system_of(ihi, kl), &cx, &cy, &h);
    for (int level = 1; level <= d; ++level) {
        const int q = digit_at(ihi, ilo, level, kl);
        h = 0.5 * h;
        cx = (q & 1) ? cx + h : cx - h;
        cy = (q & 2) ? cy - h : cy + h;
    }
    // End of the cell: first j > i whose key leaves the depth-d prefix.
    i64 lo_j = i + 1, hi_j = n;
    while (lo_j < hi_j) {
        const i64 mid = (lo_j + hi_j) >> 1;
        if (lcp_levels(ihi, ilo, khi[mid], klo ? klo[mid] : 0ull, nkey, kl) >= d)
            lo_j = mid + 1;
        else
            hi_j = mid;
    }
    const i64 e = lo_j;
    i64 bound[5];
    bound[0] = i;
    bound[4] = e;
    for (int q = 1; q < 4; ++q) {
        i64 a = bound[q - 1], z = e;
        while (a < z) {
            const i64 mid = (a + z) >> 1;
            if (digit_at(khi[mid], klo ? klo[mid] : 0ull, d + 1, kl) < q)
                a = mid + 1;
            else
                z = mid;
        }
        bound[q] = a;
    }
    const double hc = 0.5 * h;
    const i64 base = levels[d + 1] + 4 * r;
    for (int q = 0; q < 4; ++q) {
        const i64 node = base + q;
        const i64 start = bound[q], end = bound[q + 1];
        i64 child = -1;
        if (end - start > cap && d + 1 < max_depth) {
            // rank of the entry (d + 1, start) among the splitting nodes of depth d + 1
            i64 a = bstart[d + 1], z = a + depth_count[d + 1];
            while (a < z) {
                const i64 mid = (a + z) >> 1;
                if ((i64)epos[mid] < start)
                    a = mid + 1;
                else
                    z = mid;
            }
            child = levels[d + 2] + 4 * (a - (i64)bstart[d + 1]);
        }
        node_center[2 * node] = (q & 1) ? cx + hc : cx - hc;
        node_center[2 * node + 1] = (q & 2) ? cy - hc : cy + hc;
        node_half[node] = hc;
        node_start[node] = start;
        node_end[node] = end;
        node_child[node] = child;
        node_depth[node] = d + 1;
        if (node_info) node_info[node] = make_int4((int)start, (int)end, (int)child, d + 1);
    }
}

// For every leaf, its start position, scattered to the input order of its particles.
__global__ void leaf_key_kernel(const i64* __restrict__ node_start, const i64* __restrict__ node_end,
                                const i64* __restrict__ node_child, i64 count,
                                const u32* __restrict__ sorted_idx, u64* __restrict__ key2,
                                const i64* __restrict__ sizes) {
    const i64 node = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (sizes != nullptr && node >= sizes[0]) return;
    if (node >= count || node_child[node] >= 0) return;
    const i64 s = node_start[node], e = node_end[node];
    for (i64 k = s; k < e; ++k) key2[sorted_idx[k]] = (u64)s;
}

// Input order inside each leaf of at most kRankLeaf particles, without a sort: a leaf's
// particles (input indices sorted_idx[s..e), all different) each go to the position of
// their rank among them.  Writes order[s..e) for every leaf, so all of order[0, n).
constexpr int kRankLeaf = 32;
// With sizes (the build without a read-back): the node count is sizes[0], and nothing is
// done when the nodes overflowed (sizes[2]: then identity_order_kernel writes the order).
__global__ void leaf_rank_kernel(const i64* __restrict__ node_start, const i64* __restrict__ node_end,
                                 const i64* __restrict__ node_child, i64 count,
                                 const u32* __restrict__ sorted_idx, u32* __restrict__ order,
                                 const i64* __restrict__ sizes) {
    const i64 node = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (sizes != nullptr && (sizes[2] != 0 || node >= sizes[0])) return;
    if (node >= count || node_child[node] >= 0) return;
    const i64 s = node_start[node], e = node_end[node];
    u32 v[kRankLeaf];
    // the caller launches this only when every leaf holds at most kRankLeaf particles; the
    // bound keeps every access inside v whatever the input
    const int k = (int)(e - s < kRankLeaf ? e - s : kRankLeaf);
    for (int a = 0; a < k; ++a) v[a] = sorted_idx[s + a];
    for (int a = 0; a < k; ++a) {
        int rank = 0;
        for (int b = 0; b < k; ++b) rank += v[b] < v[a] ? 1 : 0;
        order[s + rank] = v[a];
    }
}

// A forest of the roots alone (the nodes overflowed): each root's leaf holds its system's
// particles, which are contiguous and in input order already, so order[k] = k.
__global__ void identity_order_kernel(i64 n, const i64* __restrict__ sizes, u32* __restrict__ order) {
    const i64 k = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n || sizes[2] == 0) return;
    order[k] = (u32)k;
}

__global__ void gather_kernel(const u32* __restrict__ order, const double* __restrict__ pos, i64 n,
                              i64* __restrict__ perm, double* __restrict__ sorted_pos) {
    const i64 k = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const u32 j = order[k];
    perm[k] = (i64)j;
    if (sorted_pos) {
        sorted_pos[k] = pos[j];
        sorted_pos[n + k] = pos[n + j];
    }
}

struct TreeWs {
    u64 *khi, *khi_alt, *klo, *klo_alt, *ekey, *ekey_alt;
    u32 *idx, *idx_alt, *cnt, *off, *eval, *eval_alt, *scratch, *small;
};

i64 entries_cap(i64 capacity) { return (capacity - 1) / 4 + 1; }

size_t carve(TreeWs* w, void* base, i64 n, i64 capacity) {
    const i64 ne = entries_cap(capacity);
    const i64 big = n > ne ? n : ne;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return base ? (void*)((char*)base + o) : nullptr;
    };
    w->khi = (u64*)take(sizeof(u64) * (size_t)big);
    w->khi_alt = (u64*)take(sizeof(u64) * (size_t)big);
    w->klo = (u64*)take(sizeof(u64) * (size_t)n);
    w->klo_alt = (u64*)take(sizeof(u64) * (size_t)n);
    w->idx = (u32*)take(sizeof(u32) * (size_t)big);
    w->idx_alt = (u32*)take(sizeof(u32) * (size_t)big);
    w->cnt = (u32*)take(sizeof(u32) * (size_t)(n + 1));
    w->off = (u32*)take(sizeof(u32) * (size_t)(n + 1));
    w->ekey = (u64*)take(sizeof(u64) * (size_t)ne);
    w->ekey_alt = (u64*)take(sizeof(u64) * (size_t)ne);
    w->eval = (u32*)take(sizeof(u32) * (size_t)ne);
    w->eval_alt = (u32*)take(sizeof(u32) * (size_t)ne);
    w->scratch = (u32*)take(td3sort::scratch_bytes(big));
    w->small = (u32*)take(sizeof(u32) * 256);
    return off;
}

int bits_for(i64 v) {
    int b = 0;
    while (b < 63 && ((i64)1 << b) <= v) ++b;
    return b;
}

}  // namespace

extern "C" size_t td3_tree_workspace_bytes(int64_t n, int64_t capacity) {
    TreeWs w;
    return carve(&w, nullptr, n < 1 ? 1 : n, capacity < 1 ? 1 : capacity);
}

namespace {

// The build of S >= 1 systems: roots (S, 3) on the device (or the single root), psys the
// system of every input particle (non-decreasing; null for one system), sys_start (S + 1)
// the systems' particle offsets on the device (null for one system).
int build_impl(const double* pos, int64_t n, i64 roots_n, Root rt, const int* psys,
               const i64* sys_start, int64_t leaf_capacity, int64_t max_depth, int64_t capacity,
               double* node_center, double* node_half, int64_t* node_start, int64_t* node_end,
               int64_t* node_child, int64_t* node_depth, int64_t* levels, int64_t* perm,
               double* sorted_pos, void* node_info, void* workspace, size_t workspace_bytes,
               int64_t* out_sizes, int64_t* levels_host, int64_t* sizes_dev,
               int64_t key_levels, hipStream_t s) {
    // sizes_dev set: the build without a host read-back (td3_forest_build_nosync), its keys
    // holding key_levels levels
    const bool nosync = sizes_dev != nullptr;
    TD3_ARG(n > 0 && n < (int64_t)0x7fffffff, "n must be in [1, 2^31)");
    TD3_ARG(leaf_capacity >= 1 && max_depth >= 1, "leaf_capacity and max_depth must be positive");
    TD3_ARG(capacity >= roots_n, "capacity");
    TD3_ARG(nosync || out_sizes != nullptr, "out_sizes");
    TD3_ARG(levels != nullptr, "levels (int64[67])");
    TreeWs w;
    const size_t need = carve(&w, nullptr, n, capacity);
    TD3_ARG(workspace != nullptr && workspace_bytes >= need, "workspace too small");
    carve(&w, workspace, n, capacity);
    Keys kl;
    kl.sbits = roots_n > 1 ? bits_for(roots_n - 1) : 0;
    kl.l1 = (64 - kl.sbits) / 2;
    const int max_levels = kl.l1 + 32 < kMaxLevels ? kl.l1 + 32 : kMaxLevels;
    TD3_ARG(kl.sbits <= 24, "at most 2^24 systems");
    TD3_ARG(!nosync || key_levels >= 1, "key_levels must be positive");
    const int nlev = (int)(max_depth < max_levels ? max_depth : max_levels);
    const int max_split = (int)(max_depth - 1 < max_levels ? max_depth - 1 : max_levels);
    u32* depth_count = w.small;          // [65]
    u32* flag = w.small + 80;            // [1]
    u32* report = w.small + 81;          // [1] the positions' check
    u32* bstart = w.small + 96;          // [66]
    u32* s_total = w.small + 176;        // [1]
    const unsigned nb = td3_blocks(n, 256);

    // 1-2. keys and sort by the first 16 levels; when a cell that deep may still split, by
    // all levels held in hi, then by all of them.  Ties keep input order (the sort is
    // stable), and the topology depends only on prefixes no longer than the deepest split.
    int stages[3], nstage = 0;
    for (int k : {16, kl.l1, nlev}) {
        if (nosync) k = key_levels;
        const int key = k < nlev ? k : nlev;
        if (nstage == 0 || key > stages[nstage - 1]) stages[nstage++] = key;
    }
    int nkey = 0;
    bool deep = false;
    u32 host_small[82];
    for (int stage = 0; stage < nstage; ++stage) {
        nkey = stages[stage];
        deep = nkey > kl.l1;
        if (nosync) {
            // a fill kernel, not a memset: this build may be recorded into a HIP graph
            hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(128), 0, s, w.small, 96);
            TD3_LAUNCHED();
        } else {
            TD3_HIP(hipMemsetAsync(w.small, 0, sizeof(u32) * 96, s));
        }
        if (!deep) {
            TD3_TIMED(s, "tree_keys", hipLaunchKernelGGL(keys_kernel, dim3(nb), dim3(256), 0, s, pos, n, rt, psys, kl, nkey,
                               w.khi, (u64*)nullptr, report));
            TD3_LAUNCHED();
            if (int rc = td3sort::sort_pairs(w.khi, w.idx, nullptr, w.khi_alt, w.idx_alt, nullptr, n,
                                             64 - kl.sbits - 2 * nkey, 64, true, w.scratch, s))
                return rc;
        } else {
            TD3_TIMED(s, "tree_keys", hipLaunchKernelGGL(keys_kernel, dim3(nb), dim3(256), 0, s, pos, n, rt, psys, kl, nkey,
                               w.khi, w.klo, report));
            TD3_LAUNCHED();
            if (int rc = td3sort::sort_pairs(w.klo, w.idx, w.khi, w.klo_alt, w.idx_alt, w.khi_alt, n,
                                             64 - 2 * (nkey - kl.l1), 64, true, w.scratch, s))
                return rc;
            if (int rc = td3sort::sort_pairs(w.khi, w.idx, w.klo, w.khi_alt, w.idx_alt, w.klo_alt, n,
                                             0, 64, false, w.scratch, s))
                return rc;
        }
        // 3. splitting-node starts
        TD3_TIMED(s, "tree_count", hipLaunchKernelGGL(count_kernel, dim3(nb < kCountBlocks ? nb : kCountBlocks), dim3(256), 0, s, w.khi, deep ? w.klo : nullptr, n,
                           (i64)leaf_capacity, max_split, nkey, kl, w.cnt, depth_count, flag));
        TD3_LAUNCHED();
        if (nosync) break;  // a set flag is read on the device (levels_nosync_kernel)
        TD3_HIP(hipMemcpyAsync(host_small, w.small, sizeof(u32) * 82, hipMemcpyDeviceToHost, s));
        TD3_HIP(hipStreamSynchronize(s));
        if (host_small[81] & 1u) {
            td3_set_error("Particle position must be finite");
            return TD3_ERR_ARG;
        }
        if (host_small[81] & 2u) {
            td3_set_error("Particle lies outside tree bounds");
            return TD3_ERR_ARG;
        }
        if (!host_small[80]) break;
        if (stage == nstage - 1) {
            td3_set_error("the tree needs more than %d levels (max_depth %lld)", max_levels,
                          (long long)max_depth);
            return TD3_ERR_ARG;
        }
    }
    if (nosync) {
        // 4'. the same steps at the capacity's sizes, every count read on the device
        const i64 ne = entries_cap(capacity);
        const u64* klo_n = deep ? w.klo : nullptr;
        if (int rc = td3sort::exclusive_scan_u32(w.cnt, w.off, n, s_total, w.scratch, s)) return rc;
        hipLaunchKernelGGL(entries_kernel, dim3(nb), dim3(256), 0, s, w.khi, klo_n, n, nkey, kl,
                           w.cnt, w.off, ne, w.ekey, w.eval);
        TD3_LAUNCHED();
        hipLaunchKernelGGL(pad_entries_kernel, dim3(td3_blocks(ne, 256)), dim3(256), 0, s, w.ekey,
                           w.eval, ne, (const u32*)s_total);
        TD3_LAUNCHED();
        if (int rc = td3sort::sort_pairs(w.ekey, w.eval, nullptr, w.ekey_alt, w.eval_alt, nullptr,
                                         ne, 0, 8, false, w.scratch, s))
            return rc;
        hipLaunchKernelGGL(levels_nosync_kernel, dim3(1), dim3(64), 0, s, depth_count, roots_n,
                           (i64)capacity, (const u32*)flag, levels, bstart, sizes_dev);
        TD3_LAUNCHED();
        hipLaunchKernelGGL(roots_kernel, dim3(td3_blocks(roots_n, 256)), dim3(256), 0, s, roots_n,
                           sys_start, (i64)n, (i64)leaf_capacity, (int)max_depth, rt, w.eval,
                           depth_count, levels, node_center, node_half, node_start, node_end,
                           node_child, node_depth, (int4*)node_info, (const i64*)sizes_dev);
        TD3_LAUNCHED();
        TD3_TIMED(s, "tree_nodes", hipLaunchKernelGGL(nodes_kernel, dim3(td3_blocks(ne, 256)), dim3(256), 0, s, w.khi, klo_n, n,
                           (i64)leaf_capacity, (int)max_depth, nkey, kl, rt, w.ekey, w.eval,
                           ne, bstart, depth_count, levels, node_center, node_half, node_start,
                           node_end, node_child, node_depth, (int4*)node_info,
                           (const u32*)s_total, (const i64*)sizes_dev));
        TD3_LAUNCHED();
        if (leaf_capacity <= kRankLeaf && nkey < max_depth) {
            // no node is deeper than the key's levels, so none is at max_depth and every
            // leaf holds at most leaf_capacity particles (or the forest is the roots alone)
            TD3_TIMED(s, "tree_leaf_keys", hipLaunchKernelGGL(leaf_rank_kernel, dim3(td3_blocks(capacity, 256)), dim3(256), 0, s,
                               node_start, node_end, node_child, (i64)capacity, w.idx, w.idx_alt,
                               (const i64*)sizes_dev));
            TD3_LAUNCHED();
            hipLaunchKernelGGL(identity_order_kernel, dim3(nb), dim3(256), 0, s, (i64)n,
                               (const i64*)sizes_dev, w.idx_alt);
            TD3_LAUNCHED();
        } else {
            TD3_TIMED(s, "tree_leaf_keys", hipLaunchKernelGGL(leaf_key_kernel, dim3(td3_blocks(capacity, 256)), dim3(256), 0, s, node_start,
                               node_end, node_child, (i64)capacity, w.idx, w.khi_alt,
                               (const i64*)sizes_dev));
            TD3_LAUNCHED();
            if (int rc = td3sort::sort_pairs(w.khi_alt, w.idx_alt, nullptr, w.khi, w.idx, nullptr, n,
                                             0, bits_for(n), true, w.scratch, s))
                return rc;
        }
        TD3_TIMED(s, "tree_gather", hipLaunchKernelGGL(gather_kernel, dim3(nb), dim3(256), 0, s, w.idx_alt, pos, n, perm, sorted_pos));
        TD3_LAUNCHED();
        return TD3_OK;
    }
    // node count and level count
    i64 count = roots_n, s_sum = 0;
    int n_levels = 1;
    for (int d = 0; d <= kMaxLevels; ++d) {
        if (host_small[d]) {
            count += 4 * (i64)host_small[d];
            s_sum += host_small[d];
            n_levels = d + 2;
        }
    }
    out_sizes[0] = count;
    out_sizes[1] = n_levels;
    out_sizes[2] = count;
    if (levels_host != nullptr) {
        // the level starts, from the splitting nodes per depth already read back: each
        // has four children one depth down
        levels_host[0] = 0;
        levels_host[1] = roots_n;
        for (int d = 0; d + 2 <= n_levels; ++d)
            levels_host[d + 2] = levels_host[d + 1] + 4 * (i64)host_small[d];
    }
    if (count > capacity) {
        td3_set_error("node capacity %lld < %lld", (long long)capacity, (long long)count);
        return TD3_ERR_CAPACITY;
    }
    const u64* klo = deep ? w.klo : nullptr;
    // 4. entries (depth, particle) in particle order, then bucketed by depth (stable)
    if (s_sum > 0) {
        // offsets into a separate array: entries_kernel needs the counts and the offsets
        if (int rc = td3sort::exclusive_scan_u32(w.cnt, w.off, n, s_total, w.scratch, s)) return rc;
        hipLaunchKernelGGL(entries_kernel, dim3(nb), dim3(256), 0, s, w.khi, klo, n, nkey, kl, w.cnt,
                           w.off, (i64)INT64_MAX, w.ekey, w.eval);
        TD3_LAUNCHED();
        if (int rc = td3sort::sort_pairs(w.ekey, w.eval, nullptr, w.ekey_alt, w.eval_alt, nullptr,
                                         s_sum, 0, 8, false, w.scratch, s))
            return rc;
    }
    hipLaunchKernelGGL(levels_kernel, dim3(1), dim3(64), 0, s, depth_count, roots_n, levels, bstart);
    TD3_LAUNCHED();
    hipLaunchKernelGGL(roots_kernel, dim3(td3_blocks(roots_n, 256)), dim3(256), 0, s, roots_n,
                       sys_start, (i64)n, (i64)leaf_capacity, (int)max_depth, rt, w.eval,
                       depth_count, levels, node_center, node_half, node_start, node_end,
                       node_child, node_depth, (int4*)node_info, (const i64*)nullptr);
    TD3_LAUNCHED();
    if (s_sum > 0) {
        TD3_TIMED(s, "tree_nodes", hipLaunchKernelGGL(nodes_kernel, dim3(td3_blocks(s_sum, 256)), dim3(256), 0, s, w.khi, klo, n,
                           (i64)leaf_capacity, (int)max_depth, nkey, kl, rt, w.ekey, w.eval,
                           s_sum, bstart, depth_count, levels, node_center, node_half, node_start,
                           node_end, node_child, node_depth, (int4*)node_info,
                           (const u32*)nullptr, (const i64*)nullptr));
        TD3_LAUNCHED();
    }
    // 5. input order inside each leaf.  A leaf holds more than leaf_capacity particles only
    // at max_depth: when no node is that deep and leaves are small, each leaf ranks its
    // own; otherwise the input indices are sorted by their leaf's start.
    if (leaf_capacity <= kRankLeaf && n_levels - 1 < max_depth) {
        TD3_TIMED(s, "tree_leaf_keys", hipLaunchKernelGGL(leaf_rank_kernel, dim3(td3_blocks(count, 256)), dim3(256), 0, s,
                           node_start, node_end, node_child, count, w.idx, w.idx_alt,
                           (const i64*)nullptr));
        TD3_LAUNCHED();
    } else {
        TD3_TIMED(s, "tree_leaf_keys", hipLaunchKernelGGL(leaf_key_kernel, dim3(td3_blocks(count, 256)), dim3(256), 0, s, node_start,
                           node_end, node_child, count, w.idx, w.khi_alt, (const i64*)nullptr));
        TD3_LAUNCHED();
        if (int rc = td3sort::sort_pairs(w.khi_alt, w.idx_alt, nullptr, w.khi, w.idx, nullptr, n, 0,
                                         bits_for(n), true, w.scratch, s))
            return rc;
    }
    TD3_TIMED(s, "tree_gather", hipLaunchKernelGGL(gather_kernel, dim3(nb), dim3(256), 0, s, w.idx_alt, pos, n, perm, sorted_pos));
    TD3_LAUNCHED();
    return TD3_OK;
}

}  // namespace

extern "C" int td3_tree_build(const double* pos, int64_t n, double xmid, double ymid, double half,
                              int64_t leaf_capacity, int64_t max_depth, int64_t capacity,
                              double* node_center, double* node_half, int64_t* node_start,
                              int64_t* node_end, int64_t* node_child, int64_t* node_depth,
                              int64_t* levels, int64_t* perm, double* sorted_pos, void* node_info,
                              void* workspace, size_t workspace_bytes, int64_t* out_sizes,
                              int64_t* levels_host, void* stream) {
    Root rt;
    rt.roots = nullptr;
    rt.cx = xmid;
    rt.cy = ymid;
    rt.h = half;
    rt.y_down = 0;
    rt.check = 1;
    return build_impl(pos, n, 1, rt, nullptr, nullptr, leaf_capacity, max_depth, capacity,
                      node_center, node_half, node_start, node_end, node_child, node_depth, levels,
                      perm, sorted_pos, node_info, workspace, workspace_bytes, out_sizes,
                      levels_host, nullptr, 0, (hipStream_t)stream);
}

extern "C" int td3_forest_build(const double* pos, int64_t n, int64_t systems,
                                const int32_t* particle_system, const int64_t* system_start,
                                const double* roots, int64_t leaf_capacity, int64_t max_depth,
                                int64_t capacity, double* node_center, double* node_half,
                                int64_t* node_start, int64_t* node_end, int64_t* node_child,
                                int64_t* node_depth, int64_t* levels, int64_t* perm,
                                double* sorted_pos, void* node_info, void* workspace,
                                size_t workspace_bytes, int64_t* out_sizes,
                                int64_t* levels_host, void* stream) {
    TD3_ARG(systems >= 1 && systems <= (1 << 24), "systems");
    TD3_ARG(particle_system != nullptr && system_start != nullptr && roots != nullptr,
            "particle_system, system_start and roots");
    Root rt;
    rt.roots = roots;
    rt.cx = rt.cy = rt.h = 0.0;
    rt.y_down = 0;
    rt.check = 0;
    return build_impl(pos, n, systems, rt, particle_system, system_start, leaf_capacity, max_depth,
                      capacity, node_center, node_half, node_start, node_end, node_child,
                      node_depth, levels, perm, sorted_pos, node_info, workspace, workspace_bytes,
                      out_sizes, levels_host, nullptr, 0, (hipStream_t)stream);
}

extern "C" int td3_forest_build_nosync(const double* pos, int64_t n, int64_t systems,
                                       const int32_t* particle_system, const int64_t* system_start,
                                       const double* roots, int64_t leaf_capacity,
                                       int64_t max_depth, int64_t capacity,
                                       int64_t key_levels, double* node_center,
                                       double* node_half, int64_t* node_start, int64_t* node_end,
                                       int64_t* node_child, int64_t* node_depth, int64_t* levels,
                                       int64_t* perm, double* sorted_pos, void* node_info,
                                       void* workspace, size_t workspace_bytes, int64_t* sizes,
                                       void* stream) {
    TD3_ARG(systems >= 1 && systems <= (1 << 24), "systems");
    TD3_ARG(particle_system != nullptr && system_start != nullptr && roots != nullptr,
            "particle_system, system_start and roots");
    TD3_ARG(sizes != nullptr && levels != nullptr && node_info != nullptr, "sizes, levels, node_info");
    Root rt;
    rt.roots = roots;
    rt.cx = rt.cy = rt.h = 0.0;
    rt.y_down = 0;
    rt.check = 0;
    return build_impl(pos, n, systems, rt, particle_system, system_start, leaf_capacity, max_depth,
                      capacity, node_center, node_half, node_start, node_end, node_child,
                      node_depth, levels, perm, sorted_pos, node_info, workspace, workspace_bytes,
                      nullptr, nullptr, sizes, key_levels, (hipStream_t)stream);
}

extern "C" int td3_box_tree_build(const double* pos, int64_t n, double xmid, double ymid,
                                  double half, int64_t leaf_capacity, int64_t max_depth,
                                  int64_t capacity, double* node_center, double* node_half,
                                  int64_t* node_start, int64_t* node_end, int64_t* node_child,
                                  int64_t* node_depth, int64_t* levels, int64_t* perm,
                                  double* sorted_pos, void* node_info, void* workspace,
                                  size_t workspace_bytes, int64_t* out_sizes,
                                  int64_t* levels_host, void* stream) {
    Root rt;
    rt.roots = nullptr;
    rt.cx = xmid;
    rt.cy = ymid;
    rt.h = half;
    rt.y_down = 1;
    rt.check = 0;
    return build_impl(pos, n, 1, rt, nullptr, nullptr, leaf_capacity, max_depth, capacity,
                      node_center, node_half, node_start, node_end, node_child, node_depth, levels,
                      perm, sorted_pos, node_info, workspace, workspace_bytes, out_sizes,
                      levels_host, nullptr, 0, (hipStream_t)stream);
}
