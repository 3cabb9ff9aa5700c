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
// Per level, one thread per inherited pair (the pairs of all the level's nodes, in node
// order): count (each walk counted per column: M2L, P2P, hand-downs to each child -- the
// same for all four -- with its first few outputs cached), an exclusive scan of the counts
// (which gives every write position, the rows of a node being its pairs' outputs in
// order), write (the cached outputs replayed; only longer walks walk again).  A walk needs
// no stack: it moves through the other cell's subtree by child and parent links, so any
// depth works.

#include "td3_common.h"
#include "td3_sort.h"

namespace {

constexpr u32 kSecond = 0x80000000u;

struct RowsTree {
    const double4* geo;  // {centre x, centre y, half, 0}: one load per visited cell
    const int4* info;    // {start, end, child, depth}
    const int* parent;
    int sib0;            // the first non-root node: siblings come in fours from here
};

// The decision for the pair {a, x} (a is `a_second` ? second : first; ga, fa its geometry
// and info).  Returns 0 drop, 1 M2L, 2 P2P, 3 split a, 4 split x (x's first child in
// *child).
__device__ __forceinline__ int decide(const RowsTree& t, double4 ga, int4 fa, int x,
                                      bool a_second, double theta, int* child) {
    // both loads issued together (the geometry is not needed for an empty cell, but
    // waiting for the count first would serialise two memory latencies per visit)
    const int4 fx = t.info[x];
    const double4 gx = t.geo[x];
    if (fx.y == fx.x || fa.y == fa.x) return 0;
    const bool leaf_a = fa.z < 0, leaf_x = fx.z < 0;
    const double4 gf = a_second ? gx : ga, gs = a_second ? ga : gx;
    const double dx = __dadd_rn(gf.x, -gs.x);
    const double dy = __dadd_rn(gf.y, -gs.y);
    const double distance = __dsqrt_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)));
    const double radii = __dmul_rn(1.4142135623730951, __dadd_rn(gf.z, gs.z));
    if (radii < __dmul_rn(theta, distance)) return 1;
    if (leaf_a && leaf_x) return 2;
    const bool leaf_f = a_second ? leaf_x : leaf_a, leaf_s = a_second ? leaf_a : leaf_x;
    const bool split_first = leaf_s || (!leaf_f && gf.z >= gs.z);
    const bool split_a = split_first != a_second;
    *child = fx.z;
    return split_a ? 3 : 4;
}

// Per entry: M2L entries, P2P entries, and the entries handed to EACH child (a walk hands
// every pair it splits on a's side to all four children, a self pair four to each).
struct Counts {
    u32 m2l, p2p, inh;
};

__device__ __forceinline__ u32 entry_with(int x, bool second) {
    return (u32)x | (second ? kSecond : 0u);
}

// A walk's outputs, in order, are events: M2L (x), P2P (x), or a hand-down of (x, side) to
// each of a's four children.  The counting pass keeps the first kCached events of every
// entry (code: x in bits 0-28, the kind in 29-30, the side in 31), so that the writing
// pass replays them instead of walking again.  (32 covers nearly every walk at N = 1M:
// a leaf beside a dense cluster has a P2P entry per cluster leaf nearby; with 8 the
// writing pass took 0.79 ms there, with 32 0.51 ms.  Slots never written cost nothing.)
constexpr int kCached = 32;
constexpr u32 kEvM2L = 1u << 29, kEvP2P = 2u << 29, kEvSplit = 3u << 29, kEvKind = 3u << 29;

// Counts per column and caches the events (cache[i * stride + e], i < kCached).
struct CountSink {
    Counts c;
    u32 events;
    u32* cache;
    u32 stride;
    __device__ void event(u32 code) {
        if (events < (u32)kCached) cache[(size_t)events * stride] = code;
        ++events;
    }
    __device__ void m2l(int x) {
        ++c.m2l;
        event((u32)x | kEvM2L);
    }
    __device__ void p2p(int x) {
        ++c.p2p;
        event((u32)x | kEvP2P);
    }
    __device__ void split(int x, bool second) {
        ++c.inh;
        event((u32)x | kEvSplit | (second ? kSecond : 0u));
    }
    __device__ void self_leaf(int a) { ++c.p2p; }
    __device__ void self_children(int child0) { c.inh += 4; }
};

// Writes at the offsets it holds (advancing them): the rows (and the M2L entries' row
// node), the next level's pairs and their nodes.
struct WriteSink {
    Counts c;
    int a, first_child;
    int* m2l_col;
    int* m2l_dst;
    int* p2p_col;
    u32* inh_next;
    int* enode_next;
    u32 child_base, child_stride;  // child k's list of a starts at child_base + k child_stride
    __device__ void m2l(int x) {
        m2l_col[c.m2l] = x;
        if (m2l_dst != nullptr) m2l_dst[c.m2l] = a;
        ++c.m2l;
    }
    __device__ void p2p(int x) { p2p_col[c.p2p++] = x; }
    __device__ void split(int x, bool second) {
        for (int k = 0; k < 4; ++k) {
            const u32 at = child_base + k * child_stride + c.inh;
            inh_next[at] = entry_with(x, second);
            enode_next[at] = first_child + k;
        }
        ++c.inh;
    }
    __device__ void self_leaf(int x) { p2p(x); }
    // the ten child pairs (c_i, c_j), i <= j: child k gets every sibling, itself as a self
    // pair, c_j (j > k) with k first, c_i (i < k) with k second
    __device__ void self_children(int child0) {
        for (int k = 0; k < 4; ++k) {
            const u32 at = child_base + k * child_stride + c.inh;
            for (int j = 0; j < 4; ++j) {
                inh_next[at + j] = (u32)(child0 + j) | (j < k ? kSecond : 0u);
                enode_next[at + j] = child0 + k;
            }
        }
        c.inh += 4;
    }
};

// The self pair {a, a}: its leaf is one P2P entry, else the sibling pairs go down.
template <class Sink>
__device__ __forceinline__ void self_entry(int a, int4 fa, Sink& sink) {
    if (fa.y == fa.x) return;
    if (fa.z < 0)
        sink.self_leaf(a);
    else
        sink.self_children(fa.z);
}

// Walks one inherited entry (not a self pair) of node a: depth first over the subtree of
// the other cell x0, without a stack (children in order, back up through the parents once
// a fourth child is done).
template <class Sink>
__device__ __forceinline__ void walk_entry(const RowsTree& t, int a, int4 fa, u32 entry, double theta,
                           Sink& sink) {
    const int x0 = (int)(entry & ~kSecond);
    const bool second = (entry & kSecond) != 0;
    const double4 ga = t.geo[a];
    int x = x0;
    while (true) {
        int child = -1;
        const int kind = decide(t, ga, fa, x, second, theta, &child);
        if (kind == 4) {
            x = child;
            continue;
        }
        if (kind == 1)
            sink.m2l(x);
        else if (kind == 2)
            sink.p2p(x);
        else if (kind == 3)
            sink.split(x, second);
        while (x != x0 && ((x - t.sib0) & 3) == 3) x = t.parent[x];
        if (x == x0) break;
        ++x;
    }
}

// Replays cached events into a writing sink.
__device__ __forceinline__ void replay(const u32* cache, u32 stride, u32 events, WriteSink& sink) {
    for (u32 i = 0; i < events; ++i) {
        const u32 code = cache[(size_t)i * stride];
        const int x = (int)(code & ~(kEvKind | kSecond));
        const u32 kind = code & kEvKind;
        if (kind == kEvM2L)
            sink.m2l(x);
        else if (kind == kEvP2P)
            sink.p2p(x);
        else
            sink.split(x, (code & kSecond) != 0);
    }
}

__device__ __forceinline__ u32 wave_sum_u32(u32 v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// ------------------------------------------------------------------------------ per level
// Entries of a level: inh[e] (the pair), enode[e] (its node), in node order; first[i] the
// first entry of the level's i-th node (first[nodes] = the entry count); the count sits on
// the device (lv[0]).  Every node has at least one entry (its self pair), so its first
// entry's thread can write the node's row starts and its children's list starts.
constexpr int kCols = 3;  // M2L, P2P, hand-downs to each child
constexpr unsigned kEntryBlocks = 2048;
constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr u32 kScanTile = kScanThreads * kScanItems;
// The counting pass sums tiles of kCountTile entries (one per thread: walks are long and
// uneven, so small blocks retire early) and adds each sum into the super-tile of
// kSuperTiles tiles (integer sums: the same whatever the order); a scan tile's carry is the
// super-tiles before it plus at most kSuperTiles - 1 tiles.
constexpr u32 kCountTile = 256;
constexpr u32 kSuperTiles = 16;
constexpr u32 kCountPerScan = kScanTile / kCountTile;

__device__ __forceinline__ u32 level_entries(const u32* lv, u32 cap) {
    const u32 n = lv[0];
    return n < cap ? n : cap;
}

// Count step: every entry's walk, counted per column, in tiles of kCountTile entries.
template <int WAVES>
__device__ __forceinline__ void block_sum3(u32 v[kCols], u32 (*s_sum)[kCols]) {
    for (int c = 0; c < kCols; ++c) v[c] = wave_sum_u32(v[c]);
    if (td3_lane() == 0)
        for (int c = 0; c < kCols; ++c) s_sum[threadIdx.x / 64][c] = v[c];
    __syncthreads();
    for (int c = 0; c < kCols; ++c) {
        u32 t = 0;
        for (int w = 0; w < WAVES; ++w) t += s_sum[w][c];
        v[c] = t;
    }
}

__global__ __launch_bounds__(kCountTile) void entry_count_kernel(
    RowsTree t, const u32* __restrict__ inh, const int* __restrict__ enode,
    const u32* __restrict__ lv, u32 cap, double theta, u32* __restrict__ cnt,
    u32* __restrict__ cache, u32* __restrict__ part, u32* __restrict__ super,
    const u32* __restrict__ flag) {
    if (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    const u32 n = level_entries(lv, cap);
    __shared__ u32 s_sum[kCountTile / 64][kCols];
    for (u32 tile = blockIdx.x; tile * kCountTile < n; tile += gridDim.x) {
        u32 sum[kCols] = {0, 0, 0};
        const u32 e = tile * kCountTile + threadIdx.x;
        if (e < n) {
            const int a = enode[e];
            const int4 fa = t.info[a];
            const u32 entry = inh[e];
            CountSink sink = {{0, 0, 0}, 0u, cache + e, cap};
            if ((int)(entry & ~kSecond) == a)
                self_entry(a, fa, sink);
            else
                walk_entry(t, a, fa, entry, theta, sink);
            u32* o = cnt + (size_t)kCols * e;
            o[0] = sink.c.m2l;
            o[1] = sink.c.p2p;
            o[2] = sink.c.inh;
            sum[0] = sink.c.m2l;
            sum[1] = sink.c.p2p;
            sum[2] = sink.c.inh;
        }
        block_sum3<kCountTile / 64>(sum, s_sum);
        if (threadIdx.x < kCols) {
            part[(size_t)kCols * tile + threadIdx.x] = sum[threadIdx.x];
            atomicAdd(super + (size_t)kCols * (tile / kSuperTiles) + threadIdx.x, sum[threadIdx.x]);
        }
        __syncthreads();  // s_sum is reused by the next tile
    }
}

// Exclusive scan of the columns inside each tile, carried by the sums of the tiles before
// it (each block adds them up itself); block 0 also writes the columns' totals.
__global__ __launch_bounds__(kScanThreads) void col_scan_kernel(u32* __restrict__ cnt,
                                                                const u32* __restrict__ lv, u32 cap,
                                                                const u32* __restrict__ part,
                                                                const u32* __restrict__ super,
                                                                u32* __restrict__ totals,
                                                                const u32* __restrict__ flag) {
    if (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    const u32 n = level_entries(lv, cap);
    __shared__ u32 s_sum[kScanThreads / 64][kCols];
    __shared__ u32 s_wave[kScanThreads / 64][kCols];
    const int lane = td3_lane(), wave = threadIdx.x / 64;
    const u32 tiles = (n + kScanTile - 1) / kScanTile;
    for (u32 tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    const u32 base = tile * kScanTile;
    u32 carry[kCols] = {0, 0, 0};
    if (tile == 0) {
        // the totals: every super-tile
        const u32 supers = (tiles * kCountPerScan + kSuperTiles - 1) / kSuperTiles;
        for (u32 i = threadIdx.x; i < supers; i += kScanThreads)
            for (int c = 0; c < kCols; ++c) carry[c] += super[(size_t)kCols * i + c];
    } else {
        // the super-tiles, then the count tiles, before this tile's first count tile
        const u32 first = tile * kCountPerScan, whole = first / kSuperTiles;
        for (u32 i = threadIdx.x; i < whole; i += kScanThreads)
            for (int c = 0; c < kCols; ++c) carry[c] += super[(size_t)kCols * i + c];
        for (u32 i = whole * kSuperTiles + threadIdx.x; i < first; i += kScanThreads)
            for (int c = 0; c < kCols; ++c) carry[c] += part[(size_t)kCols * i + c];
    }
    block_sum3<kScanThreads / 64>(carry, s_sum);
    if (tile == 0) {
        if (threadIdx.x < kCols) totals[threadIdx.x] = carry[threadIdx.x];
        for (int c = 0; c < kCols; ++c) carry[c] = 0;
    }
    for (int r = 0; r < kScanItems; ++r) {
        const u32 e = base + r * kScanThreads + threadIdx.x;
        u32 v[kCols], incl[kCols];
        for (int c = 0; c < kCols; ++c) {
            v[c] = e < n ? cnt[(size_t)kCols * e + c] : 0u;
            incl[c] = v[c];
        }
#pragma unroll
        for (int off = 1; off < 64; off <<= 1)
            for (int c = 0; c < kCols; ++c) {
                const u32 y = __shfl_up(incl[c], off, 64);
                if (lane >= off) incl[c] += y;
            }
        if (lane == 63)
            for (int c = 0; c < kCols; ++c) s_wave[wave][c] = incl[c];
        __syncthreads();
        for (int c = 0; c < kCols; ++c) {
            u32 before = 0, all = 0;
            for (int w = 0; w < kScanThreads / 64; ++w) {
                const u32 y = s_wave[w][c];
                before += w < wave ? y : 0u;
                all += y;
            }
            if (e < n) cnt[(size_t)kCols * e + c] = carry[c] + before + incl[c] - v[c];
            carry[c] += all;
        }
        __syncthreads();
    }
    }
}

// Write step: every entry's walk again, writing at its offsets; the first entry of a node
// also writes the node's row starts and its children's list starts, entry 0 the next
// level's entry count and row bases.
__global__ __launch_bounds__(256) void entry_write_kernel(
    RowsTree t, const u32* __restrict__ inh, const int* __restrict__ enode,
    const u32* __restrict__ first, i64 lo, const u32* __restrict__ lv, u32* __restrict__ lv_next,
    u32 cap, double theta, const u32* __restrict__ scan, const u32* __restrict__ totals,
    u32* __restrict__ base, int level, u32* __restrict__ m2l_row, int* __restrict__ m2l_col,
    int* __restrict__ m2l_dst, u32* __restrict__ p2p_row, int* __restrict__ p2p_col,
    u32* __restrict__ inh_next, int* __restrict__ enode_next, u32* __restrict__ first_next,
    i64 next_lo, i64 next_nodes, u32 cap_m2l, u32 cap_p2p, const u32* __restrict__ cache,
    u32* __restrict__ flag) {
    if (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    const u32 n = level_entries(lv, cap);
    const u32 bm = base[2 * level], bp = base[2 * level + 1];
    const u32 inh_total = 4 * totals[2];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        base[2 * level + 2] = bm + totals[0];
        base[2 * level + 3] = bp + totals[1];
        lv_next[0] = inh_total;
        if (next_nodes > 0) first_next[next_nodes] = inh_total;
    }
    for (u32 e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
        const int a = enode[e];
        const int4 fa = t.info[a];
        const u32 f = first[a - lo], f1 = first[a - lo + 1];
        const u32* se = scan + (size_t)kCols * e;
        const u32* sf = scan + (size_t)kCols * f;
        // children's lists: a hands h entries to each child, child k's list after the
        // earlier children's
        u32 child_base = 0, h = 0;
        if (fa.z >= 0) {
            h = (f1 < n ? scan[(size_t)kCols * f1 + 2] : totals[2]) - sf[2];
            child_base = 4 * sf[2];
            if (e == f)
                for (int k = 0; k < 4; ++k) first_next[fa.z + k - next_lo] = child_base + k * h;
        }
        if (e == f) {
            m2l_row[a] = bm + sf[0];
            p2p_row[a] = bp + sf[1];
        }
        const Counts at = {bm + se[0], bp + se[1], se[2] - sf[2]};
        // the outputs of this entry end where the next entry's begin
        const u32* nx = e + 1 < n ? scan + (size_t)kCols * (e + 1) : totals;
        bool fits = bm + nx[0] <= cap_m2l && bp + nx[1] <= cap_p2p;
        if (fa.z >= 0) fits = fits && child_base + 3 * h + (nx[2] - sf[2]) <= cap;
        if (!fits) {
            atomicOr(flag, 2u);
            continue;
        }
        WriteSink sink = {at, a, fa.z, m2l_col, m2l_dst, p2p_col, inh_next, enode_next,
                          child_base, h};
        const u32 entry = inh[e];
        if ((int)(entry & ~kSecond) == a) {
            self_entry(a, fa, sink);
        } else {
            // a walk's events: one per M2L, P2P or hand-down (each hand-down is one entry
            // in every child's column)
            const u32 events = (nx[0] - se[0]) + (nx[1] - se[1]) + (nx[2] - se[2]);
            if (events <= (u32)kCached)
                replay(cache + e, cap, events, sink);
            else
                walk_entry(t, a, fa, entry, theta, sink);
        }
    }
}

// Parent links, and the count of non-empty leaves (each holds one self pair in P2P).
__global__ void rows_parent_kernel(const int4* __restrict__ info, i64 count, int* __restrict__ parent,
                                   u32* __restrict__ leaves) {
    const i64 node = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= count) return;
    if (node == 0) parent[0] = -1;
    const int4 f = info[node];
    if (f.z >= 0)
        for (int k = 0; k < 4; ++k) parent[f.z + k] = (int)node;
    else if (f.y > f.x)
        atomicAdd(leaves, 1u);
}

__global__ void rows_geo_kernel(const double* __restrict__ center, const double* __restrict__ half,
                                i64 count, double4* __restrict__ geo) {
    const i64 node = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (node < count) geo[node] = make_double4(center[2 * node], center[2 * node + 1], half[node], 0.0);
}

// Every level's row bases start at zero: after an overflow the writes stop, and the
// counts read back are then those of the levels written so far (a lower bound).
__global__ void rows_root_kernel(u32* inh, int* enode, u32* first, u32* lv, u32* base,
                                 int levels) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        inh[0] = 0u;  // the pair (root, root)
        enode[0] = 0;
        first[0] = 0u;
        first[1] = 1u;
        lv[0] = 1u;
        for (int k = 0; k < 2 * levels + 2; ++k) base[k] = 0u;
    }
}

__global__ void rows_empty_kernel(const u32* __restrict__ flag, i64 count,
                                  u32* __restrict__ m2l_row, u32* __restrict__ p2p_row) {
    if (!(__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 2u)) return;
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > count) return;
    m2l_row[i] = 0u;
    p2p_row[i] = 0u;
}

// {M2L entries, P2P entries, non-empty leaves, overflow}
__global__ void rows_counts_kernel(const u32* base, int levels, const u32* flag, int64_t* out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        out[0] = base[2 * levels];
        out[1] = base[2 * levels + 1];
        out[2] = flag[2];
        out[3] = (flag[0] & 2u) ? 1 : 0;
    }
}

__global__ void rows_end_kernel(const u32* base, int levels, i64 count, u32* m2l_row,
                                u32* p2p_row) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        m2l_row[count] = base[2 * levels];
        p2p_row[count] = base[2 * levels + 1];
    }
}

struct RowsWs {
    u32 *inh_a, *inh_b, *first_a, *first_b, *cnt, *cache, *part, *super, *totals, *lv, *base,
        *flag;
    i64 super_stride;  // u32 per level
    int *enode_a, *enode_b, *parent;
    double4* geo;
};

size_t rows_carve(RowsWs* w, void* p, i64 count, i64 width, i64 cap_inh, int levels) {
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return p ? (void*)((char*)p + o) : nullptr;
    };
    const i64 tiles = (cap_inh + kScanTile - 1) / kScanTile;
    w->inh_a = (u32*)take(sizeof(u32) * (size_t)cap_inh);
    w->inh_b = (u32*)take(sizeof(u32) * (size_t)cap_inh);
    w->enode_a = (int*)take(sizeof(int) * (size_t)cap_inh);
    w->enode_b = (int*)take(sizeof(int) * (size_t)cap_inh);
    w->first_a = (u32*)take(sizeof(u32) * (size_t)(width + 1));
    w->first_b = (u32*)take(sizeof(u32) * (size_t)(width + 1));
    w->cnt = (u32*)take(sizeof(u32) * kCols * (size_t)cap_inh);
    w->cache = (u32*)take(sizeof(u32) * kCached * (size_t)cap_inh);
    const i64 count_tiles = (cap_inh + kCountTile - 1) / kCountTile;
    w->part = (u32*)take(sizeof(u32) * kCols * (size_t)count_tiles);
    w->super_stride = (count_tiles + kSuperTiles - 1) / kSuperTiles * kCols;
    w->super = (u32*)take(sizeof(u32) * (size_t)w->super_stride * (size_t)levels);
    w->totals = (u32*)take(sizeof(u32) * kCols);
    w->lv = (u32*)take(sizeof(u32) * 2);
    w->base = (u32*)take(sizeof(u32) * (size_t)(2 * levels + 4));
    w->flag = (u32*)take(sizeof(u32) * 4);
    w->parent = (int*)take(sizeof(int) * (size_t)count);
    w->geo = (double4*)take(sizeof(double4) * (size_t)count);
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
                        int64_t* out_counts, int64_t* device_counts, void* stream) {
    TD3_ARG(levels_host != nullptr && n_levels >= 1 && n_levels <= 64, "levels");
    const i64 count = levels_host[n_levels];
    TD3_ARG(count >= 1 && count < (1 << 29), "node count");  // event codes hold 29 bits
    TD3_ARG(cap_m2l >= 1 && cap_m2l < 0xffffffffll && cap_p2p >= 1 && cap_p2p < 0xffffffffll &&
                cap_inherit >= 16 && cap_inherit < 0xffffffffll, "capacities");
    TD3_ARG(out_counts != nullptr || device_counts != nullptr, "out_counts");
    TD3_ARG(workspace_bytes >= td3_rows_workspace_bytes(levels_host, n_levels, cap_inherit),
            "workspace too small");
    hipStream_t s = (hipStream_t)stream;
    RowsWs w;
    const i64 width = widest_level(levels_host, (int)n_levels);
    rows_carve(&w, workspace, count, width, cap_inherit, (int)n_levels);
    RowsTree t;
    t.geo = w.geo;
    t.info = (const int4*)node_info;
    t.parent = w.parent;
    t.sib0 = (int)(levels_host[1] - levels_host[0]);
    hipLaunchKernelGGL(rows_geo_kernel, dim3(td3_blocks(count, 256)), dim3(256), 0, s, node_center,
                       node_half, count, w.geo);
    TD3_LAUNCHED();
    TD3_HIP(hipMemsetAsync(w.flag, 0, sizeof(u32) * 4, s));
    TD3_HIP(hipMemsetAsync(w.super, 0, sizeof(u32) * (size_t)w.super_stride * n_levels, s));
    hipLaunchKernelGGL(rows_parent_kernel, dim3(td3_blocks(count, 256)), dim3(256), 0, s,
                       (const int4*)node_info, count, w.parent, w.flag + 2);
    TD3_LAUNCHED();
    hipLaunchKernelGGL(rows_root_kernel, dim3(1), dim3(1), 0, s, w.inh_a, w.enode_a, w.first_a,
                       w.lv, w.base, (int)n_levels);
    TD3_LAUNCHED();
    u32 *inh = w.inh_a, *inh_next = w.inh_b, *first = w.first_a, *first_next = w.first_b;
    int *enode = w.enode_a, *enode_next = w.enode_b;
    u32 *lv = w.lv, *lv_next = w.lv + 1;
    const u32 cap = (u32)cap_inherit;
    const unsigned tiles = (unsigned)((cap_inherit + kScanTile - 1) / kScanTile);
    // blocks stride over the tiles: enough to fill the chip, few to launch when a level
    // is small (the entry counts stay on the device)
    const unsigned grid_tiles = tiles < kEntryBlocks ? tiles : kEntryBlocks;
    for (int d = 0; d < n_levels; ++d) {
        const i64 lo = levels_host[d], hi = levels_host[d + 1];
        const i64 next_lo = hi, next_hi = d + 1 < n_levels ? levels_host[d + 2] : hi;
        const i64 next = next_hi - next_lo;
        u32* super = w.super + (size_t)w.super_stride * d;
        TD3_TIMED(s, "rows_count", hipLaunchKernelGGL(entry_count_kernel, dim3(kEntryBlocks), dim3(kCountTile), 0, s, t, inh, enode, lv,
                           cap, theta, w.cnt, w.cache, w.part, super, w.flag));
        TD3_LAUNCHED();
        TD3_TIMED(s, "rows_scan", hipLaunchKernelGGL(col_scan_kernel, dim3(grid_tiles), dim3(kScanThreads), 0, s, w.cnt, lv, cap, w.part,
                           super, w.totals, w.flag));
        TD3_LAUNCHED();
        TD3_TIMED(s, "rows_write", hipLaunchKernelGGL(entry_write_kernel, dim3(kEntryBlocks), dim3(256), 0, s, t, inh, enode, first,
                           lo, lv, lv_next, cap, theta, w.cnt, w.totals, w.base, d, m2l_row,
                           (int*)m2l_col, (int*)m2l_dst, p2p_row, (int*)p2p_col, inh_next,
                           enode_next, first_next, next_lo, next, (u32)cap_m2l, (u32)cap_p2p,
                           w.cache, w.flag));
        TD3_LAUNCHED();
        u32* tmp = inh;
        inh = inh_next;
        inh_next = tmp;
        tmp = first;
        first = first_next;
        first_next = tmp;
        int* ti = enode;
        enode = enode_next;
        enode_next = ti;
        tmp = lv;
        lv = lv_next;
        lv_next = tmp;
        (void)hi;
    }
    hipLaunchKernelGGL(rows_end_kernel, dim3(1), dim3(1), 0, s, w.base, (int)n_levels, count,
                       m2l_row, p2p_row);
    TD3_LAUNCHED();
    if (device_counts != nullptr) {
        // no read-back: the counts go to the device, and an overflow leaves every row
        // empty (safe to evaluate; the caller checks the counts and builds again)
        hipLaunchKernelGGL(rows_empty_kernel, dim3(td3_blocks(count + 1, 256)), dim3(256), 0, s,
                           w.flag, count, m2l_row, p2p_row);
        TD3_LAUNCHED();
        hipLaunchKernelGGL(rows_counts_kernel, dim3(1), dim3(1), 0, s, w.base, (int)n_levels,
                           w.flag, device_counts);
        TD3_LAUNCHED();
        return TD3_OK;
    }
    u32 host[5];
    TD3_HIP(hipMemcpyAsync(host, w.flag, sizeof(u32) * 3, hipMemcpyDeviceToHost, s));
    TD3_HIP(hipMemcpyAsync(host + 3, w.base + 2 * n_levels, sizeof(u32) * 2, hipMemcpyDeviceToHost, s));
    TD3_HIP(hipStreamSynchronize(s));
    out_counts[0] = host[3];
    out_counts[1] = host[4];
    out_counts[2] = host[2];
    out_counts[3] = host[0] & 2u ? 1 : 0;
    if (host[0] & 2u) {
        td3_set_error("interaction rows overflowed (m2l %u, p2p %u)", host[3], host[4]);
        return TD3_ERR_CAPACITY;
    }
    return TD3_OK;
}
