// Interaction rows by node, straight from the traversal rules.
//
// The reference's traversal (flat_tree.py:1318-1358, kernels/fmm_traverse/
// fmm_traverse_pyccel.py:15-106) walks unordered cell pairs from (root, root) and emits
// M2L and P2P pairs; the rows by node come after (kernels/pair_csr/pair_csr_pyccel.py).
// Here every node a gathers its own rows: the pairs {a, x} the traversal visits are, for a
// fixed a, those inherited from a's parent (pairs the parent split on its own side, and
// the sibling pairs of the parent's self pair) closed under splitting the other side.
// Each inherited entry is walked depth first over x's subtree:
//   empty cell -> dropped; self pair: leaf -> P2P, else its children get their sibling
//   pairs; MAC (sqrt(2)(h_a + h_b) < theta d) -> M2L; two leaves -> P2P; else the rule
//   splits the first cell if the second is a leaf or (the first is not a leaf and
//   h_first >= h_second), else the second: splitting a hands (x, side) down to a's four
//   children, splitting x continues the walk with x's children.
// The side (a first or second) follows each pair from its creation, as in the reference,
// because the tie h_first == h_second splits the first cell.  The pair sets are those of
// td3_traverse (both rows of a mutual pair come from the same decision, evaluated with the
// same FP64 operations).
//
// Layout: one wavefront per node, level by level.  Every node owns fixed-width slots
// (`width` entries each) for its M2L row, its P2P row and its hand-downs (stored once: its
// four children read them there).  A node's entries are its parent's hand-downs, then the
// four sibling pairs of the parent's self pair.  The lanes walk 64 entries at a time once,
// keeping their first eight events in registers, take offsets from a wave prefix sum and
// write them (a lane with more events walks again).  Nothing is scanned per level: after
// the last level one kernel sums the node counts per tile of 256 nodes and one places
// every row (the tiles before it, a block scan inside) and packs the slots into them.
// (Per-node atomic adds into the tile sums, from the level kernels, measured 1.8 ms against
// 1.0 ms for the levels at N = 1M: same-address atomics serialise.)  Rows are in a fixed
// order (entries in order, each walked depth first), so every sum over them is the same on
// every run.  A walk needs no stack: it moves through the other cell's subtree by child and
// parent links, so any depth works.

#include "td3_common.h"

namespace {

constexpr u32 kSecond = 0x80000000u;
constexpr int kWaves = 4;       // nodes (wavefronts) per block of the level kernel
constexpr int kBuffered = 8;    // events per lane kept in registers during the walk
constexpr int kTileNodes = 256; // nodes per tile of the row placement
constexpr int kRing = 512;      // live queued cells per wavefront of the breadth-first kernel
constexpr u32 kEvM2L = 1u << 29, kEvP2P = 2u << 29, kEvHand = 3u << 29, kEvKind = 3u << 29;

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
    const double d2 = __dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy));
    const double radii = __dmul_rn(1.4142135623730951, __dadd_rn(gf.z, gs.z));
    // The reference's test is radii < fl(theta fl(sqrt(d2))); its right side lies within a
    // relative 3e-16 of theta sqrt(d2).  Squares decide it without the square root unless
    // radii^2 and theta^2 d2 agree to 1e-12 (their own roundings are a few 1e-16); only
    // then is the reference's expression evaluated.
    const double lhs = radii * radii, rhs = theta * theta * d2;
    bool accept;
    if (lhs < rhs * (1.0 - 1e-12))
        accept = true;
    else if (lhs > rhs * (1.0 + 1e-12))
        accept = false;
    else
        accept = radii < __dmul_rn(theta, __dsqrt_rn(d2));
    if (accept) return 1;
    if (leaf_a && leaf_x) return 2;
    const bool leaf_f = a_second ? leaf_x : leaf_a, leaf_s = a_second ? leaf_a : leaf_x;
    const bool split_first = leaf_s || (!leaf_f && gf.z >= gs.z);
    const bool split_a = split_first != a_second;
    *child = fx.z;
    return split_a ? 3 : 4;
}

__device__ __forceinline__ u32 entry_with(int x, bool second) {
    return (u32)x | (second ? kSecond : 0u);
}

// Slots and counts of every node (counts uncapped: a count above the width reports the
// width the node needs).
struct Slots {
    u32* m2l;    // node * width
    u32* p2p;
    u32* inh;    // the node's hand-downs, which its children inherit
    u32* n_m2l;  // per node
    u32* n_p2p;
    u32* n_inh;
    u32* tile_m2l;  // per kTileNodes nodes
    u32* tile_p2p;
    u32* tile_leaf;
    u32* carry_m2l;  // per tile: its first row entry
    u32* carry_p2p;
    u32 width;
};

// Counts a walk's events and keeps the first kBuffered (codes: x in bits 0-28, the kind in
// 29-30, the side in 31) in eight scalar registers (selects, not an indexed array, which
// the compiler would place in scratch memory).
struct CountSink {
    u32 m2l = 0, p2p = 0, hand = 0, events = 0;
    u32 e0 = 0, e1 = 0, e2 = 0, e3 = 0, e4 = 0, e5 = 0, e6 = 0, e7 = 0;
    __device__ void keep(u32 code) {
        e0 = events == 0 ? code : e0;
        e1 = events == 1 ? code : e1;
        e2 = events == 2 ? code : e2;
        e3 = events == 3 ? code : e3;
        e4 = events == 4 ? code : e4;
        e5 = events == 5 ? code : e5;
        e6 = events == 6 ? code : e6;
        e7 = events == 7 ? code : e7;
        ++events;
    }
    __device__ void on_m2l(int x) {
        ++m2l;
        keep((u32)x | kEvM2L);
    }
    __device__ void on_p2p(int x) {
        ++p2p;
        keep((u32)x | kEvP2P);
    }
    __device__ void on_hand(int x, bool second) {
        ++hand;
        keep((u32)x | kEvHand | (second ? kSecond : 0u));
    }
};

// Writes a walk's events from the offsets it holds: M2L and P2P entries into the node's
// slots, a hand-down into its hand-down slot.  Entries past the width are
// counted but not written.
struct WriteSink {
    u32 m2l, p2p, hand, width;
    u32* m2l_slot;
    u32* p2p_slot;
    u32* inh;     // the node's hand-down slot
    __device__ void on_m2l(int x) {
        if (m2l < width) m2l_slot[m2l] = (u32)x;
        ++m2l;
    }
    __device__ void on_p2p(int x) {
        if (p2p < width) p2p_slot[p2p] = (u32)x;
        ++p2p;
    }
    __device__ void on_hand(int x, bool second) {
        if (hand < width) inh[hand] = entry_with(x, second);
        ++hand;
    }
    __device__ void replay(u32 code) {
        const int x = (int)(code & ~(kEvKind | kSecond));
        const u32 kind = code & kEvKind;
        if (kind == kEvM2L)
            on_m2l(x);
        else if (kind == kEvP2P)
            on_p2p(x);
        else
            on_hand(x, (code & kSecond) != 0);
    }
};

// One entry of node a: its self pair (a leaf's own P2P entry; an internal node's children
// get their sibling pairs at the end of its wave), or a walk depth first over the
// subtree of the other cell x0, without a stack (children in order, back up through the
// parents once a fourth child is done).
template <class Sink>
__device__ __forceinline__ void run_entry(const RowsTree& t, int a, int4 fa, double4 ga,
                                          u32 entry, double theta, Sink& sink) {
    const int x0 = (int)(entry & ~kSecond);
    if (x0 == a) {
        if (fa.z < 0) sink.on_p2p(a);
        return;
    }
    const bool second = (entry & kSecond) != 0;
    int x = x0;
    while (true) {
        int child = -1;
        const int kind = decide(t, ga, fa, x, second, theta, &child);
        if (kind == 4) {
            x = child;
            continue;
        }
        if (kind == 1)
            sink.on_m2l(x);
        else if (kind == 2)
            sink.on_p2p(x);
        else if (kind == 3)
            sink.on_hand(x, second);
        while (x != x0 && ((x - t.sib0) & 3) == 3) x = t.parent[x];
        if (x == x0) break;
        ++x;
    }
}

__device__ __forceinline__ u32 wave_excl_u32(u32 v, u32* total) {
    const int lane = td3_lane();
    u32 incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const u32 y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
    }
    *total = __shfl(incl, 63, 64);
    return incl - v;
}

// The nodes [lo, hi) of one level, a wavefront each: entries = the parent's hand-downs,
// then the four sibling pairs (the root: its own self pair).
__global__ __launch_bounds__(64 * kWaves) void rows_level_kernel(RowsTree t, i64 lo, i64 hi,
                                                                  double theta, Slots sl,
                                                                  u32* __restrict__ flag) {
    const int lane = td3_lane();
    const int wave = threadIdx.x / 64;
    const i64 node = lo + (i64)blockIdx.x * kWaves + wave;
    if (node >= hi) return;  // the whole wavefront
    const int a = (int)node;
    const size_t slot = (size_t)a * sl.width;
    const int4 fa = t.info[a];
    const double4 ga = t.geo[a];
    u32 total_m = 0, total_p = 0, total_h = 0;
    if (fa.y > fa.x) {
        u32 entries = 1, inh_n = 0;
        int first = 0, k = 0;
        const u32* hand = nullptr;
        if (a != 0) {
            const int p = t.parent[a];
            k = (a - t.sib0) & 3;
            first = a - k;
            inh_n = sl.n_inh[p];
            inh_n = inh_n < sl.width ? inh_n : sl.width;
            hand = sl.inh + (size_t)p * sl.width;
            entries = inh_n + 4;
        }
        for (u32 base = 0; base < entries; base += 64) {
            const u32 i = base + lane;
            const bool valid = i < entries;
            u32 entry = 0u;
            if (valid && a != 0) {
                const int j = (int)(i - inh_n);
                entry = i < inh_n ? hand[i] : entry_with(first + j, j < k);
            }
            CountSink c;
            if (valid) run_entry(t, a, fa, ga, entry, theta, c);
            u32 sum_m, sum_p, sum_h;
            const u32 om = wave_excl_u32(c.m2l, &sum_m);
            const u32 op = wave_excl_u32(c.p2p, &sum_p);
            const u32 oh = wave_excl_u32(c.hand, &sum_h);
            if (c.events) {
                WriteSink w = {total_m + om, total_p + op, total_h + oh, sl.width,
                               sl.m2l + slot, sl.p2p + slot, sl.inh + slot};
                if (c.events <= (u32)kBuffered) {
                    const u32 ev[kBuffered] = {c.e0, c.e1, c.e2, c.e3, c.e4, c.e5, c.e6, c.e7};
#pragma unroll
                    for (int e = 0; e < kBuffered; ++e)
                        if ((u32)e < c.events) w.replay(ev[e]);
                } else
                    run_entry(t, a, fa, ga, entry, theta, w);
            }
            total_m += sum_m;
            total_p += sum_p;
            total_h += sum_h;
        }
    }
    if (lane == 0) {
        sl.n_m2l[a] = total_m;
        sl.n_p2p[a] = total_p;
        sl.n_inh[a] = total_h;
        u32 most = total_m > total_p ? total_m : total_p;
        most = most > total_h ? most : total_h;
        if (most > sl.width) {
            atomicOr(flag, 1u);
            atomicMax(flag + 1, most);
        }
    }
}

// The nodes [lo, hi) of one level, a wavefront each, breadth first: every round the 64
// lanes take the next 64 items (the node's entries in order, then the cells queued by
// earlier rounds), decide them together (one memory latency per round instead of one per
// visited cell of a lane's depth-first walk), write their events at offsets from ballots,
// and queue the four children of every split cell behind the rest (a ring in LDS).  Rows
// come in round order, the same on every run.  A queue that would outgrow the ring sets
// bit 2 of flag[0]; td3_rows then builds the rows again with rows_level_kernel.
__global__ __launch_bounds__(64 * kWaves) void rows_level_bfs_kernel(RowsTree t, i64 lo, i64 hi,
                                                                      double theta, Slots sl,
                                                                      u32* __restrict__ flag) {
    __shared__ u32 s_ring[kWaves * kRing];
    const int lane = td3_lane();
    const int wave = threadIdx.x / 64;
    const i64 node = lo + (i64)blockIdx.x * kWaves + wave;
    if (node >= hi) return;  // the whole wavefront (no block barrier below)
    u32* ring = s_ring + wave * kRing;
    const int a = (int)node;
    const size_t slot = (size_t)a * sl.width;
    const int4 fa = t.info[a];
    const double4 ga = t.geo[a];
    const u64 below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    u32 total_m = 0, total_p = 0, total_h = 0;
    bool full = false;
    if (fa.y > fa.x) {
        u32 entries = 1, inh_n = 0;
        int first = 0, k = 0;
        const u32* hand = nullptr;
        if (a != 0) {
            const int p = t.parent[a];
            k = (a - t.sib0) & 3;
            first = a - k;
            inh_n = sl.n_inh[p];
            inh_n = inh_n < sl.width ? inh_n : sl.width;
            hand = sl.inh + (size_t)p * sl.width;
            entries = inh_n + 4;
        }
        u32 taken = 0, head = 0, tail = 0;  // entries taken; the ring's live items
        while (taken < entries || head < tail) {
            u32 n, item = 0u;
            if (taken < entries) {
                n = entries - taken < 64u ? entries - taken : 64u;
                if ((u32)lane < n) {
                    const u32 i = taken + lane;
                    const int j = (int)(i - inh_n);
                    item = a == 0 ? 0u : (i < inh_n ? hand[i] : entry_with(first + j, j < k));
                }
                taken += n;
            } else {
                n = tail - head < 64u ? tail - head : 64u;
                if ((u32)lane < n) item = ring[(head + lane) & (kRing - 1)];
                head += n;
            }
            const bool valid = (u32)lane < n;
            const int x = (int)(item & ~kSecond);
            const bool second = (item & kSecond) != 0;
            int kind = 0, child = -1;
            if (valid) {
                if (x == a)
                    kind = fa.z < 0 ? 2 : 0;  // a leaf's own P2P entry
                else
                    kind = decide(t, ga, fa, x, second, theta, &child);
            }
            const u64 bm = __ballot(kind == 1), bp = __ballot(kind == 2),
                      bh = __ballot(kind == 3), bs = __ballot(kind == 4);
            if (kind == 1) {
                const u32 pos = total_m + (u32)__popcll(bm & below);
                if (pos < sl.width) sl.m2l[slot + pos] = (u32)x;
            } else if (kind == 2) {
                const u32 pos = total_p + (u32)__popcll(bp & below);
                if (pos < sl.width) sl.p2p[slot + pos] = (u32)x;
            } else if (kind == 3) {
                const u32 pos = total_h + (u32)__popcll(bh & below);
                if (pos < sl.width) sl.inh[slot + pos] = entry_with(x, second);
            }
            total_m += (u32)__popcll(bm);
            total_p += (u32)__popcll(bp);
            total_h += (u32)__popcll(bh);
            const u32 pushed = 4u * (u32)__popcll(bs);
            if (pushed) {
                if (tail - head + pushed > (u32)kRing) {
                    full = true;
                    break;
                }
                // every lane has read its item from the ring before any lane writes to it
                __builtin_amdgcn_wave_barrier();
                if (kind == 4) {
                    const u32 pos = tail + 4u * (u32)__popcll(bs & below);
                    for (int c = 0; c < 4; ++c)
                        ring[(pos + c) & (kRing - 1)] = entry_with(child + c, second);
                }
                tail += pushed;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
            }
        }
    }
    if (lane == 0) {
        sl.n_m2l[a] = total_m;
        sl.n_p2p[a] = total_p;
        sl.n_inh[a] = total_h;
        u32 most = total_m > total_p ? total_m : total_p;
        most = most > total_h ? most : total_h;
        if (full) atomicOr(flag, 4u);
        if (most > sl.width) {
            atomicOr(flag, 1u);
            atomicMax(flag + 1, most);
        }
    }
}

// Sum of a u32 over the block (kTileNodes threads); every thread gets it.
__device__ __forceinline__ u32 block_sum_u32(u32 v, u32* s_wave) {
    u32 total;
    (void)wave_excl_u32(v, &total);
    __syncthreads();
    if (td3_lane() == 0) s_wave[threadIdx.x / 64] = total;
    __syncthreads();
    u32 all = 0;
    for (int w = 0; w < kTileNodes / 64; ++w) all += s_wave[w];
    return all;
}

// The M2L and P2P entries of every tile of kTileNodes nodes.
// Exclusive prefix over the block in thread order; *total gets the block's sum.
__device__ __forceinline__ u32 block_excl_u32(u32 v, u32* s_wave, u32* total) {
    u32 wave_total;
    const u32 in_wave = wave_excl_u32(v, &wave_total);
    __syncthreads();
    if (td3_lane() == 0) s_wave[threadIdx.x / 64] = wave_total;
    __syncthreads();
    u32 before = 0, all = 0;
    const int wave = threadIdx.x / 64;
    for (int w = 0; w < kTileNodes / 64; ++w) {
        before += w < wave ? s_wave[w] : 0u;
        all += s_wave[w];
    }
    *total = all;
    return before + in_wave;
}

// The M2L entries, P2P entries and non-empty leaves (each holds one self pair in P2P) of
// every tile of kTileNodes nodes.
__global__ __launch_bounds__(kTileNodes) void rows_tile_kernel(i64 count, Slots sl,
                                                               const int4* __restrict__ info) {
    __shared__ u32 s_wave[kTileNodes / 64];
    const i64 node = (i64)blockIdx.x * kTileNodes + threadIdx.x;
    bool full_leaf = false;
    if (node < count) {
        const int4 f = info[node];
        full_leaf = f.z < 0 && f.y > f.x;
    }
    const u32 m = block_sum_u32(node < count ? sl.n_m2l[node] : 0u, s_wave);
    const u32 p = block_sum_u32(node < count ? sl.n_p2p[node] : 0u, s_wave);
    const u32 l = block_sum_u32(full_leaf ? 1u : 0u, s_wave);
    if (threadIdx.x == 0) {
        sl.tile_m2l[blockIdx.x] = m;
        sl.tile_p2p[blockIdx.x] = p;
        sl.tile_leaf[blockIdx.x] = l;
    }
}

// The tiles' row starts (an exclusive scan of their counts, one block) and the non-empty
// leaves of the tree (flag[2]).
__global__ __launch_bounds__(kTileNodes) void rows_carry_kernel(i64 tiles, Slots sl,
                                                                u32* __restrict__ flag) {
    __shared__ u32 s_wave[kTileNodes / 64];
    u32 run_m = 0, run_p = 0, leaves = 0;
    for (i64 base = 0; base < tiles; base += kTileNodes) {
        const i64 t = base + threadIdx.x;
        const bool in = t < tiles;
        u32 sum_m, sum_p;
        const u32 em = block_excl_u32(in ? sl.tile_m2l[t] : 0u, s_wave, &sum_m);
        const u32 ep = block_excl_u32(in ? sl.tile_p2p[t] : 0u, s_wave, &sum_p);
        leaves += block_sum_u32(in ? sl.tile_leaf[t] : 0u, s_wave);
        if (in) {
            sl.carry_m2l[t] = run_m + em;
            sl.carry_p2p[t] = run_p + ep;
        }
        run_m += sum_m;
        run_p += sum_p;
    }
    if (threadIdx.x == 0) flag[2] = leaves;
}

// Packs the slots of the tile's nodes into their rows, which follow each other: output
// entry idx of the tile belongs to the last node whose row starts at or before it.
__device__ __forceinline__ void pack_rows(const u32* s_start, u32 tile_total, u32 carry, i64 tile0,
                                          i64 count, const u32* __restrict__ slots, u32 width,
                                          u32 cap, int* __restrict__ col, int* __restrict__ dst) {
    const int nodes = (int)((count - tile0) < kTileNodes ? (count - tile0) : kTileNodes);
    for (u32 idx = threadIdx.x; idx < tile_total; idx += kTileNodes) {
        int lo = 0, hi = nodes - 1;  // the last j with s_start[j] <= idx
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_start[mid] <= idx)
                lo = mid;
            else
                hi = mid - 1;
        }
        const u32 i = idx - s_start[lo];
        const u32 pos = carry + idx;
        if (i < width && pos < cap) {
            const i64 node = tile0 + lo;
            col[pos] = (int)slots[(size_t)node * width + i];
            if (dst != nullptr) dst[pos] = (int)node;
        }
    }
}

// One block per tile of kTileNodes nodes: the rows before the tile (the tile sums), the
// row starts of its nodes (a block scan), then the packed rows.  The last tile writes the
// totals and row[count]; totals past the capacities set the overflow flag.
__global__ __launch_bounds__(kTileNodes) void rows_place_kernel(
    i64 count, Slots sl, u32 cap_m2l, u32 cap_p2p, u32* __restrict__ m2l_row,
    int* __restrict__ m2l_col, int* __restrict__ m2l_dst, u32* __restrict__ p2p_row,
    int* __restrict__ p2p_col, u32* __restrict__ totals, u32* __restrict__ flag) {
    __shared__ u32 s_wave[kTileNodes / 64];
    __shared__ u32 s_start_m[kTileNodes], s_start_p[kTileNodes];
    const u32 tile = blockIdx.x;
    const i64 tile0 = (i64)tile * kTileNodes;
    const u32 carry_m = sl.carry_m2l[tile], carry_p = sl.carry_p2p[tile];
    const i64 node = tile0 + threadIdx.x;
    const u32 nm = node < count ? sl.n_m2l[node] : 0u;
    const u32 np = node < count ? sl.n_p2p[node] : 0u;
    u32 tile_m, tile_p;
    const u32 em = block_excl_u32(nm, s_wave, &tile_m);
    const u32 ep = block_excl_u32(np, s_wave, &tile_p);
    s_start_m[threadIdx.x] = em;
    s_start_p[threadIdx.x] = ep;
    if (node < count) {
        m2l_row[node] = carry_m + em;
        p2p_row[node] = carry_p + ep;
    }
    if (node == count - 1) {
        const u32 all_m = carry_m + tile_m, all_p = carry_p + tile_p;
        m2l_row[count] = all_m;
        p2p_row[count] = all_p;
        totals[0] = all_m;
        totals[1] = all_p;
        if (all_m > cap_m2l || all_p > cap_p2p) atomicOr(flag, 2u);
    }
    __syncthreads();
    pack_rows(s_start_m, tile_m, carry_m, tile0, count, sl.m2l, sl.width, cap_m2l, m2l_col, m2l_dst);
    pack_rows(s_start_p, tile_p, carry_p, tile0, count, sl.p2p, sl.width, cap_p2p, p2p_col, nullptr);
}

// Parent links.
__global__ void rows_parent_kernel(const int4* __restrict__ info, i64 count, int* __restrict__ parent) {
    const i64 node = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= count) return;
    if (node == 0) parent[0] = -1;
    const int4 f = info[node];
    if (f.z >= 0)
        for (int k = 0; k < 4; ++k) parent[f.z + k] = (int)node;
}

__global__ void rows_geo_kernel(const double* __restrict__ center, const double* __restrict__ half,
                                i64 count, double4* __restrict__ geo) {
    const i64 node = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (node < count) geo[node] = make_double4(center[2 * node], center[2 * node + 1], half[node], 0.0);
}

__global__ void rows_empty_kernel(const u32* __restrict__ flag, i64 count,
                                  u32* __restrict__ m2l_row, u32* __restrict__ p2p_row) {
    if (!__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > count) return;
    m2l_row[i] = 0u;
    p2p_row[i] = 0u;
}

// {M2L entries, P2P entries, non-empty leaves, overflow: 0, else the width needed (when a
// node's slots overflowed) or 1 (when only the rows' capacity did)}
__global__ void rows_counts_kernel(const u32* totals, const u32* flag, int64_t* out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        out[0] = totals[0];
        out[1] = totals[1];
        out[2] = flag[2];
        out[3] = (flag[0] & 1u) ? (int64_t)flag[1] : (flag[0] ? 1 : 0);
    }
}

struct RowsWs {
    Slots sl;
    u32 *flag, *totals;  // flag[4] and the tile sums are one block, cleared together
    int* parent;
    double4* geo;
    i64 tiles;
};

size_t rows_carve(RowsWs* w, void* p, i64 count, i64 width) {
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return p ? (void*)((char*)p + o) : nullptr;
    };
    const size_t slots = sizeof(u32) * (size_t)count * (size_t)width;
    w->tiles = (count + kTileNodes - 1) / kTileNodes;
    w->sl.m2l = (u32*)take(slots);
    w->sl.p2p = (u32*)take(slots);
    w->sl.inh = (u32*)take(slots);
    w->sl.n_m2l = (u32*)take(sizeof(u32) * (size_t)count);
    w->sl.n_p2p = (u32*)take(sizeof(u32) * (size_t)count);
    w->sl.n_inh = (u32*)take(sizeof(u32) * (size_t)count);
    w->sl.width = (u32)width;
    // flag[4], then the two tile-sum arrays, contiguous (one memset)
    w->flag = (u32*)take(sizeof(u32) * (4 + 2 * (size_t)w->tiles));
    w->sl.tile_m2l = w->flag ? w->flag + 4 : nullptr;
    w->sl.tile_p2p = w->flag ? w->flag + 4 + w->tiles : nullptr;
    w->totals = (u32*)take(sizeof(u32) * 2);
    w->sl.tile_leaf = (u32*)take(sizeof(u32) * (size_t)w->tiles);
    w->sl.carry_m2l = (u32*)take(sizeof(u32) * (size_t)w->tiles);
    w->sl.carry_p2p = (u32*)take(sizeof(u32) * (size_t)w->tiles);
    w->parent = (int*)take(sizeof(int) * (size_t)count);
    w->geo = (double4*)take(sizeof(double4) * (size_t)count);
    return off;
}

}  // namespace

extern "C" size_t td3_rows_workspace_bytes(const int64_t* levels_host, int64_t n_levels,
                                           int64_t width) {
    RowsWs w;
    return rows_carve(&w, nullptr, levels_host[n_levels], width);
}

extern "C" int td3_rows(const double* node_center, const double* node_half, const void* node_info,
                        const int64_t* levels_host, int64_t n_levels, double theta,
                        int64_t cap_m2l, int64_t cap_p2p, int64_t width,
                        uint32_t* m2l_row, int32_t* m2l_col, int32_t* m2l_dst, uint32_t* p2p_row,
                        int32_t* p2p_col, void* workspace, size_t workspace_bytes,
                        int64_t* out_counts, int64_t* device_counts, void* stream) {
    TD3_ARG(levels_host != nullptr && n_levels >= 1 && n_levels <= 64, "levels");
    const i64 count = levels_host[n_levels];
    TD3_ARG(count >= 1 && count < (1 << 29), "node count");  // event codes hold 29 bits
    TD3_ARG(cap_m2l >= 1 && cap_m2l < 0xffffffffll && cap_p2p >= 1 && cap_p2p < 0xffffffffll &&
                width >= 8 && width < (1 << 24), "capacities");
    TD3_ARG((u64)count * (u64)width < (1ull << 40), "slot width");
    TD3_ARG(out_counts != nullptr || device_counts != nullptr, "out_counts");
    TD3_ARG(workspace_bytes >= td3_rows_workspace_bytes(levels_host, n_levels, width),
            "workspace too small");
    hipStream_t s = (hipStream_t)stream;
    RowsWs w;
    rows_carve(&w, workspace, count, width);
    RowsTree t;
    t.geo = w.geo;
    t.info = (const int4*)node_info;
    t.parent = w.parent;
    t.sib0 = (int)(levels_host[1] - levels_host[0]);
    hipLaunchKernelGGL(rows_geo_kernel, dim3(td3_blocks(count, 256)), dim3(256), 0, s, node_center,
                       node_half, count, w.geo);
    TD3_LAUNCHED();
  for (int depth_first = 0; depth_first < 2; ++depth_first) {
    TD3_HIP(hipMemsetAsync(w.flag, 0, sizeof(u32) * (4 + 2 * (size_t)w.tiles), s));
    hipLaunchKernelGGL(rows_parent_kernel, dim3(td3_blocks(count, 256)), dim3(256), 0, s,
                       (const int4*)node_info, count, w.parent);
    TD3_LAUNCHED();
    for (int d = 0; d < n_levels; ++d) {
        const i64 lo = levels_host[d], hi = levels_host[d + 1];
        if (hi <= lo) continue;
        if (depth_first)
            TD3_TIMED(s, "rows_level",
                      hipLaunchKernelGGL(rows_level_kernel, dim3(td3_blocks(hi - lo, kWaves)),
                                         dim3(64 * kWaves), 0, s, t, lo, hi, theta, w.sl,
                                         w.flag));
        else
            TD3_TIMED(s, "rows_level",
                      hipLaunchKernelGGL(rows_level_bfs_kernel, dim3(td3_blocks(hi - lo, kWaves)),
                                         dim3(64 * kWaves), 0, s, t, lo, hi, theta, w.sl,
                                         w.flag));
        TD3_LAUNCHED();
    }
    TD3_TIMED(s, "rows_place",
              hipLaunchKernelGGL(rows_tile_kernel, dim3((unsigned)w.tiles), dim3(kTileNodes), 0,
                                 s, count, w.sl, (const int4*)node_info));
    TD3_LAUNCHED();
    TD3_TIMED(s, "rows_place",
              hipLaunchKernelGGL(rows_carry_kernel, dim3(1), dim3(kTileNodes), 0, s, w.tiles,
                                 w.sl, w.flag));
    TD3_LAUNCHED();
    TD3_TIMED(s, "rows_place",
              hipLaunchKernelGGL(rows_place_kernel, dim3((unsigned)w.tiles), dim3(kTileNodes), 0,
                                 s, count, w.sl, (u32)cap_m2l, (u32)cap_p2p, m2l_row,
                                 (int*)m2l_col, (int*)m2l_dst, p2p_row, (int*)p2p_col, w.totals,
                                 w.flag));
    TD3_LAUNCHED();
    if (device_counts != nullptr) {
        // no read-back: the counts go to the device, and an overflow leaves every row
        // empty (safe to evaluate; the caller checks the counts and builds again)
        hipLaunchKernelGGL(rows_empty_kernel, dim3(td3_blocks(count + 1, 256)), dim3(256), 0, s,
                           w.flag, count, m2l_row, p2p_row);
        TD3_LAUNCHED();
        hipLaunchKernelGGL(rows_counts_kernel, dim3(1), dim3(1), 0, s, w.totals, w.flag,
                           device_counts);
        TD3_LAUNCHED();
        return TD3_OK;
    }
    u32 host[6];
    TD3_HIP(hipMemcpyAsync(host, w.flag, sizeof(u32) * 4, hipMemcpyDeviceToHost, s));
    TD3_HIP(hipMemcpyAsync(host + 4, w.totals, sizeof(u32) * 2, hipMemcpyDeviceToHost, s));
    TD3_HIP(hipStreamSynchronize(s));
    if ((host[0] & 4u) && !depth_first) continue;  // a queue outgrew its ring: depth first
    out_counts[0] = host[4];
    out_counts[1] = host[5];
    out_counts[2] = host[2];
    out_counts[3] = (host[0] & 1u) ? (int64_t)host[1] : (host[0] ? 1 : 0);
    if (host[0]) {
        td3_set_error("interaction rows overflowed (m2l %u, p2p %u, slot width %lld, needed %u)",
                      host[4], host[5], (long long)width, host[1]);
        return TD3_ERR_CAPACITY;
    }
    return TD3_OK;
  }
  return TD3_OK;
}
