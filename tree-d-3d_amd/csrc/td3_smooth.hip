// Smooth-membership FMM of line charges: energy and exact-gradient fields on static cells.
//
// Replaces smooth.py:122-193 (BoxTree.memberships), :280-380 (SmoothFMM._compute: moments
// over weighted memberships, the dual walk over occupied cells, M2L, L2L, L2P and the near
// field over membership routes), :410-422 (energy, fields) and :196-201 / :424-440 (the
// direct sums with the compact softening).  The NumPy prototype loops over routes; here
// every membership is one slot of a particle's four, and the FMM between them reuses the
// log kernel's stages:
//   1. memberships: one lane per particle descends the cells keeping every child whose box
//      grown by the band holds it (at most four cells per level when the band is at most
//      the smallest leaf half-width), and weighs each leaf found by the product of four
//      quintic edge steps, in the reference's operation order without contraction, so the
//      weights and the membership sets are those of the reference bit for bit;
//   2. layout: the memberships are sorted by the depth-first rank of their leaf, so every
//      cell's memberships are one contiguous range ({start, end, child, depth} per cell,
//      the node_info of td3_moments / td3_rows) and a cell is occupied when its range is
//      not empty (the reference's `occupied`);
//   3. td3_moments (log) on the memberships' positions and charges q w, td3_rows on that
//      node_info (the same traversal rules as smooth.py:246-277), td3_m2l with either
//      truncation, td3_l2l;
//   4. one lane per membership: L2P at its leaf plus the exact compact kernel over the
//      memberships of the leaf's P2P row, other particles only (Phi_m, G_m);
//   5. one lane per particle: phi = -2 sum w Phi_m, E = 2 sum (w G_m + Phi_m grad w) over
//      its slots, and its share -q sum w Phi_m of the energy, summed by a fixed tree of
//      partial sums (deterministic).

#include "td3_sort.h"

namespace {

constexpr int kSlots = 4;

inline int bits_for(i64 v) {
    int b = 1;
    while (b < 63 && (1ll << b) <= v) ++b;
    return b;
}

// ---------------------------------------------------------------- leaf order of the cells
// span[c] = leaves below cell c (1 for a leaf), one level at a time from the deepest
__global__ void span_kernel(const int4* __restrict__ info, i64 lo, i64 hi, u32* __restrict__ span) {
    const i64 c = lo + (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= hi) return;
    const int f = info[c].z;
    span[c] = f < 0 ? 1u : span[f] + span[f + 1] + span[f + 2] + span[f + 3];
}

// rank[child k] = rank[c] + spans of the children before it, one level at a time from the
// root: the depth-first rank of a cell's first leaf
__global__ void rank_kernel(const int4* __restrict__ info, i64 lo, i64 hi,
                            const u32* __restrict__ span, u32* __restrict__ rank) {
    const i64 c = lo + (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= hi) return;
    if (c == 0) rank[0] = 0;
    const int f = info[c].z;
    if (f < 0) return;
    u32 r = rank[c];
    for (int k = 0; k < 4; ++k) {
        rank[f + k] = r;
        r += span[f + k];
    }
}

// ---------------------------------------------------------------- memberships
struct Weights {
    double w, dwx, dwy;
};

// smooth.py:172-191 for one particle and one leaf, operation by operation
__device__ Weights leaf_weights(double x, double y, double cx, double cy, double h, double band,
                                double rx, double ry, double rh) {
#pragma clang fp contract(off)
    double s[4], d[4];
    const double t[4] = {(x - (cx - h)) / band, ((cx + h) - x) / band, (y - (cy - h)) / band,
                         ((cy + h) - y) / band};
    const double tol = 1e-12 * rh;
    const bool on_root[4] = {cx - h <= rx - rh + tol, cx + h >= rx + rh - tol,
                             cy - h <= ry - rh + tol, cy + h >= ry + rh - tol};
    for (int e = 0; e < 4; ++e) {
        double u = (t[e] + 1.0) / 2.0;
        u = u < 0.0 ? 0.0 : (u > 1.0 ? 1.0 : u);
        const double one_minus = 1.0 - u;
        const double si = u * u * u * (10.0 + u * (-15.0 + 6.0 * u));
        const double di = 0.5 * (30.0 * u * u * (one_minus * one_minus));
        s[e] = on_root[e] ? 1.0 : si;
        d[e] = on_root[e] ? 0.0 : di;
    }
    const double wx = s[0] * s[1], wy = s[2] * s[3];
    Weights out;
    out.w = wx * wy;
    out.dwx = (d[0] * s[1] - s[0] * d[1]) / band * wy;
    out.dwy = (d[2] * s[3] - s[2] * d[3]) / band * wx;
    return out;
}

// One lane per particle: up to four (leaf, w, dw/dx, dw/dy) slots, and the sort key of each
// slot (its leaf's depth-first rank; leaf_count for an empty slot).  status: 1 + the index
// of a particle that reached more than four cells on one level (only possible when
// rounding puts it in five grown cells at once).
__global__ void members_kernel(const double* __restrict__ pos, i64 n,
                               const double* __restrict__ center, const double* __restrict__ half,
                               const int4* __restrict__ info, const u32* __restrict__ rank,
                               i64 leaf_count, double rx, double ry, double rh, double band,
                               int* __restrict__ slot_leaf, double* __restrict__ slot_weight,
                               int* __restrict__ slot_member, u64* __restrict__ keys,
                               i64* __restrict__ status) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = pos[i], y = pos[n + i];
    int found[kSlots];
    int nf = 0;
    if (band <= 0.0) {
        // smooth.py:125-143: right when x > cx, down when y <= cy
        int c = 0;
        for (int f = info[0].z; f >= 0; f = info[c].z) {
            const double cx = center[2 * c], cy = center[2 * c + 1];
            c = f + (x > cx ? 1 : 0) + (y <= cy ? 2 : 0);
        }
        found[nf++] = c;
    } else {
        // smooth.py:147-167: the frontier holds cells of one depth
        int front[kSlots], next[kSlots];
        int nfr = 1;
        front[0] = 0;
        bool over = false;
        while (nfr > 0) {
            int nn = 0;
            for (int k = 0; k < nfr; ++k) {
                const int c = front[k];
                const int f = info[c].z;
                if (f < 0) {
                    if (nf < kSlots) found[nf++] = c;
                    else over = true;
                    continue;
                }
                for (int q = 0; q < 4; ++q) {
                    const int cc = f + q;
                    const double reach = half[cc] + band;
                    if (fabs(x - center[2 * cc]) < reach && fabs(y - center[2 * cc + 1]) < reach) {
                        if (nn < kSlots) next[nn++] = cc;
                        else over = true;
                    }
                }
            }
            for (int k = 0; k < nn; ++k) front[k] = next[k];
            nfr = nn;
        }
        if (over) td3_report(status, i + 1);
    }
    int used = 0;
    for (int k = 0; k < nf; ++k) {
        const int c = found[k];
        Weights wt;
        if (band <= 0.0) {
            wt.w = 1.0;
            wt.dwx = wt.dwy = 0.0;
        } else {
            wt = leaf_weights(x, y, center[2 * c], center[2 * c + 1], half[c], band, rx, ry, rh);
            if (!(wt.w > 0.0)) continue;  // smooth.py:192: w > 0 only
        }
        const i64 s = kSlots * i + used++;
        slot_leaf[s] = c;
        slot_weight[s] = wt.w;
        slot_weight[kSlots * n + s] = wt.dwx;
        slot_weight[2 * kSlots * n + s] = wt.dwy;
        keys[s] = rank[c];
    }
    for (; used < kSlots; ++used) {
        const i64 s = kSlots * i + used;
        slot_leaf[s] = -1;
        slot_weight[s] = 0.0;
        slot_weight[kSlots * n + s] = 0.0;
        slot_weight[2 * kSlots * n + s] = 0.0;
        keys[s] = (u64)leaf_count;
    }
    for (int k = 0; k < kSlots; ++k) slot_member[kSlots * i + k] = -1;
}

// mstart[r] = first sorted slot whose leaf rank is >= r, for r = 0..leaf_count
__global__ void starts_kernel(const u64* __restrict__ keys, i64 total, i64 leaf_count,
                              u32* __restrict__ mstart) {
    const i64 r = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (r > leaf_count) return;
    i64 lo = 0, hi = total;
    while (lo < hi) {
        const i64 mid = (lo + hi) >> 1;
        if ((i64)keys[mid] < r)
            lo = mid + 1;
        else
            hi = mid;
    }
    mstart[r] = (u32)lo;
}

// Sorted membership j: position, q w, particle, leaf, and the slot's sorted index
__global__ void gather_kernel(const u64* __restrict__ keys, const u32* __restrict__ vals,
                              i64 total, i64 leaf_count, const double* __restrict__ pos,
                              const double* __restrict__ charge, i64 n,
                              const int* __restrict__ slot_leaf,
                              const double* __restrict__ slot_weight, int* __restrict__ slot_member,
                              double* __restrict__ mpos, double* __restrict__ mcharge,
                              double4* __restrict__ msrc, int* __restrict__ mleaf) {
    const i64 j = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= total) return;
    if ((i64)keys[j] >= leaf_count) {
        mleaf[j] = -1;
        return;
    }
    const u32 slot = vals[j];
    const i64 i = slot / kSlots;
    const double x = pos[i], y = pos[n + i];
    const double qw = charge[i] * slot_weight[slot];  // smooth.py:295: q[part] * w
    mpos[j] = x;
    mpos[total + j] = y;
    mcharge[j] = qw;
    msrc[j] = make_double4(x, y, qw, (double)i);
    mleaf[j] = slot_leaf[slot];
    slot_member[slot] = (int)j;
}

// node_info of the memberships: the cell's range of sorted memberships, its children and
// depth from the cells' own node_info; member_count[0] = all memberships
__global__ void member_info_kernel(const int4* __restrict__ cell_info, i64 count,
                                   const u32* __restrict__ rank, const u32* __restrict__ span,
                                   const u32* __restrict__ mstart, i64 leaf_count,
                                   int4* __restrict__ member_info, i64* __restrict__ member_count) {
    const i64 c = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= count) return;
    const int4 ci = cell_info[c];
    const u32 r = rank[c];
    member_info[c] = make_int4((int)mstart[r], (int)mstart[r + span[c]], ci.z, ci.w);
    if (c == 0) member_count[0] = (i64)mstart[leaf_count];
}

// ---------------------------------------------------------------- evaluation
struct Compact {
    double eps2, phi0;  // phi0 = ln a - 11/12 (a^2 = eps2)
};

// q Phi(d) and q grad Phi(d) of the compact cloud (smooth.py:52-72), accumulated
__device__ __forceinline__ bool add_compact(double dx, double dy, double q, Compact k,
                                           double* p, double* gx, double* gy) {
    const double r2 = dx * dx + dy * dy;
    if (k.eps2 > 0.0 && r2 < k.eps2) {
        const double u = r2 / k.eps2;
        *p += q * (k.phi0 + (3.0 * u - 1.5 * u * u + u * u * u / 3.0) / 2.0);
        const double g = q * (3.0 - 3.0 * u + u * u) / k.eps2;
        *gx += g * dx;
        *gy += g * dy;
        return true;
    }
    if (r2 == 0.0) return false;
    *p += q * (0.5 * log(r2));
    const double g = q / r2;
    *gx += g * dx;
    *gy += g * dy;
    return true;
}

// One lane per sorted membership: Phi_m = Re W + near, G_m = conj(W') + near (smooth.py:
// 344-379).  Locals are the radius-scaled B_l of td3_l2l.  status: 1 + the particle of a
// membership that met another particle at distance 0 without softening.
__global__ __launch_bounds__(256) void smooth_eval_kernel(
    const double4* __restrict__ msrc, const int* __restrict__ mleaf, i64 cap,
    const i64* __restrict__ member_count, const double* __restrict__ center,
    const double* __restrict__ half, const int4* __restrict__ minfo,
    const double* __restrict__ lre, const double* __restrict__ lim, int p1,
    const u32* __restrict__ prow, const int* __restrict__ pcol, Compact kern,
    double* __restrict__ phi_m, double* __restrict__ g_m, i64* __restrict__ status) {
    const i64 j = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cap || j >= member_count[0]) return;
    const double4 me = msrc[j];
    const int leaf = mleaf[j];
    double ap = 0.0, agx = 0.0, agy = 0.0;
    bool bad = false;
    const u32 r0 = prow[leaf], r1 = prow[leaf + 1];
    for (u32 r = r0; r < r1; ++r) {
        const int4 sf = minfo[pcol[r]];
        for (int k = sf.x; k < sf.y; ++k) {
            const double4 s = msrc[k];
            if (s.w == me.w) continue;  // the same particle (itself, or in another leaf)
            if (!add_compact(me.x - s.x, me.y - s.y, s.z, kern, &ap, &agx, &agy)) bad = true;
        }
    }
    if (bad) td3_report(status, (i64)me.w + 1);
    // Horner for W and W' on the scaled locals at u / r
    const double inv_r = 1.0 / (1.4142135623730951 * half[leaf]);
    const double ur = (me.x - center[2 * leaf]) * inv_r, ui = (me.y - center[2 * leaf + 1]) * inv_r;
    const double* bre = lre + (i64)leaf * p1;
    const double* bim = lim + (i64)leaf * p1;
    double w_re = bre[p1 - 1], w_im = bim[p1 - 1], d_re = 0.0, d_im = 0.0;
    for (int l = p1 - 2; l >= 0; --l) {
        const double nd = d_re * ur - d_im * ui + w_re;
        d_im = d_re * ui + d_im * ur + w_im;
        d_re = nd;
        const double nw = w_re * ur - w_im * ui + bre[l];
        w_im = w_re * ui + w_im * ur + bim[l];
        w_re = nw;
    }
    phi_m[j] = w_re + ap;
    g_m[j] = d_re * inv_r + agx;
    g_m[cap + j] = -d_im * inv_r + agy;
}

// Sum of partial[0:count) by one block in a fixed order into out[0]
__global__ __launch_bounds__(256) void final_sum_kernel(const double* __restrict__ partial, i64 count,
                                                        double* __restrict__ out) {
    __shared__ double s_wave[4];
    double v = 0.0;
    for (i64 k = threadIdx.x; k < count; k += 256) v += partial[k];
    v = td3_wave_sum(v);
    if (td3_lane() == 0) s_wave[threadIdx.x / 64] = v;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = (s_wave[0] + s_wave[1]) + (s_wave[2] + s_wave[3]);
}

// block total of v into partial[blockIdx.x]
__device__ __forceinline__ void block_partial(double v, double* partial) {
    __shared__ double s_wave[4];
    v = td3_wave_sum(v);
    if (td3_lane() == 0) s_wave[threadIdx.x / 64] = v;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = (s_wave[0] + s_wave[1]) + (s_wave[2] + s_wave[3]);
}

// One lane per particle over its slots (smooth.py:410-422)
__global__ __launch_bounds__(256) void combine_kernel(
    i64 n, i64 cap, const double* __restrict__ charge, const int* __restrict__ slot_member,
    const double* __restrict__ slot_weight, const double* __restrict__ phi_m,
    const double* __restrict__ g_m, double* __restrict__ phi, double* __restrict__ field,
    double* __restrict__ partial) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    double share = 0.0;
    if (i < n) {
        double p = 0.0, ex = 0.0, ey = 0.0, wphi = 0.0;
        for (int k = 0; k < kSlots; ++k) {
            const i64 s = kSlots * i + k;
            const int j = slot_member[s];
            if (j < 0) continue;
            const double w = slot_weight[s];
            const double dwx = slot_weight[kSlots * n + s], dwy = slot_weight[2 * kSlots * n + s];
            const double pm = phi_m[j];
            p += -2.0 * w * pm;
            ex += 2.0 * (w * g_m[j] + pm * dwx);
            ey += 2.0 * (w * g_m[cap + j] + pm * dwy);
            wphi += w * pm;
        }
        if (phi != nullptr) phi[i] = p;
        if (field != nullptr) {
            field[i] = ex;
            field[n + i] = ey;
        }
        share = -charge[i] * wphi;
    }
    if (partial != nullptr) block_partial(share, partial);
}

// Direct sums of the compact kernel (smooth.py:196-201, :424-440): one lane per target,
// sources staged through LDS 256 at a time.  self_pairs: the targets are the sources, and
// source t is skipped for target t.  partial (may be null): block sums of q_t phi_t / 2.
__global__ __launch_bounds__(256) void compact_direct_kernel(
    const double* __restrict__ targets, i64 m, int self_pairs, const double* __restrict__ src,
    const double* __restrict__ charge, i64 n, Compact kern, double* __restrict__ phi,
    double* __restrict__ field, double* __restrict__ partial, i64* __restrict__ status) {
    __shared__ double4 tile[256];
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = t < m;
    const double tx = live ? targets[t] : 0.0, ty = live ? targets[m + t] : 0.0;
    double ap = 0.0, agx = 0.0, agy = 0.0;
    bool bad = false;
    for (i64 base = 0; base < n; base += 256) {
        const i64 k = base + threadIdx.x;
        __syncthreads();
        tile[threadIdx.x] = k < n ? make_double4(src[k], src[n + k], charge[k], 0.0)
                                  : make_double4(0.0, 0.0, 0.0, 0.0);
        __syncthreads();
        const int count = (int)(n - base < 256 ? n - base : 256);
        if (!live) continue;
        for (int e = 0; e < count; ++e) {
            if (self_pairs && base + e == t) continue;
            const double4 s = tile[e];
            if (!add_compact(tx - s.x, ty - s.y, s.z, kern, &ap, &agx, &agy)) bad = true;
        }
    }
    double share = 0.0;
    if (live) {
        if (bad) td3_report(status, t + 1);
        if (phi != nullptr) phi[t] = -2.0 * ap;
        if (field != nullptr) {
            field[t] = 2.0 * agx;
            field[m + t] = 2.0 * agy;
        }
        if (partial != nullptr) share = 0.5 * charge[t] * (-2.0 * ap);
    }
    if (partial != nullptr) block_partial(share, partial);
}

Compact compact_of(double softening) {
    Compact k;
    k.eps2 = softening * softening;
    k.phi0 = k.eps2 > 0.0 ? 0.5 * log(k.eps2) - 11.0 / 12.0 : 0.0;
    return k;
}

}  // namespace

extern "C" int td3_box_leaf_order(const void* node_info, int64_t node_count,
                                  const int64_t* levels_host, int64_t n_levels, uint32_t* leaf_rank,
                                  uint32_t* leaf_span, void* stream) {
    TD3_ARG(node_info != nullptr && leaf_rank != nullptr && leaf_span != nullptr, "arrays");
    TD3_ARG(node_count >= 1 && n_levels >= 1 && levels_host != nullptr &&
                levels_host[n_levels] == node_count, "levels");
    hipStream_t s = (hipStream_t)stream;
    const int4* info = (const int4*)node_info;
    for (i64 d = n_levels - 1; d >= 0; --d) {
        const i64 lo = levels_host[d], hi = levels_host[d + 1];
        if (hi <= lo) continue;
        hipLaunchKernelGGL(span_kernel, dim3(td3_blocks(hi - lo, 256)), dim3(256), 0, s, info, lo,
                           hi, leaf_span);
        TD3_LAUNCHED();
    }
    for (i64 d = 0; d < n_levels; ++d) {
        const i64 lo = levels_host[d], hi = levels_host[d + 1];
        if (hi <= lo) continue;
        hipLaunchKernelGGL(rank_kernel, dim3(td3_blocks(hi - lo, 256)), dim3(256), 0, s, info, lo,
                           hi, (const u32*)leaf_span, leaf_rank);
        TD3_LAUNCHED();
    }
    return TD3_OK;
}

extern "C" size_t td3_smooth_members_workspace_bytes(int64_t n, int64_t leaf_count) {
    const i64 total = kSlots * (n > 0 ? n : 1);
    size_t bytes = 0;
    bytes += 2 * (size_t)total * sizeof(u64);                   // keys, keys_alt
    bytes += 2 * (size_t)total * sizeof(u32);                   // vals, vals_alt
    bytes += ((size_t)(leaf_count + 1) * sizeof(u32) + 255) / 256 * 256;  // mstart
    bytes += td3sort::scratch_bytes(total) + 1024;
    return bytes;
}

extern "C" int td3_smooth_members(const double* pos, const double* charge, int64_t n,
                                  const double* node_center, const double* node_half,
                                  const void* node_info, int64_t node_count,
                                  const uint32_t* leaf_rank, const uint32_t* leaf_span,
                                  int64_t leaf_count, double root_x, double root_y,
                                  double root_half, double band, int32_t* slot_leaf,
                                  double* slot_weight, int32_t* slot_member, double* member_pos,
                                  double* member_charge, double* member_src, int32_t* member_leaf,
                                  void* member_info, int64_t* member_count, int64_t* status,
                                  void* workspace, size_t workspace_bytes, void* stream) {
    TD3_ARG(n >= 1 && n < (1ll << 29), "n (1 <= N < 2^29)");
    TD3_ARG(node_count >= 1 && leaf_count >= 1 && leaf_count <= node_count, "cell counts");
    TD3_ARG(!(band > 0.0) || band == band, "band");
    TD3_ARG(workspace_bytes >= td3_smooth_members_workspace_bytes(n, leaf_count), "workspace");
    hipStream_t s = (hipStream_t)stream;
    const i64 total = kSlots * n;
    char* ws = (char*)workspace;
    u64* keys = (u64*)ws;
    ws += total * sizeof(u64);
    u64* keys_alt = (u64*)ws;
    ws += total * sizeof(u64);
    u32* vals = (u32*)ws;
    ws += total * sizeof(u32);
    u32* vals_alt = (u32*)ws;
    ws += total * sizeof(u32);
    ws = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
    u32* mstart = (u32*)ws;
    ws += ((size_t)(leaf_count + 1) * sizeof(u32) + 255) / 256 * 256;
    u32* scratch = (u32*)ws;
    const int4* info = (const int4*)node_info;
    TD3_TIMED(s, "smooth_members", hipLaunchKernelGGL(members_kernel, dim3(td3_blocks(n, 256)), dim3(256), 0, s, pos, (i64)n,
                       node_center, node_half, info, leaf_rank, (i64)leaf_count, root_x, root_y,
                       root_half, band, slot_leaf, slot_weight, slot_member, keys, status));
    TD3_LAUNCHED();
    if (int rc = td3sort::sort_pairs(keys, vals, nullptr, keys_alt, vals_alt, nullptr, total, 0,
                                     bits_for(leaf_count + 1), true, scratch, s))
        return rc;
    hipLaunchKernelGGL(starts_kernel, dim3(td3_blocks(leaf_count + 1, 256)), dim3(256), 0, s,
                       (const u64*)keys, total, (i64)leaf_count, mstart);
    TD3_LAUNCHED();
    TD3_TIMED(s, "smooth_gather", hipLaunchKernelGGL(gather_kernel, dim3(td3_blocks(total, 256)), dim3(256), 0, s,
                       (const u64*)keys, (const u32*)vals, total, (i64)leaf_count, pos, charge,
                       (i64)n, (const int*)slot_leaf, (const double*)slot_weight, slot_member,
                       member_pos, member_charge, (double4*)member_src, member_leaf));
    TD3_LAUNCHED();
    hipLaunchKernelGGL(member_info_kernel, dim3(td3_blocks(node_count, 256)), dim3(256), 0, s, info,
                       (i64)node_count, leaf_rank, leaf_span, (const u32*)mstart, (i64)leaf_count,
                       (int4*)member_info, member_count);
    TD3_LAUNCHED();
    return TD3_OK;
}

extern "C" int td3_smooth_eval(int64_t order, const double* member_src, const int32_t* member_leaf,
                               int64_t cap, const int64_t* member_count, const double* node_center,
                               const double* node_half, const void* member_info,
                               const double* local_re, const double* local_im,
                               const uint32_t* p2p_row, const int32_t* p2p_col, double softening,
                               double* phi_m, double* g_m, int64_t* status, void* stream) {
    TD3_ARG(order >= 1 && order <= 64, "order");
    TD3_ARG(cap >= 1 && softening >= 0.0, "sizes, softening");
    hipStream_t s = (hipStream_t)stream;
    TD3_TIMED(s, "smooth_eval", hipLaunchKernelGGL(smooth_eval_kernel, dim3(td3_blocks(cap, 256)), dim3(256), 0, s,
                       (const double4*)member_src, member_leaf, (i64)cap, member_count, node_center,
                       node_half, (const int4*)member_info, local_re, local_im, (int)order + 1,
                       p2p_row, p2p_col, compact_of(softening), phi_m, g_m, status));
    TD3_LAUNCHED();
    return TD3_OK;
}

extern "C" size_t td3_smooth_combine_workspace_bytes(int64_t n) {
    return (size_t)td3_blocks(n, 256) * sizeof(double) + 256;
}

extern "C" int td3_smooth_combine(int64_t n, int64_t cap, const double* charge,
                                  const int32_t* slot_member, const double* slot_weight,
                                  const double* phi_m, const double* g_m, double* phi,
                                  double* field, double* energy, void* workspace,
                                  size_t workspace_bytes, void* stream) {
    TD3_ARG(n >= 1 && cap >= kSlots * n, "sizes");
    TD3_ARG(energy == nullptr || workspace_bytes >= td3_smooth_combine_workspace_bytes(n),
            "workspace");
    hipStream_t s = (hipStream_t)stream;
    const unsigned blocks = td3_blocks(n, 256);
    double* partial = energy != nullptr ? (double*)workspace : nullptr;
    TD3_TIMED(s, "smooth_combine", hipLaunchKernelGGL(combine_kernel, dim3(blocks), dim3(256), 0, s, (i64)n, (i64)cap, charge,
                       slot_member, slot_weight, phi_m, g_m, phi, field, partial));
    TD3_LAUNCHED();
    if (energy != nullptr) {
        hipLaunchKernelGGL(final_sum_kernel, dim3(1), dim3(256), 0, s, (const double*)partial,
                           (i64)blocks, energy);
        TD3_LAUNCHED();
    }
    return TD3_OK;
}

extern "C" size_t td3_compact_direct_workspace_bytes(int64_t m) {
    return (size_t)td3_blocks(m, 256) * sizeof(double) + 256;
}

extern "C" int td3_compact_direct(const double* targets, int64_t m, int self_pairs,
                                  const double* sources, const double* charge, int64_t n,
                                  double softening, double* phi, double* field, double* energy,
                                  int64_t* status, void* workspace, size_t workspace_bytes,
                                  void* stream) {
    TD3_ARG(m >= 1 && n >= 1 && softening >= 0.0, "sizes, softening");
    TD3_ARG(!self_pairs || m == n, "self_pairs needs the targets to be the sources");
    TD3_ARG(energy == nullptr || (self_pairs &&
                                  workspace_bytes >= td3_compact_direct_workspace_bytes(m)),
            "energy needs self_pairs and the workspace");
    hipStream_t s = (hipStream_t)stream;
    const unsigned blocks = td3_blocks(m, 256);
    double* partial = energy != nullptr ? (double*)workspace : nullptr;
    TD3_TIMED(s, "compact_direct", hipLaunchKernelGGL(compact_direct_kernel, dim3(blocks), dim3(256), 0, s, targets, (i64)m,
                       self_pairs, sources, charge, (i64)n, compact_of(softening), phi, field,
                       partial, status));
    TD3_LAUNCHED();
    if (energy != nullptr) {
        hipLaunchKernelGGL(final_sum_kernel, dim3(1), dim3(256), 0, s, (const double*)partial,
                           (i64)blocks, energy);
        TD3_LAUNCHED();
    }
    return TD3_OK;
}
