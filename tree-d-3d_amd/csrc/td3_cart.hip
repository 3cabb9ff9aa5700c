// The Coulomb-kernel FMM with Cartesian expansions of the softened kernel
// f(t) = (|t|^2 + eps^2)^(-1/2), to total degree p.
//
// Replaces the reference's CPU-only Coulomb FMM (flat_tree.py:1360-1428 and fmm_at's
// Coulomb branch :1247-1295; kernels/cart_moments_upward, cart_m2l_host, cart_downward,
// cart_evaluate_host, cart_targets_host; tables of _cartesian.py:40-107).  Coefficient
// (a, b) of total degree n = a + b lives at n (n + 1) / 2 + b, as in the reference.
//
// The M2L  T_b += sum_a (-1)^|a| C(a + b, a) m_a A_{a+b}(t)  is rewritten with
//   m'_a = (-1)^|a| m_a / a!   (per cell, once)   and   A'_g = g! A_g   (per pair),
// so  T_b = (1 / b!) sum_a m'_a A'_{a+b}: a coefficient-free correlation.  A' follows the
// reference's recurrence multiplied through by g!:
//   n R^2 A'_g = -(2n-1)(t_x g_x A'_{g-e_x} + t_y g_y A'_{g-e_y})
//                - (n-1)(g_x (g_x-1) A'_{g-2e_x} + g_y (g_y-1) A'_{g-2e_y}),  A'_0 = 1/R,
// whose integer coefficients are compile-time constants.  For p <= 12 the translation is
// fully unrolled per order (A' and the accumulators in registers); one lane per directed
// pair and the segmented wave reduction of the log-kernel M2L.

#include "td3_common.h"

namespace {

__host__ __device__ constexpr int cidx(int a, int b) { return (a + b) * (a + b + 1) / 2 + b; }
__host__ __device__ constexpr int csize(int p) { return (p + 1) * (p + 2) / 2; }
// n! for n < 32 as a table the compiler folds where n is known
struct CartFactorials {
    double v[32];
    constexpr CartFactorials() : v() {
        double f = 1.0;
        v[0] = 1.0;
        for (int i = 1; i < 32; ++i) v[i] = f *= i;
    }
};
__device__ constexpr CartFactorials kCartFact{};

// -------------------------------------------------------------------------- moments
// m_g = sum_j q_j dx^gx dy^gy over a leaf's particles; one thread per (leaf, g).
__global__ void cart_p2m_kernel(const double* __restrict__ spos, const double* __restrict__ scharge,
                                i64 n, const double* __restrict__ center,
                                const int4* __restrict__ info, i64 count, int p, int nc,
                                double* __restrict__ mom) {
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count * nc) return;
    const i64 node = t / nc;
    const int i = (int)(t % nc);
    const int4 nf = info[node];
    if (nf.z >= 0) return;
    int deg = 0;
    while (csize(deg) <= i) ++deg;
    const int b = i - deg * (deg + 1) / 2, a = deg - b;
    const double cx = center[2 * node], cy = center[2 * node + 1];
    double acc = 0.0;
    for (int j = nf.x; j < nf.y; ++j) {
        const double dx = spos[j] - cx, dy = spos[n + j] - cy;
        double v = scharge[j];
        for (int k = 0; k < a; ++k) v *= dx;
        for (int k = 0; k < b; ++k) v *= dy;
        acc += v;
    }
    mom[node * nc + i] = acc;
}

// cart_p2m_kernel with a thread per (leaf, a) for orders up to 15: the thread forms
// q dx^a once per particle and runs up the powers of dy, adding q dx^a dy^b into each
// m_(a,b) of its column.  Every term is the same product in the same order as there
// (q, then a factors dx, then b factors dy), summed over the particles in the same order:
// the moments are bit for bit cart_p2m_kernel's.
constexpr int kCartP2mCol = 16;
__global__ __launch_bounds__(256) void cart_p2m_col_kernel(
    const double* __restrict__ spos, const double* __restrict__ scharge, i64 n,
    const double* __restrict__ center, const int4* __restrict__ info, i64 count, int p, int nc,
    double* __restrict__ mom) {
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count * (p + 1)) return;
    const i64 node = t / (p + 1);
    const int a = (int)(t - node * (p + 1));
    const int4 nf = info[node];
    if (nf.z >= 0) return;
    const double cx = center[2 * node], cy = center[2 * node + 1];
    double acc[kCartP2mCol];
#pragma unroll
    for (int b = 0; b < kCartP2mCol; ++b) acc[b] = 0.0;
    for (int j = nf.x; j < nf.y; ++j) {
        const double dx = spos[j] - cx, dy = spos[n + j] - cy;
        double v = scharge[j];
        for (int k = 0; k < a; ++k) v *= dx;
#pragma unroll
        for (int b = 0; b < kCartP2mCol; ++b) {
            if (b <= p - a) {
                acc[b] += v;
                v *= dy;
            }
        }
    }
#pragma unroll
    for (int b = 0; b < kCartP2mCol; ++b) {
        if (b <= p - a) {
            const int deg = a + b;
            mom[node * nc + deg * (deg + 1) / 2 + b] = acc[b];
        }
    }
}

// M2M for the nodes [lo, hi): m_g += sum_children sum_{b <= g} C(g,b) m_b(child) s^(g-b).
__global__ void cart_m2m_kernel(const double* __restrict__ center, const int4* __restrict__ info,
                                i64 lo, i64 hi, int p, int nc, const double* __restrict__ binom,
                                int bsz, double* __restrict__ mom) {
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (hi - lo) * nc) return;
    const i64 node = lo + t / nc;
    const int i = (int)(t % nc);
    const int4 nf = info[node];
    if (nf.z < 0) return;
    int deg = 0;
    while (csize(deg) <= i) ++deg;
    const int gy = i - deg * (deg + 1) / 2, gx = deg - gy;
    const double cx = center[2 * node], cy = center[2 * node + 1];
    double acc = 0.0;
    for (int c = nf.z; c < nf.z + 4; ++c) {
        const int4 cf = info[c];
        if (cf.y == cf.x) continue;
        const double sx = center[2 * c] - cx, sy = center[2 * c + 1] - cy;
        // s^(g - b) built up as b runs down from g (no power loops per term)
        double px = 1.0;
        for (int bx = gx; bx >= 0; --bx) {
            const double fx = binom[gx * bsz + bx] * px;
            double py = 1.0;
            for (int by = gy; by >= 0; --by) {
                acc += fx * binom[gy * bsz + by] * mom[(i64)c * nc + cidx(bx, by)] * py;
                py *= sy;
            }
            px *= sx;
        }
    }
    mom[node * nc + i] = acc;
}

// The same with a wavefront per node (orders up to 13: nc <= kCartStage): the four
// children's moments, the powers of their offsets and the binomials are staged in LDS once,
// and the shift is done one axis at a time (sum_{b <= g} C(g, b) s^(g - b) m_b factors into
// a shift in x followed by one in y): at most 4 (p + 1) terms per output and pass instead of
// 4 (gx + 1)(gy + 1).  Lane l forms output l, and output 64 + l past 64 (at order 10 and 11
// the cheapest low-degree lanes).
constexpr int kCartStage = 105;  // csize(13)
constexpr int kCartDeg = 14;     // 13 + 1

// binomials C(n, k), n, k <= p, from the caller's table into LDS (every thread of the block)
__device__ __forceinline__ void cart_stage_binomials(double (*s_bin)[kCartDeg],
                                                     const double* __restrict__ binom, int bsz,
                                                     int p) {
    for (int i = threadIdx.x; i < (p + 1) * (p + 1); i += blockDim.x)
        s_bin[i / (p + 1)][i % (p + 1)] = binom[(i / (p + 1)) * bsz + i % (p + 1)];
    __syncthreads();
}

__device__ __forceinline__ void cart_degree_split(int i, int& gx, int& gy) {
    int deg = 0;
    while (csize(deg) <= i) ++deg;
    gy = i - deg * (deg + 1) / 2;
    gx = deg - gy;
}

__global__ __launch_bounds__(256) void cart_m2m_lds_kernel(
    const double* __restrict__ center, const int4* __restrict__ info, i64 lo, i64 hi, int p,
    int nc, const double* __restrict__ binom, int bsz, double* __restrict__ mom) {
    __shared__ double s_bin[kCartDeg][kCartDeg];
    __shared__ double s_m[4][4][kCartStage];
    __shared__ double s_pw[4][4][2][kCartDeg];
    cart_stage_binomials(s_bin, binom, bsz, p);
    const int w = (int)(threadIdx.x / 64), lane = td3_lane();
    const i64 node = lo + ((i64)blockIdx.x * blockDim.x + threadIdx.x) / 64;
    if (node >= hi) return;  // uniform over the wavefront
    const int4 nf = info[node];
    if (nf.z < 0) return;
    const double cx = center[2 * node], cy = center[2 * node + 1];
    for (int c = 0; c < 4; ++c) {
        const int4 cf = info[nf.z + c];
        for (int i = lane; i < nc; i += 64)
            s_m[w][c][i] = cf.y == cf.x ? 0.0 : mom[(i64)(nf.z + c) * nc + i];
        // s^k of the child's offset, k <= p, by repeated products
        if (lane < 2) {
            const double sv = center[2 * (nf.z + c) + lane] - (lane ? cy : cx);
            double pw = 1.0;
            for (int k = 0; k <= p; ++k) {
                s_pw[w][c][lane][k] = pw;
                pw *= sv;
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // the shift in x: Y_c(gx, by) = sum_{bx <= gx} C(gx, bx) s_x^(gx - bx) m_c(bx, by), in
    // place of each child's moments (entries of the same degree-ordered layout; entry
    // (gx, by) reads only (bx <= gx, by), so each child's entries are formed from the
    // highest gx down, one gx at a time)
    for (int gx = p; gx >= 0; --gx) {
        for (int it = lane; it < 4 * (p - gx + 1); it += 64) {
            const int c = it / (p - gx + 1), by = it % (p - gx + 1);
            double acc = 0.0;
            for (int bx = gx; bx >= 0; --bx)
                acc += s_bin[gx][bx] * s_pw[w][c][0][gx - bx] * s_m[w][c][cidx(bx, by)];
            s_m[w][c][cidx(gx, by)] = acc;  // read by lower gx only after this row is done
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    // the shift in y and the sum over the children
    for (int i = lane; i < nc; i += 64) {
        int gx, gy;
        cart_degree_split(i, gx, gy);
        double acc = 0.0;
        for (int c = 0; c < 4; ++c)
            for (int by = gy; by >= 0; --by)
                acc += s_bin[gy][by] * s_pw[w][c][1][gy - by] * s_m[w][c][cidx(gx, by)];
        mom[node * nc + i] = acc;
    }
}

// m'_a = (-1)^|a| m_a / a!
__global__ void cart_prime_kernel(const double* __restrict__ mom, i64 count, int nc,
                                  double* __restrict__ out) {
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count * nc) return;
    const int i = (int)(t % nc);
    int deg = 0;
    while (csize(deg) <= i) ++deg;
    const int b = i - deg * (deg + 1) / 2, a = deg - b;
    double f = 1.0;
    for (int k = 2; k <= a; ++k) f *= k;
    for (int k = 2; k <= b; ++k) f *= k;
    out[t] = ((deg & 1) ? -1.0 : 1.0) * mom[t] / f;
}

// -------------------------------------------------------------------------- M2L
// One lane per directed pair over 64 consecutive pairs of the destination-sorted CSR: the
// lane's T_b stay in registers, then go through LDS 32 at a time, where the lanes of each
// destination are summed in lane order (rows within the 64 written divided by b!, the
// partial sums of rows that cross a wave boundary to side, primed, for the fixup).
//
// T = double: the reference's arithmetic on m'_a.  T = float (mixed precision): the same
// translation on radius-scaled quantities, so that every term stays in FP32 range at any
// depth.  The softened kernel is homogeneous, A'_g(t; eps) = h^-(|g|+1) A'_g(t/h; eps/h),
// so with h the destination's half size and mscaled_a = m'_a / h_src^|a| (per cell, once):
//   T_b = (1 / b!) h^-(|b|+1) sum_a (mscaled_a kappa^|a|) A'_{a+b}(t/h; eps/h),
// kappa = h_src / h.  A'(t/h) is bounded by the separation of the pair, mscaled by the
// cell's charge: the recurrence, the products and the lane's sums run in FP32, the sums
// over a destination's pairs and the final scaling in FP64.
constexpr int kCartGroup = 32;

template <int P, typename T>
__global__ __launch_bounds__(256) void cart_m2l_kernel(
    const int* __restrict__ col, const int* __restrict__ dst, const u32* __restrict__ row_start,
    u32 e_lo, u32 e_hi, const double* __restrict__ dcenter, const double* __restrict__ scenter,
    const double* __restrict__ dhalf, const double* __restrict__ shalf,
    const T* __restrict__ mprime, double eps2, double* __restrict__ local,
    double* __restrict__ side) {
    constexpr int NC = csize(P);
    constexpr bool kScaled = sizeof(T) == 4;
    __shared__ double s_t[4][kCartGroup][64];
    __shared__ int s_seg[4][65];
    const int lane = td3_lane(), w = (int)(threadIdx.x / 64);
    const u32 wave = (u32)((blockIdx.x * blockDim.x + threadIdx.x) / 64);
    const u32 w0 = e_lo + wave * 64;
    if (w0 >= e_hi) return;  // uniform over the wavefront
    const u32 e = w0 + lane;
    const bool valid = e < e_hi;
    const int d = valid ? dst[e] : -1;
    T Tb[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) Tb[i] = (T)0;
    if (valid) {
        const int src = col[e];
        double tx = dcenter[2 * d] - scenter[2 * src];
        double ty = dcenter[2 * d + 1] - scenter[2 * src + 1];
        double e2 = eps2;
        T kappa = (T)1;
        if (kScaled) {
            const double inv_h = 1.0 / dhalf[d];
            tx *= inv_h;
            ty *= inv_h;
            e2 *= inv_h * inv_h;
            kappa = (T)(shalf[src] * inv_h);
        }
        const double r2 = tx * tx + ty * ty + e2;
        const T txt = (T)tx, tyt = (T)ty;
        const T inv_r2 = (T)(1.0 / r2);
        T A[NC];
        A[0] = (T)(1.0 / sqrt(r2));
#pragma unroll
        for (int n = 1; n <= P; ++n) {
#pragma unroll
            for (int gy = 0; gy <= n; ++gy) {
                const int gx = n - gy;
                T s = (T)0;
                if (gx >= 1) s += (T)((2 * n - 1) * gx) * txt * A[cidx(gx - 1, gy)];
                if (gy >= 1) s += (T)((2 * n - 1) * gy) * tyt * A[cidx(gx, gy - 1)];
                if (gx >= 2) s += (T)((n - 1) * gx * (gx - 1)) * A[cidx(gx - 2, gy)];
                if (gy >= 2) s += (T)((n - 1) * gy * (gy - 1)) * A[cidx(gx, gy - 2)];
                A[cidx(gx, gy)] = -s * inv_r2 * (T)(1.0 / n);
            }
        }
        const T* m = mprime + (i64)src * NC;
        T kpow = (T)1;
#pragma unroll
        for (int na = 0; na <= P; ++na) {
#pragma unroll
            for (int ay = 0; ay <= na; ++ay) {
                const int ax = na - ay;
                const T ma = kScaled ? m[cidx(ax, ay)] * kpow : m[cidx(ax, ay)];
#pragma unroll
                for (int nb = 0; nb + na <= P; ++nb) {
#pragma unroll
                    for (int by = 0; by <= nb; ++by) {
                        const int bx = nb - by;
                        Tb[cidx(bx, by)] += ma * A[cidx(ax + bx, ay + by)];
                    }
                }
            }
            kpow *= kappa;
        }
    }
    // segments of equal destination among the valid lanes, by their first lane
    const int prev_d = __shfl_up(d, 1, 64);
    const bool head = valid && (lane == 0 || prev_d != d);
    const u64 heads = __ballot(head);
    const int nseg = __popcll(heads);
    if (head) s_seg[w][__popcll(heads & ((1ull << lane) - 1ull))] = lane;
    if (lane == 0) s_seg[w][nseg] = (int)(e_hi - w0 < 64 ? e_hi - w0 : 64);
#pragma unroll
    for (int g0 = 0; g0 < NC; g0 += kCartGroup) {
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
        for (int i = 0; i < kCartGroup; ++i)
            if (g0 + i < NC) s_t[w][i][lane] = (double)Tb[g0 + i];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const int width = NC - g0 < kCartGroup ? NC - g0 : kCartGroup;
        for (int it = lane; it < nseg * width; it += 64) {
            const int sg = it / width, i = it % width, b = g0 + i;
            const int l0 = s_seg[w][sg], l1 = s_seg[w][sg + 1];
            double v = 0.0;
            for (int l = l0; l < l1; ++l) v += s_t[w][i][l];
            const int dd = dst[w0 + l0];
            int deg = 0;
            while (csize(deg) <= b) ++deg;
            if (kScaled) {
                // h^-(|b| + 1)
                const double inv_h = 1.0 / dhalf[dd];
                double sc = inv_h;
                for (int k = 0; k < deg; ++k) sc *= inv_h;
                v *= sc;
            }
            const u32 rs0 = row_start[dd], rs1 = row_start[dd + 1];
            const bool before = rs0 < w0, after = rs1 > w0 + 64;
            if (!before && !after) {
                const int by = b - deg * (deg + 1) / 2, bx = deg - by;
                local[(i64)dd * NC + b] = v / (kCartFact.v[bx] * kCartFact.v[by]);
            } else {
                // partial sums stay primed: the fixup divides by b!
                side[((i64)((w0 - e_lo) / 64) * 2 + (before ? 0 : 1)) * NC + b] = v;
            }
        }
    }
}

// mscaled_a = m'_a / h^|a| of every cell, in FP32 (the mixed-precision M2L's moments).
__global__ void cart_scale_prime_kernel(const double* __restrict__ mprime,
                                        const double* __restrict__ half, i64 count, int nc,
                                        float* __restrict__ out) {
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count * nc) return;
    const i64 node = t / nc;
    const int i = (int)(t % nc);
    int deg = 0;
    while (csize(deg) <= i) ++deg;
    const double inv_h = 1.0 / half[node];
    double sc = 1.0;
    for (int k = 0; k < deg; ++k) sc *= inv_h;
    out[t] = (float)(mprime[t] * sc);
}

__global__ void cart_m2l_fixup_kernel(const u32* __restrict__ row_start, i64 node_lo, i64 node_hi,
                                      u32 e_lo, int p, int nc, const double* __restrict__ side,
                                      double* __restrict__ local) {
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    const i64 node = node_lo + t / nc;
    const int i = (int)(t % nc);
    if (node >= node_hi) return;
    const u32 rs0 = row_start[node], rs1 = row_start[node + 1];
    if (rs1 == rs0) return;
    const u32 wa = (rs0 - e_lo) / 64, wb = (rs1 - 1 - e_lo) / 64;
    if (wa == wb) return;
    double v = side[((i64)wa * 2 + 1) * nc + i];
    for (u32 w = wa + 1; w <= wb; ++w) v += side[((i64)w * 2) * nc + i];
    int deg = 0;
    while (csize(deg) <= i) ++deg;
    const int by = i - deg * (deg + 1) / 2, bx = deg - by;
    local[node * nc + i] = v / (kCartFact.v[bx] * kCartFact.v[by]);
}

// Orders above 12: the same correlation with runtime loops (slower, any p); one wavefront
// per destination (grid-stride), A' built by lane 0 in a per-wavefront scratch row.
constexpr unsigned kGenericBlocks = 2048;
__global__ __launch_bounds__(64) void cart_m2l_generic_kernel(
    const int* __restrict__ col, const u32* __restrict__ row_start, i64 node_lo, i64 node_hi,
    const double* __restrict__ dcenter, const double* __restrict__ scenter,
    const double* __restrict__ mprime, int p, int nc, double eps2, double* __restrict__ scratch,
    double* __restrict__ local) {
    double* A = scratch + (i64)blockIdx.x * nc;
    for (i64 dnode = node_lo + blockIdx.x; dnode < node_hi; dnode += gridDim.x) {
        const u32 r0 = row_start[dnode], r1 = row_start[dnode + 1];
        if (r1 == r0) continue;
        for (int i0 = 0; i0 < nc; i0 += 64) {
            const int i = i0 + td3_lane();
            double acc = 0.0;
            int bdeg = 0;
            if (i < nc)
                while (csize(bdeg) <= i) ++bdeg;
            const int by = i < nc ? i - bdeg * (bdeg + 1) / 2 : 0, bx = bdeg - by;
            for (u32 e = r0; e < r1; ++e) {
                const int src = col[e];
                const double tx = dcenter[2 * dnode] - scenter[2 * src];
                const double ty = dcenter[2 * dnode + 1] - scenter[2 * src + 1];
                const double r2 = tx * tx + ty * ty + eps2;
                __builtin_amdgcn_wave_barrier();
                if (td3_lane() == 0) {
                    A[0] = 1.0 / sqrt(r2);
                    for (int n = 1; n <= p; ++n)
                        for (int gy = 0; gy <= n; ++gy) {
                            const int gx = n - gy;
                            double acc2 = 0.0;
                            if (gx >= 1) acc2 += (double)((2 * n - 1) * gx) * tx * A[cidx(gx - 1, gy)];
                            if (gy >= 1) acc2 += (double)((2 * n - 1) * gy) * ty * A[cidx(gx, gy - 1)];
                            if (gx >= 2) acc2 += (double)((n - 1) * gx * (gx - 1)) * A[cidx(gx - 2, gy)];
                            if (gy >= 2) acc2 += (double)((n - 1) * gy * (gy - 1)) * A[cidx(gx, gy - 2)];
                            A[cidx(gx, gy)] = -acc2 / (n * r2);
                        }
                }
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                if (i < nc) {
                    const double* m = mprime + (i64)src * nc;
                    for (int na = 0; na + bdeg <= p; ++na)
                        for (int ay = 0; ay <= na; ++ay)
                            acc += m[cidx(na - ay, ay)] * A[cidx(na - ay + bx, ay + by)];
                }
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            }
            if (i < nc) {
                double f = 1.0;
                for (int k = 2; k <= bx; ++k) f *= k;
                for (int k = 2; k <= by; ++k) f *= k;
                local[dnode * nc + i] = acc / f;
            }
        }
    }
}

// -------------------------------------------------------------------------- L2L
// T_b(child) += sum_{g >= b} C(gx,bx) C(gy,by) T_g(parent) s^(g-b); thread per (child, b)
// of the children of the nodes [lo, hi).
__global__ void cart_l2l_kernel(const double* __restrict__ center, const int4* __restrict__ info,
                                i64 lo, i64 hi, int p, int nc, const double* __restrict__ binom,
                                int bsz, double* __restrict__ local) {
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    const i64 parent = lo + t / (4 * nc);
    if (parent >= hi) return;
    const int4 pf = info[parent];
    if (pf.z < 0) return;
    const int rem = (int)(t % (4 * nc));
    const i64 child = pf.z + rem / nc;
    const int i = rem % nc;
    int deg = 0;
    while (csize(deg) <= i) ++deg;
    const int by = i - deg * (deg + 1) / 2, bx = deg - by;
    const double sx = center[2 * child] - center[2 * parent];
    const double sy = center[2 * child + 1] - center[2 * parent + 1];
    double acc = 0.0;
    // s^(g - b) built up with g (no power loops per term)
    double px = 1.0;
    for (int gx = bx; gx <= p; ++gx) {
        const double fx = binom[gx * bsz + bx] * px;
        double py = 1.0;
        for (int gy = by; gx + gy <= p; ++gy) {
            acc += fx * binom[gy * bsz + by] * local[parent * nc + cidx(gx, gy)] * py;
            py *= sy;
        }
        px *= sx;
    }
    local[child * nc + i] += acc;
}

// The same with a wavefront per parent: its coefficients, the powers of its children's
// offsets and the binomials staged in LDS once for the four children's 4 nc outputs.  Same
// terms in the same order as cart_l2l_kernel.
__global__ __launch_bounds__(256) void cart_l2l_lds_kernel(
    const double* __restrict__ center, const int4* __restrict__ info, i64 lo, i64 hi, int p,
    int nc, const double* __restrict__ binom, int bsz, double* __restrict__ local) {
    __shared__ double s_bin[kCartDeg][kCartDeg];
    __shared__ double s_l[4][kCartStage];
    __shared__ double s_pw[4][4][2][kCartDeg];
    cart_stage_binomials(s_bin, binom, bsz, p);
    const int w = (int)(threadIdx.x / 64), lane = td3_lane();
    const i64 parent = lo + ((i64)blockIdx.x * blockDim.x + threadIdx.x) / 64;
    if (parent >= hi) return;  // uniform over the wavefront
    const int4 pf = info[parent];
    if (pf.z < 0) return;
    for (int i = lane; i < nc; i += 64) s_l[w][i] = local[parent * nc + i];
    if (lane < 8) {
        const int c = lane >> 1, axis = lane & 1;
        const double sv = center[2 * (pf.z + c) + axis] - center[2 * parent + axis];
        double pw = 1.0;
        for (int k = 0; k <= p; ++k) {
            s_pw[w][c][axis][k] = pw;
            pw *= sv;
        }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    for (int o = lane; o < 4 * nc; o += 64) {
        const int c = o / nc, i = o % nc;
        int bx, by;
        cart_degree_split(i, bx, by);
        double acc = 0.0;
        for (int gx = bx; gx <= p; ++gx) {
            const double fx = s_bin[gx][bx] * s_pw[w][c][0][gx - bx];
            for (int gy = by; gx + gy <= p; ++gy)
                acc += fx * s_bin[gy][by] * s_l[w][cidx(gx, gy)] * s_pw[w][c][1][gy - by];
        }
        local[(pf.z + c) * (i64)nc + i] += acc;
    }
}

template <int P, typename T>
int launch_cart_m2l(hipStream_t s, const int* col, const int* dst, const u32* row_start, u32 e_lo,
                    u32 e_hi, i64 node_lo, i64 node_hi, const double* dcenter,
                    const double* scenter, const double* dhalf, const double* shalf,
                    const T* mprime, double eps2, double* local, double* side) {
    constexpr int NC = csize(P);
    const i64 waves = ((i64)(e_hi - e_lo) + 63) / 64;
    TD3_TIMED(s, "cart_m2l", hipLaunchKernelGGL((cart_m2l_kernel<P, T>), dim3(td3_blocks(waves * 64, 256)), dim3(256), 0, s,
                       col, dst, row_start, e_lo, e_hi, dcenter, scenter, dhalf, shalf, mprime, eps2,
                       local, side));
    TD3_LAUNCHED();
    const i64 items = (node_hi - node_lo) * NC;
    hipLaunchKernelGGL(cart_m2l_fixup_kernel, dim3(td3_blocks(items, 256)), dim3(256), 0, s, row_start,
                       node_lo, node_hi, e_lo, P, NC, side, local);
    TD3_LAUNCHED();
    return TD3_OK;
}

}  // namespace

extern "C" size_t td3_cart_m2l_workspace_bytes(int64_t pair_count, int64_t order,
                                              int64_t node_count) {
    const i64 nc = csize((int)order);
    if (order <= 12) {
        const i64 waves = (pair_count + 63) / 64 + 1;
        return (size_t)waves * 2 * (size_t)nc * sizeof(double) + 256;
    }
    return (size_t)kGenericBlocks * (size_t)nc * sizeof(double) + 256;
}

extern "C" int td3_cart_moments(int64_t order, const double* sorted_pos, const double* sorted_charge,
                                int64_t n, const double* node_center, const void* node_info,
                                int64_t node_count, const int64_t* levels_host, int64_t n_levels,
                                const double* binomial, int64_t bsz, double* moments,
                                double* mprime, void* stream) {
    TD3_ARG(order >= 1 && order <= 64 && bsz >= order + 1, "order");
    TD3_ARG(n >= 1 && node_count >= 1 && levels_host[n_levels] == node_count, "sizes");
    hipStream_t s = (hipStream_t)stream;
    const int p = (int)order, nc = csize(p);
    const int4* info = (const int4*)node_info;
    if (p + 1 <= kCartP2mCol)
        TD3_TIMED(s, "cart_p2m", hipLaunchKernelGGL(cart_p2m_col_kernel, dim3(td3_blocks(node_count * (p + 1), 256)), dim3(256), 0, s,
                           sorted_pos, sorted_charge, n, node_center, info, node_count, p, nc, moments));
    else
        TD3_TIMED(s, "cart_p2m", hipLaunchKernelGGL(cart_p2m_kernel, dim3(td3_blocks(node_count * nc, 256)), dim3(256), 0, s,
                           sorted_pos, sorted_charge, n, node_center, info, node_count, p, nc, moments));
    TD3_LAUNCHED();
    for (i64 d = n_levels - 2; d >= 0; --d) {
        const i64 lo = levels_host[d], hi = levels_host[d + 1];
        if (nc <= kCartStage)
            TD3_TIMED(s, "cart_m2m", hipLaunchKernelGGL(cart_m2m_lds_kernel, dim3(td3_blocks((hi - lo) * 64, 256)), dim3(256), 0, s,
                      node_center, info, lo, hi, p, nc, binomial, (int)bsz, moments));
        else
            TD3_TIMED(s, "cart_m2m", hipLaunchKernelGGL(cart_m2m_kernel, dim3(td3_blocks((hi - lo) * nc, 256)), dim3(256), 0, s,
                      node_center, info, lo, hi, p, nc, binomial, (int)bsz, moments));
        TD3_LAUNCHED();
    }
    hipLaunchKernelGGL(cart_prime_kernel, dim3(td3_blocks(node_count * nc, 256)), dim3(256), 0, s, moments,
                       node_count, nc, mprime);
    TD3_LAUNCHED();
    return TD3_OK;
}

extern "C" int td3_cart_m2l(int64_t order, const int32_t* col, const int32_t* dst,
                            const uint32_t* row_start, int64_t node_lo, int64_t node_hi,
                            const double* dst_center, const double* src_center,
                            const double* mprime, double softening, double* local,
                            void* workspace, size_t workspace_bytes, void* stream) {
    TD3_ARG(order >= 1 && order <= 64, "order");
    TD3_ARG(node_lo >= 0 && node_hi >= node_lo, "node range");
    hipStream_t s = (hipStream_t)stream;
    if (node_hi == node_lo) return TD3_OK;
    u32 bounds[2];
    TD3_HIP(hipMemcpyAsync(&bounds[0], row_start + node_lo, sizeof(u32), hipMemcpyDeviceToHost, s));
    TD3_HIP(hipMemcpyAsync(&bounds[1], row_start + node_hi, sizeof(u32), hipMemcpyDeviceToHost, s));
    TD3_HIP(hipStreamSynchronize(s));
    const i64 pairs = (i64)bounds[1] - (i64)bounds[0];
    if (pairs <= 0) return TD3_OK;
    TD3_ARG(workspace_bytes >= td3_cart_m2l_workspace_bytes(pairs, order, node_hi - node_lo),
            "workspace too small");
    const double eps2 = softening * softening;
    double* side = (double*)workspace;
    switch ((int)order) {
#define TD3_CART_CASE(PP) \
    case PP:              \
        return launch_cart_m2l<PP, double>(s, col, dst, row_start, bounds[0], bounds[1], node_lo,    \
                                           node_hi, dst_center, src_center, nullptr, nullptr,      \
                                           mprime, eps2, local, side);
        TD3_CART_CASE(1)
        TD3_CART_CASE(2)
        TD3_CART_CASE(3)
        TD3_CART_CASE(4)
        TD3_CART_CASE(5)
        TD3_CART_CASE(6)
        TD3_CART_CASE(7)
        TD3_CART_CASE(8)
        TD3_CART_CASE(9)
        TD3_CART_CASE(10)
        TD3_CART_CASE(11)
        TD3_CART_CASE(12)
#undef TD3_CART_CASE
        default:
            break;
    }
    const int p = (int)order, nc = csize(p);
    const unsigned gblocks = (node_hi - node_lo) < (i64)kGenericBlocks ? (unsigned)(node_hi - node_lo) : kGenericBlocks;
    TD3_TIMED(s, "cart_m2l", hipLaunchKernelGGL(cart_m2l_generic_kernel, dim3(gblocks), dim3(64), 0, s,
                       col, row_start, node_lo, node_hi, dst_center, src_center, mprime, p, nc, eps2,
                       side, local));
    TD3_LAUNCHED();
    return TD3_OK;
}

extern "C" int td3_cart_scale_prime(int64_t order, const double* mprime, const double* node_half,
                                    int64_t node_count, float* out, void* stream) {
    TD3_ARG(order >= 1 && order <= 12 && node_count >= 0, "order (1..12), sizes");
    if (node_count == 0) return TD3_OK;
    const int nc = csize((int)order);
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(cart_scale_prime_kernel, dim3(td3_blocks(node_count * nc, 256)), dim3(256),
                       0, s, mprime, node_half, node_count, nc, out);
    TD3_LAUNCHED();
    return TD3_OK;
}

extern "C" int td3_cart_m2l_f32(int64_t order, const int32_t* col, const int32_t* dst,
                                const uint32_t* row_start, int64_t node_lo, int64_t node_hi,
                                const double* dst_center, const double* dst_half,
                                const double* src_center, const double* src_half,
                                const float* mscaled, double softening, double* local,
                                void* workspace, size_t workspace_bytes, void* stream) {
    TD3_ARG(order >= 1 && order <= 12, "order (1..12)");
    TD3_ARG(node_lo >= 0 && node_hi >= node_lo, "node range");
    hipStream_t s = (hipStream_t)stream;
    if (node_hi == node_lo) return TD3_OK;
    u32 bounds[2];
    TD3_HIP(hipMemcpyAsync(&bounds[0], row_start + node_lo, sizeof(u32), hipMemcpyDeviceToHost, s));
    TD3_HIP(hipMemcpyAsync(&bounds[1], row_start + node_hi, sizeof(u32), hipMemcpyDeviceToHost, s));
    TD3_HIP(hipStreamSynchronize(s));
    const i64 pairs = (i64)bounds[1] - (i64)bounds[0];
    if (pairs <= 0) return TD3_OK;
    TD3_ARG(workspace_bytes >= td3_cart_m2l_workspace_bytes(pairs, order, node_hi - node_lo),
            "workspace too small");
    const double eps2 = softening * softening;
    double* side = (double*)workspace;
    switch ((int)order) {
#define TD3_CART_CASE(PP) \
    case PP:              \
        return launch_cart_m2l<PP, float>(s, col, dst, row_start, bounds[0], bounds[1], node_lo,    \
                                          node_hi, dst_center, src_center, dst_half, src_half,     \
                                          mscaled, eps2, local, side);
        TD3_CART_CASE(1)
        TD3_CART_CASE(2)
        TD3_CART_CASE(3)
        TD3_CART_CASE(4)
        TD3_CART_CASE(5)
        TD3_CART_CASE(6)
        TD3_CART_CASE(7)
        TD3_CART_CASE(8)
        TD3_CART_CASE(9)
        TD3_CART_CASE(10)
        TD3_CART_CASE(11)
        TD3_CART_CASE(12)
#undef TD3_CART_CASE
        default:
            break;
    }
    return TD3_ERR_ARG;
}

extern "C" int td3_cart_l2l(int64_t order, const double* node_center, const void* node_info,
                            const int64_t* levels_host, int64_t n_levels, const double* binomial,
                            int64_t bsz, double* local, void* stream) {
    TD3_ARG(order >= 1 && order <= 64 && bsz >= order + 1, "order");
    hipStream_t s = (hipStream_t)stream;
    const int p = (int)order, nc = csize(p);
    for (i64 d = 0; d + 1 < n_levels; ++d) {
        const i64 lo = levels_host[d], hi = levels_host[d + 1];
        if (nc <= kCartStage)
            TD3_TIMED(s, "cart_l2l", hipLaunchKernelGGL(cart_l2l_lds_kernel, dim3(td3_blocks((hi - lo) * 64, 256)), dim3(256), 0, s,
                      node_center, (const int4*)node_info, lo, hi, p, nc, binomial, (int)bsz, local));
        else
            TD3_TIMED(s, "cart_l2l", hipLaunchKernelGGL(cart_l2l_kernel, dim3(td3_blocks((hi - lo) * 4 * nc, 256)), dim3(256), 0, s,
                      node_center, (const int4*)node_info, lo, hi, p, nc, binomial, (int)bsz, local));
        TD3_LAUNCHED();
    }
    return TD3_OK;
}
