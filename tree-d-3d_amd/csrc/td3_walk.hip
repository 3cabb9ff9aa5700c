// Barnes-Hut walk, one wavefront per 64 targets.
//
// Replaces kernels/evaluate_quadtree/evaluate_quadtree_cuda.cu:9-107 and
// kernels/evaluate_log_quadtree/evaluate_log_quadtree_cuda.cu:9-104 (one thread per
// target, a 512-entry stack per thread in scratch memory, sources read through perm).
// The design follows the reference's CPU walk by groups,
// kernels/evaluate_groups/evaluate_groups_pyccel.py:16-297: a group of targets shares
// one stack of (node, lane mask); here the group is the 64 lanes of a wavefront, the
// masks are ballots, and the stack lives in LDS.  Every lane makes its own acceptance
// test (in FP64, without FMA contraction, as the reference's expression), so each target
// gets exactly the terms of its own walk, in the same depth-first order.  Targets come
// in Morton order (td3_target_order) so that a wave's targets are close; sources are
// read in tree order (sorted_pos, sorted charge), contiguous per leaf.

#include "td3_common.h"

namespace {

constexpr int kWaves = 4;
constexpr int kMaxStack = 3 * 64 + 8;

struct WalkTree {
    const double* spos;     // (2, n) tree order
    const double* scharge;  // (n,) tree order
    const i64* perm;        // (n,) input index of each tree slot
    i64 n;
    const double* center;   // (C, 2)
    const double* half;     // (C,)
    const int4* info;       // {start, end, child, depth}
    const double* qabs;
    const double* q;        // coulomb
    const double* p;        // coulomb (C, 2)
    const double* s;        // coulomb (C, 3)
    const double* mre;      // log (C, p1)
    const double* mim;
    int p1;
};

// Lane i's value in every lane (i the same in all lanes): two v_readlane_b32.
__device__ __forceinline__ u64 lane_bits(u64 b, int i) {
    const u32 lo = (u32)__builtin_amdgcn_readlane((int)(u32)b, i);
    const u32 hi = (u32)__builtin_amdgcn_readlane((int)(u32)(b >> 32), i);
    return ((u64)hi << 32) | lo;
}
__device__ __forceinline__ double lane_value(double v, int i) {
    return __builtin_bit_cast(double, lane_bits(__builtin_bit_cast(u64, v), i));
}
__device__ __forceinline__ i64 lane_value(i64 v, int i) { return (i64)lane_bits((u64)v, i); }

// COUNT: the same walk also adds up its work into counts[0..3] (td3_walk_counts): cells
// accepted per target, source terms per target, cells accepted per wavefront (a cell some
// lane of the wave accepts: its data is read once for the wave) and sources read per
// wavefront (every source of every leaf the wave visits).
template <int KERNEL, int PREC, int WAVES, bool COUNT = false>
__global__ __launch_bounds__(64 * WAVES) void walk_kernel(
    WalkTree tr, const double* __restrict__ targets, i64 m, const i64* __restrict__ order,
    const i64* __restrict__ exclude, const int* __restrict__ target_root, double theta,
    double eps2, int with_potential, int stack_size, double* __restrict__ phi,
    double* __restrict__ field, i64* __restrict__ status,
    unsigned long long* __restrict__ counts = nullptr) {
    __shared__ int s_node[WAVES][kMaxStack];
    __shared__ u64 s_mask[WAVES][kMaxStack];
    const int wave = threadIdx.x / TD3_WAVE;
    const int lane = td3_lane();
    // each XCD walks one contiguous eighth of the target order (in a forest, whole
    // systems), so its L2 holds those trees only
    const i64 slot = ((i64)td3_xcd_block() * WAVES + wave) * TD3_WAVE + lane;
    const bool active = slot < m;
    const i64 ti = active ? order[slot] : 0;
    const double tx = active ? targets[ti] : 0.0;
    const double ty = active ? targets[m + ti] : 0.0;
    const i64 skip = active ? exclude[ti] : -1;
    const u64 me = 1ull << lane;
    const u64 full = __ballot(active);
    if (full == 0) return;
    double pot = 0.0, ex = 0.0, ey = 0.0;
    bool bad = false;
    u32 n_accept = 0, n_terms = 0;       // COUNT: this lane's
    u32 w_accept = 0, w_sources = 0;     // COUNT: the wave's (uniform)
    int* stn = s_node[wave];
    u64* stm = s_mask[wave];
    // one stack entry per root among the wave's targets (node 0 for a single tree; the
    // target's system root in a forest), with the mask of the lanes it serves
    int top = 0;
    bool overflow = false;
    {
        const int my_root = active ? (target_root ? target_root[ti] : 0) : -1;
        u64 left = full;
        while (left) {
            const int lead = __ffsll((unsigned long long)left) - 1;
            const int r = __shfl(my_root, lead, 64);
            const u64 mk = __ballot(active && my_root == r);
            if (top + 4 > stack_size) {
                overflow = true;
                break;
            }
            if (lane == 0) {
                stn[top] = r;
                stm[top] = mk;
            }
            ++top;
            left &= ~mk;
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (overflow) top = 0;
    while (top > 0) {
        --top;
        __builtin_amdgcn_wave_barrier();
        const int node = __builtin_amdgcn_readfirstlane(stn[top]);
        const u64 mask = stm[top];
        if (tr.qabs[node] == 0.0) continue;
        const int4 nf = tr.info[node];
        const bool mine = (mask & me) != 0;
        if (nf.z < 0) {
            // The leaf's sources are loaded once, a lane each, and handed to every lane in
            // turn (v_readlane): one round of loads per 64 sources instead of four dependent
            // loads per source.  Each target adds the same terms in the same order.
            const int k0 = __builtin_amdgcn_readfirstlane(nf.x);
            const int k1 = __builtin_amdgcn_readfirstlane(nf.y);
            if (COUNT) w_sources += (u32)(k1 - k0);
            for (int base = k0; base < k1; base += TD3_WAVE) {
                const int kl = base + lane;
                double lc = 0.0, lx = 0.0, ly = 0.0;
                i64 lp = -1;
                if (kl < k1) {
                    lc = tr.scharge[kl];
                    lp = tr.perm[kl];
                    lx = tr.spos[kl];
                    ly = tr.spos[tr.n + kl];
                }
                const int cnt = k1 - base < TD3_WAVE ? k1 - base : TD3_WAVE;
                for (int i = 0; i < cnt; ++i) {
                    const double c = lane_value(lc, i);
                    const i64 pk = lane_value(lp, i);
                    const double sx = lane_value(lx, i), sy = lane_value(ly, i);
                    if (!mine || bad) continue;
                    if (pk == skip || c == 0.0) continue;
                    const double dx = tx - sx;
                    const double dy = ty - sy;
                    const double r2 = dx * dx + dy * dy + eps2;
                    if (r2 == 0.0) {
                        bad = true;
                        continue;
                    }
                    if (COUNT) ++n_terms;
                    if (PREC == TD3_FP64) {
                        if (KERNEL == TD3_LOG) {
                            if (with_potential) pot -= c * log(r2);
                            const double inv = 2.0 * c / r2;
                            ex += inv * dx;
                            ey += inv * dy;
                        } else {
                            // HIP's FP64 rsqrt (1 ulp) in place of 1 / sqrt (two
                            // correctly rounded operations, each an iterative sequence)
                            const double inv_r = rsqrt(r2);
                            const double cr3 = c * inv_r * inv_r * inv_r;
                            pot += c * inv_r;
                            ex += cr3 * dx;
                            ey += cr3 * dy;
                        }
                    } else {
                        const float r2f = (float)r2, cf = (float)c;
                        if (KERNEL == TD3_LOG) {
                            if (with_potential) pot -= (double)(cf * __logf(r2f));
                            const float inv = 2.0f * cf * __frcp_rn(r2f);
                            ex += (double)(inv * (float)dx);
                            ey += (double)(inv * (float)dy);
                        } else {
                            const float inv_r = __frsqrt_rn(r2f);
                            const float cr3 = cf * inv_r * inv_r * inv_r;
                            pot += (double)(cf * inv_r);
                            ex += (double)(cr3 * (float)dx);
                            ey += (double)(cr3 * (float)dy);
                        }
                    }
                }
            }
            continue;
        }
        const double xmid = tr.center[2 * node], ymid = tr.center[2 * node + 1];
        const double h = tr.half[node];
        const double rx = tx - xmid, ry = ty - ymid;
        const bool inside = tx >= xmid - h && tx <= xmid + h && ty >= ymid - h && ty <= ymid + h;
        bool acc;
        // the reference's expressions, evaluated without contraction
        const double d2 = __dadd_rn(__dmul_rn(rx, rx), __dmul_rn(ry, ry));
        if (KERNEL == TD3_LOG)
            acc = !inside && __dmul_rn(__dmul_rn(4.0, h), h) < __dmul_rn(__dmul_rn(theta, theta), d2);
        else {
            // 2h < fl(theta fl(sqrt(d2))): decided on squares unless they agree to 1e-12
            // (the right side lies within 3e-16 of theta sqrt(d2)), where the reference's
            // expression is evaluated
            const double lhs = 4.0 * h * h, rhs = theta * theta * d2;
            const bool near_tie = !(lhs < rhs * (1.0 - 1e-12)) && !(lhs > rhs * (1.0 + 1e-12));
            acc = !inside && (near_tie ? __dmul_rn(2.0, h) < __dmul_rn(theta, __dsqrt_rn(d2))
                                       : lhs < rhs);
        }
        const u64 accept = __ballot(mine && acc);
        const u64 reject = __ballot(mine && !acc);
        if (COUNT) {
            w_accept += accept != 0 ? 1u : 0u;
            n_accept += (mine && acc && !bad) ? 1u : 0u;
        }
        if (mine && acc && !bad) {
            if (KERNEL == TD3_LOG) {
                const double* are = tr.mre + (i64)node * tr.p1;
                const double* aim = tr.mim + (i64)node * tr.p1;
                if (PREC == TD3_FP64) {
                    const double w_re = rx / d2, w_im = -ry / d2;
                    double prev_re = 1.0, prev_im = 0.0, cur_re = w_re, cur_im = w_im;
                    double f_re = 0.0, f_im = 0.0, series = 0.0;
                    for (int k = 0; k < tr.p1; ++k) {
                        const double a_re = are[k], a_im = aim[k];
                        f_re += a_re * cur_re - a_im * cur_im;
                        f_im += a_re * cur_im + a_im * cur_re;
                        if (k > 0) series += (a_re * prev_re - a_im * prev_im) / k;
                        prev_re = cur_re;
                        prev_im = cur_im;
                        const double next_re = cur_re * w_re - cur_im * w_im;
                        cur_im = cur_re * w_im + cur_im * w_re;
                        cur_re = next_re;
                    }
                    if (with_potential) pot += -are[0] * log(d2) + 2.0 * series;
                    ex += 2.0 * f_re;
                    ey -= 2.0 * f_im;
                } else {
                    // FP32 series in w' = h w (|w'| < theta / 2) on moments scaled by h^-k
                    const float w_re = (float)(rx / d2 * h), w_im = (float)(-ry / d2 * h);
                    float prev_re = 1.0f, prev_im = 0.0f, cur_re = w_re, cur_im = w_im;
                    float f_re = 0.0f, f_im = 0.0f, series = 0.0f;
                    // a_k h^-k with h = mant 2^e: mant^-k (at most 2^64) times an exact
                    // power of two applied by ldexp, so no intermediate h^-k overflows
                    // (deep cells at high order) and the scaled moment stays finite
                    int h_exp;
                    const double h_mant = frexp(h, &h_exp);
                    const double inv_mant = 1.0 / h_mant;
                    double mant_pow = 1.0;
                    for (int k = 0; k < tr.p1; ++k) {
                        const float a_re = (float)ldexp(are[k] * mant_pow, -h_exp * k);
                        const float a_im = (float)ldexp(aim[k] * mant_pow, -h_exp * k);
                        mant_pow *= inv_mant;
                        f_re += a_re * cur_re - a_im * cur_im;
                        f_im += a_re * cur_im + a_im * cur_re;
                        if (k > 0) series += (a_re * prev_re - a_im * prev_im) / (float)k;
                        prev_re = cur_re;
                        prev_im = cur_im;
                        const float next_re = cur_re * w_re - cur_im * w_im;
                        cur_im = cur_re * w_im + cur_im * w_re;
                        cur_re = next_re;
                    }
                    if (with_potential) pot += -are[0] * log(d2) + 2.0 * (double)series;
                    ex += 2.0 * (double)f_re / h;
                    ey -= 2.0 * (double)f_im / h;
                }
            } else {
                const double inv_r = rsqrt(d2 + eps2);
                const double inv_r2 = inv_r * inv_r;
                const double inv_r3 = inv_r2 * inv_r;
                const double inv_r5 = inv_r3 * inv_r2;
                const double inv_r7 = inv_r5 * inv_r2;
                const double q = tr.q[node];
                const double px = tr.p[2 * node], py = tr.p[2 * node + 1];
                const double sxx = tr.s[3 * node], sxy = tr.s[3 * node + 1], syy = tr.s[3 * node + 2];
                const double pr = px * rx + py * ry;
                const double srx = sxx * rx + sxy * ry;
                const double sry = sxy * rx + syy * ry;
                const double rsr = rx * srx + ry * sry;
                const double trace = sxx + syy;
                pot += q * inv_r + pr * inv_r3 + 0.5 * (3.0 * rsr * inv_r5 - trace * inv_r3);
                const double radial = q * inv_r3 + 3.0 * pr * inv_r5 + 7.5 * rsr * inv_r7 - 1.5 * trace * inv_r5;
                ex += radial * rx - px * inv_r3 - 3.0 * srx * inv_r5;
                ey += radial * ry - py * inv_r3 - 3.0 * sry * inv_r5;
            }
        }
        if (reject) {
            if (top + 4 > stack_size) {
                overflow = true;
                break;
            }
            __builtin_amdgcn_wave_barrier();
            if (lane < 4) {
                stn[top + lane] = nf.z + lane;
                stm[top + lane] = reject;
            }
            __builtin_amdgcn_wave_barrier();
            top += 4;
        }
    }
    if (overflow) {
        if (lane == 0) td3_report(status, -1);
        return;
    }
    if (COUNT) {
        u32 a = n_accept, t = n_terms;
        for (int off = TD3_WAVE / 2; off > 0; off >>= 1) {
            a += __shfl_xor(a, off, TD3_WAVE);
            t += __shfl_xor(t, off, TD3_WAVE);
        }
        if (lane == 0) {
            atomicAdd(&counts[0], (unsigned long long)a);
            atomicAdd(&counts[1], (unsigned long long)t);
            atomicAdd(&counts[2], (unsigned long long)w_accept);
            atomicAdd(&counts[3], (unsigned long long)w_sources);
        }
    }
    if (!active) return;
    if (bad) td3_report(status, ti + 1);
    phi[ti] = (KERNEL == TD3_LOG && !with_potential) ? 0.0 : pot;
    field[ti] = ex;
    field[m + ti] = ey;
}

// The reference's 16-bit-per-axis Morton key of the targets (flat_tree.py:891-912).
__global__ void morton_kernel(const double* __restrict__ t, i64 m, double lox, double loy,
                              double scale, u64* __restrict__ key) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    u64 k = 0;
    for (int axis = 0; axis < 2; ++axis) {
        const double v = (t[axis * m + i] - (axis ? loy : lox)) * scale;
        u64 c = (u64)v;
        c = (c | (c << 8)) & 0x00FF00FFull;
        c = (c | (c << 4)) & 0x0F0F0F0Full;
        c = (c | (c << 2)) & 0x33333333ull;
        c = (c | (c << 1)) & 0x55555555ull;
        k |= c << (axis == 0 ? 1 : 0);
    }
    key[i] = k;
}

__global__ void widen_kernel(const u32* __restrict__ v, i64 n, i64* __restrict__ out) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (i64)v[i];
}

template <int KERNEL, int PREC, int WAVES>
void launch_walk_w(hipStream_t s, const WalkTree& tr, const double* targets, i64 m,
                   const i64* order, const i64* exclude, const int* target_root, double theta,
                   double eps2, int with_potential, int stack_size, double* phi, double* field,
                   i64* status) {
    // a whole number of blocks per XCD (the extra blocks' targets are past m: they exit)
    const unsigned blocks = td3_xcd_grid(td3_blocks(m, 64 * WAVES));
    TD3_TIMED(s, "walk", hipLaunchKernelGGL((walk_kernel<KERNEL, PREC, WAVES>), dim3(blocks), dim3(64 * WAVES), 0, s, tr, targets,
                       m, order, exclude, target_root, theta, eps2, with_potential, stack_size, phi,
                       field, status));
}

// The walk's wavefronts per block (1, 2, 4 or 8): a launch shape chosen by the caller's
// tuning (td3_walk_waves), the same for every walk of the process.
int g_walk_waves = kWaves;

template <int KERNEL, int PREC>
void launch_walk(hipStream_t s, const WalkTree& tr, const double* targets, i64 m,
                 const i64* order, const i64* exclude, const int* target_root, double theta,
                 double eps2, int with_potential, int stack_size, double* phi, double* field,
                 i64* status) {
    switch (g_walk_waves) {
        case 1:
            launch_walk_w<KERNEL, PREC, 1>(s, tr, targets, m, order, exclude, target_root, theta,
                                           eps2, with_potential, stack_size, phi, field, status);
            break;
        case 2:
            launch_walk_w<KERNEL, PREC, 2>(s, tr, targets, m, order, exclude, target_root, theta,
                                           eps2, with_potential, stack_size, phi, field, status);
            break;
        case 8:
            launch_walk_w<KERNEL, PREC, 8>(s, tr, targets, m, order, exclude, target_root, theta,
                                           eps2, with_potential, stack_size, phi, field, status);
            break;
        default:
            launch_walk_w<KERNEL, PREC, 4>(s, tr, targets, m, order, exclude, target_root, theta,
                                           eps2, with_potential, stack_size, phi, field, status);
    }
}

}  // namespace

#include "td3_sort.h"

extern "C" size_t td3_target_order_workspace_bytes(int64_t m) {
    if (m <= 0) return 0;
    return (size_t)m * (2 * sizeof(u64) + 2 * sizeof(u32)) + td3sort::scratch_bytes(m) + 1024;
}

extern "C" int td3_target_order(const double* targets, int64_t m, double lox, double loy,
                                double scale, int64_t* order, void* workspace,
                                size_t workspace_bytes, void* stream) {
    TD3_ARG(m >= 0, "m");
    if (m == 0) return TD3_OK;
    TD3_ARG(workspace_bytes >= td3_target_order_workspace_bytes(m), "workspace too small");
    hipStream_t s = (hipStream_t)stream;
    char* w = (char*)workspace;
    u64* key = (u64*)w;
    u64* key_alt = key + m;
    u32* val = (u32*)(key_alt + m);
    u32* val_alt = val + m;
    u32* scratch = (u32*)(((uintptr_t)(val_alt + m) + 255) & ~(uintptr_t)255);
    hipLaunchKernelGGL(morton_kernel, dim3(td3_blocks(m, 256)), dim3(256), 0, s, targets, m, lox, loy,
                       scale, key);
    TD3_LAUNCHED();
    if (int rc = td3sort::sort_pairs(key, val, nullptr, key_alt, val_alt, nullptr, m, 0, 32, true,
                                     scratch, s))
        return rc;
    // widen the u32 order to int64
    hipLaunchKernelGGL(widen_kernel, dim3(td3_blocks(m, 256)), dim3(256), 0, s, val, m, order);
    TD3_LAUNCHED();
    return TD3_OK;
}

extern "C" int td3_walk(int kernel, int precision, const double* targets, int64_t m,
                        const int64_t* order, const int64_t* exclude, const double* sorted_pos,
                        const double* sorted_charge, const int64_t* perm, int64_t n,
                        const double* node_center, const double* node_half, const void* node_info,
                        const double* node_qabs, const double* node_q, const double* node_p,
                        const double* node_s, const double* moment_re, const double* moment_im,
                        int64_t order_p, int64_t tree_depth, const int32_t* target_root,
                        double theta, double softening, int with_potential, double* phi,
                        double* field, int64_t* status, void* stream) {
    TD3_ARG(kernel == TD3_COULOMB || kernel == TD3_LOG, "kernel");
    TD3_ARG(precision == TD3_FP64 || precision == TD3_MIXED, "precision");
    TD3_ARG(m >= 0 && n >= 1, "sizes");
    TD3_ARG(tree_depth >= 0 && tree_depth <= 64, "tree depth");
    if (m == 0) return TD3_OK;
    WalkTree tr;
    tr.spos = sorted_pos;
    tr.scharge = sorted_charge;
    tr.perm = perm;
    tr.n = n;
    tr.center = node_center;
    tr.half = node_half;
    tr.info = (const int4*)node_info;
    tr.qabs = node_qabs;
    tr.q = node_q;
    tr.p = node_p;
    tr.s = node_s;
    tr.mre = moment_re;
    tr.mim = moment_im;
    tr.p1 = (int)order_p + 1;
    // a forest's waves may start from several roots: give them the whole stack
    const int stack_size = target_root != nullptr ? kMaxStack
                           : (int)(3 * tree_depth + 8) < kMaxStack ? (int)(3 * tree_depth + 8)
                                                                   : kMaxStack;
    const double eps2 = softening * softening;
    hipStream_t s = (hipStream_t)stream;
    if (kernel == TD3_LOG) {
        if (precision == TD3_FP64)
            launch_walk<TD3_LOG, TD3_FP64>(s, tr, targets, m, order, exclude, target_root, theta, eps2,
                                           with_potential, stack_size, phi, field, status);
        else
            launch_walk<TD3_LOG, TD3_MIXED>(s, tr, targets, m, order, exclude, target_root, theta, eps2,
                                            with_potential, stack_size, phi, field, status);
    } else {
        if (precision == TD3_FP64)
            launch_walk<TD3_COULOMB, TD3_FP64>(s, tr, targets, m, order, exclude, target_root, theta, eps2,
                                               with_potential, stack_size, phi, field, status);
        else
            launch_walk<TD3_COULOMB, TD3_MIXED>(s, tr, targets, m, order, exclude, target_root, theta,
                                                eps2, with_potential, stack_size, phi, field, status);
    }
    TD3_LAUNCHED();
    return TD3_OK;
}

extern "C" int td3_walk_counts(int kernel, int precision, const double* targets, int64_t m,
                               const int64_t* order, const int64_t* exclude,
                               const double* sorted_pos, const double* sorted_charge,
                               const int64_t* perm, int64_t n, const double* node_center,
                               const double* node_half, const void* node_info,
                               const double* node_qabs, const double* node_q,
                               const double* node_p, const double* node_s,
                               const double* moment_re, const double* moment_im, int64_t order_p,
                               int64_t tree_depth, const int32_t* target_root, double theta,
                               double softening, int with_potential, double* phi, double* field,
                               uint64_t* counts, int64_t* status, void* stream) {
    TD3_ARG(kernel == TD3_COULOMB || kernel == TD3_LOG, "kernel");
    TD3_ARG(precision == TD3_FP64 || precision == TD3_MIXED, "precision");
    TD3_ARG(m >= 0 && n >= 1, "sizes");
    TD3_ARG(tree_depth >= 0 && tree_depth <= 64, "tree depth");
    TD3_ARG(counts != nullptr, "counts");
    if (m == 0) return TD3_OK;
    WalkTree tr;
    tr.spos = sorted_pos;
    tr.scharge = sorted_charge;
    tr.perm = perm;
    tr.n = n;
    tr.center = node_center;
    tr.half = node_half;
    tr.info = (const int4*)node_info;
    tr.qabs = node_qabs;
    tr.q = node_q;
    tr.p = node_p;
    tr.s = node_s;
    tr.mre = moment_re;
    tr.mim = moment_im;
    tr.p1 = (int)order_p + 1;
    const int stack_size = target_root != nullptr ? kMaxStack
                           : (int)(3 * tree_depth + 8) < kMaxStack ? (int)(3 * tree_depth + 8)
                                                                   : kMaxStack;
    const double eps2 = softening * softening;
    hipStream_t s = (hipStream_t)stream;
    const unsigned blocks = td3_xcd_grid(td3_blocks(m, 64 * kWaves));
    unsigned long long* c = (unsigned long long*)counts;
#define TD3_WALK_COUNTED(K, P)                                                                  \
    hipLaunchKernelGGL((walk_kernel<K, P, kWaves, true>), dim3(blocks), dim3(64 * kWaves), 0, s, \
                       tr, targets, m, order, exclude, target_root, theta, eps2, with_potential,  \
                       stack_size, phi, field, status, c)
    if (kernel == TD3_LOG) {
        if (precision == TD3_FP64)
            TD3_WALK_COUNTED(TD3_LOG, TD3_FP64);
        else
            TD3_WALK_COUNTED(TD3_LOG, TD3_MIXED);
    } else {
        if (precision == TD3_FP64)
            TD3_WALK_COUNTED(TD3_COULOMB, TD3_FP64);
        else
            TD3_WALK_COUNTED(TD3_COULOMB, TD3_MIXED);
    }
#undef TD3_WALK_COUNTED
    TD3_LAUNCHED();
    return TD3_OK;
}

extern "C" int td3_walk_waves(int waves) {
    TD3_ARG(waves == 0 || waves == 1 || waves == 2 || waves == 4 || waves == 8, "waves");
    const int previous = g_walk_waves;
    if (waves != 0) g_walk_waves = waves;
    return previous;
}
