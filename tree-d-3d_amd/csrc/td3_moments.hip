// Cell moments: P2M at the leaves, M2M up the tree, FP64.
//
// Replaces the reference's GPU moments (kernels/log_quadtree_moments/log_quadtree_moments_cuda.cu:7-46
// and kernels/quadtree_moments/quadtree_moments_cuda.cu:7-46: one thread per node summing
// all of its particles, the root thread looping over all N) with the upward pass of the
// reference's CPU kernels (kernels/log_moments_upward/log_moments_upward_pyccel.py:14-93,
// kernels/quadtree_moments_upward/quadtree_moments_upward_pyccel.py:13-83): every leaf is
// summed from its particles in one launch, then one launch per level shifts the four
// children into each parent.  Log moments use one thread per (node, coefficient n).

#include "td3_common.h"

namespace {

// (x + i y)^k by binary powering, times (pr + i pi)
__device__ __forceinline__ void cpow_times(double br, double bi, int k, double* pr, double* pi) {
    while (k) {
        if (k & 1) {
            const double tr = *pr * br - *pi * bi;
            *pi = *pr * bi + *pi * br;
            *pr = tr;
        }
        k >>= 1;
        if (k) {
            const double tr = br * br - bi * bi;
            bi = 2.0 * br * bi;
            br = tr;
        }
    }
}

// a_n = sum_j q_j (z_j - c)^n for every leaf; qabs = sum |q_j|.  With dipoles D_j
// (sdip (2, N), may be null) also a_n -= n D_j (z_j - c)^(n-1): the moments of the
// potential -2 Re(D / (z - z_j)) of a dipole, the derivative of a charge's in z_j.
__global__ void log_p2m_kernel(const double* __restrict__ spos, const double* __restrict__ scharge,
                               const double* __restrict__ sdip, i64 n,
                               const double* __restrict__ center,
                               const int4* __restrict__ info, i64 count, int p1,
                               double* __restrict__ qabs, double* __restrict__ mre,
                               double* __restrict__ mim) {
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count * p1) return;
    const i64 node = t / p1;
    const int k = (int)(t - node * p1);
    const int4 nf = info[node];
    if (nf.z >= 0) return;
    const double cx = center[2 * node], cy = center[2 * node + 1];
    double are = 0.0, aim = 0.0, qa = 0.0;
    for (int j = nf.x; j < nf.y; ++j) {
        const double q = scharge[j];
        const double dx = spos[j] - cx, dy = spos[n + j] - cy;
        double pr = q, pi = 0.0;
        cpow_times(dx, dy, k, &pr, &pi);
        are += pr;
        aim += pi;
        qa += fabs(q);
        if (sdip != nullptr && k > 0) {
            double dr = -k * sdip[j], di = -k * sdip[n + j];
            cpow_times(dx, dy, k - 1, &dr, &di);
            are += dr;
            aim += di;
        }
    }
    mre[node * p1 + k] = are;
    mim[node * p1 + k] = aim;
    if (k == 0) qabs[node] = qa;
}

// log_p2m_kernel for orders up to 15 with a thread per leaf: the powers (z_j - c)^n run
// upward by one multiplication each (log_moments_upward_pyccel.py:14-93 forms them the
// same way) and the p + 1 sums stay in registers, so a leaf's particles are read once and
// no lane raises z to a power of its own.  Each a_n sums its terms in particle order.
constexpr int kP2mMax = 16;
__global__ __launch_bounds__(256) void log_p2m_leaf_kernel(
    const double* __restrict__ spos, const double* __restrict__ scharge,
    const double* __restrict__ sdip, i64 n, const double* __restrict__ center,
    const int4* __restrict__ info, i64 count, int p1, double* __restrict__ qabs,
    double* __restrict__ mre, double* __restrict__ mim) {
    const i64 node = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= count) return;
    const int4 nf = info[node];
    if (nf.z >= 0) return;
    const double cx = center[2 * node], cy = center[2 * node + 1];
    double ar[kP2mMax], ai[kP2mMax];
#pragma unroll
    for (int k = 0; k < kP2mMax; ++k) ar[k] = ai[k] = 0.0;
    double qa = 0.0;
    for (int j = nf.x; j < nf.y; ++j) {
        const double q = scharge[j];
        const double dx = spos[j] - cx, dy = spos[n + j] - cy;
        qa += fabs(q);
        double pr = q, pi = 0.0;  // q z^k
        double er = 0.0, ei = 0.0;  // D z^(k-1), k >= 1
        if (sdip != nullptr) {
            er = sdip[j];
            ei = sdip[n + j];
        }
#pragma unroll
        for (int k = 0; k < kP2mMax; ++k) {
            if (k < p1) {
                ar[k] += pr;
                ai[k] += pi;
                if (sdip != nullptr && k > 0) {
                    ar[k] += -k * er;
                    ai[k] += -k * ei;
                    const double t = er * dx - ei * dy;
                    ei = er * dy + ei * dx;
                    er = t;
                }
                const double t = pr * dx - pi * dy;
                pi = pr * dy + pi * dx;
                pr = t;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < kP2mMax; ++k) {
        if (k < p1) {
            mre[node * p1 + k] = ar[k];
            mim[node * p1 + k] = ai[k];
        }
    }
    qabs[node] = qa;
}

// a_n += sum_m C(n, m) a_m(child) s^(n-m), s = c_child - c, for the nodes [lo, hi).
__global__ void log_m2m_kernel(const double* __restrict__ center, const int4* __restrict__ info,
                               i64 lo, i64 hi, int p1, const double* __restrict__ binomial,
                               int bsz, double* __restrict__ qabs, double* __restrict__ mre,
                               double* __restrict__ mim) {
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (hi - lo) * p1) return;
    const i64 node = lo + t / p1;
    const int k = (int)(t % p1);
    const int4 nf = info[node];
    if (nf.z < 0) return;
    const double cx = center[2 * node], cy = center[2 * node + 1];
    double are = 0.0, aim = 0.0, qa = 0.0;
    for (int c = nf.z; c < nf.z + 4; ++c) {
        qa += qabs[c];
        const int4 cf = info[c];
        if (cf.y == cf.x) continue;
        const double sr = center[2 * c] - cx, si = center[2 * c + 1] - cy;
        // sum over m = k .. 0 with s^(k-m) built upward
        double pr = 1.0, pi = 0.0, acc_r = 0.0, acc_i = 0.0;
        for (int m = k; m >= 0; --m) {
            const double b = binomial[k * bsz + m];
            const double ar = mre[(i64)c * p1 + m] * b, ai = mim[(i64)c * p1 + m] * b;
            acc_r += ar * pr - ai * pi;
            acc_i += ar * pi + ai * pr;
            const double tr = pr * sr - pi * si;
            pi = pr * si + pi * sr;
            pr = tr;
        }
        are += acc_r;
        aim += acc_i;
    }
    mre[node * p1 + k] = are;
    mim[node * p1 + k] = aim;
    if (k == 0) qabs[node] = qa;
}

// The same for p + 1 <= 16 coefficients, a wavefront per node: lane (c, k) = (lane / 16,
// lane % 16) loads child c's coefficient k once and takes the others from its neighbours
// (shuffles within the 16), forming child c's term of coefficient k; the four children go
// at once instead of one after another (a node's latency is one child's chain), and lane
// (0, k) adds them in child order, skipping empty children: the same terms and the same
// order of sums as log_m2m_kernel.
__global__ void log_m2m64_kernel(const double* __restrict__ center, const int4* __restrict__ info,
                                 i64 lo, i64 hi, int p1, const double* __restrict__ binomial,
                                 int bsz, double* __restrict__ qabs, double* __restrict__ mre,
                                 double* __restrict__ mim) {
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    const i64 node = lo + t / 64;
    const int lane = (int)(t % 64);
    const int ci = lane >> 4, k = lane & 15;
    if (node >= hi) return;  // uniform over the wavefront
    const int4 nf = info[node];
    if (nf.z < 0) return;
    const double cx = center[2 * node], cy = center[2 * node + 1];
    const int c = nf.z + ci;
    const double cqa = qabs[c];
    const int4 cf = info[c];
    const bool has = cf.y != cf.x;
    double acc_r = 0.0, acc_i = 0.0;
    const double sr = center[2 * c] - cx, si = center[2 * c + 1] - cy;
    const double my_re = has && k < p1 ? mre[(i64)c * p1 + k] : 0.0;
    const double my_im = has && k < p1 ? mim[(i64)c * p1 + k] : 0.0;
    double pr = 1.0, pi = 0.0;
    for (int j = 0; j < 16; ++j) {
        const int m = k - j;
        const double a_re = __shfl(my_re, m < 0 ? 0 : m, 16);
        const double a_im = __shfl(my_im, m < 0 ? 0 : m, 16);
        if (has && m >= 0 && k < p1) {
            const double b = binomial[k * bsz + m];
            const double ar = a_re * b, ai = a_im * b;
            acc_r += ar * pr - ai * pi;
            acc_i += ar * pi + ai * pr;
            const double tr = pr * sr - pi * si;
            pi = pr * si + pi * sr;
            pr = tr;
        }
    }
    double cr[4], cim[4], cq[4];
    int ch[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        cr[q] = __shfl(acc_r, 16 * q + k, 64);
        cim[q] = __shfl(acc_i, 16 * q + k, 64);
        cq[q] = __shfl(cqa, 16 * q + k, 64);
        ch[q] = __shfl(has ? 1 : 0, 16 * q + k, 64);
    }
    if (ci != 0) return;
    double are = 0.0, aim = 0.0, qa = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        qa += cq[q];
        if (!ch[q]) continue;
        are += cr[q];
        aim += cim[q];
    }
    if (k < p1) {
        mre[node * p1 + k] = are;
        mim[node * p1 + k] = aim;
    }
    if (k == 0) qabs[node] = qa;
}

// Coulomb moments (q, |q|, p, raw S) of one node: a leaf from its particles (leaves_only)
// or a parent from its children.
__device__ __forceinline__ void coulomb_node(const double* __restrict__ spos,
                                             const double* __restrict__ scharge, i64 n,
                                             const double* __restrict__ center,
                                             const int4* __restrict__ info, i64 node,
                                             int leaves_only, double* __restrict__ nq,
                                             double* __restrict__ nqabs, double* __restrict__ np,
                                             double* __restrict__ ns) {
    const int4 nf = info[node];
    const bool leaf = nf.z < 0;
    if (leaf != (leaves_only != 0)) return;
    const double xmid = center[2 * node], ymid = center[2 * node + 1];
    double q = 0, qa = 0, px = 0, py = 0, sxx = 0, sxy = 0, syy = 0;
    if (leaf) {
        for (int j = nf.x; j < nf.y; ++j) {
            const double c = scharge[j];
            const double dx = spos[j] - xmid, dy = spos[n + j] - ymid;
            q += c;
            qa += fabs(c);
            px += c * dx;
            py += c * dy;
            sxx += c * dx * dx;
            sxy += c * dx * dy;
            syy += c * dy * dy;
        }
    } else {
        for (int ch = nf.z; ch < nf.z + 4; ++ch) {
            const double dx = center[2 * ch] - xmid, dy = center[2 * ch + 1] - ymid;
            const double cq = nq[ch], cpx = np[2 * ch], cpy = np[2 * ch + 1];
            q += cq;
            qa += nqabs[ch];
            px += cpx + cq * dx;
            py += cpy + cq * dy;
            sxx += ns[3 * ch] + 2.0 * cpx * dx + cq * dx * dx;
            sxy += ns[3 * ch + 1] + cpx * dy + cpy * dx + cq * dx * dy;
            syy += ns[3 * ch + 2] + 2.0 * cpy * dy + cq * dy * dy;
        }
    }
    nq[node] = q;
    nqabs[node] = qa;
    np[2 * node] = px;
    np[2 * node + 1] = py;
    ns[3 * node] = sxx;
    ns[3 * node + 1] = sxy;
    ns[3 * node + 2] = syy;
}

// The nodes [lo, hi), a thread each.
__global__ void coulomb_kernel(const double* __restrict__ spos, const double* __restrict__ scharge,
                               i64 n, const double* __restrict__ center,
                               const int4* __restrict__ info, i64 lo, i64 hi, int leaves_only,
                               double* __restrict__ nq, double* __restrict__ nqabs,
                               double* __restrict__ np, double* __restrict__ ns) {
    const i64 node = lo + (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= hi) return;
    coulomb_node(spos, scharge, n, center, info, node, leaves_only, nq, nqabs, np, ns);
}

// The nodes [levels[d], levels[d + 1]) with the level starts read on the device (d < 0:
// every node, [0, levels[66])), grid-stride: a fixed grid whatever the level's width.
__global__ void coulomb_level_kernel(const double* __restrict__ spos,
                                     const double* __restrict__ scharge, i64 n,
                                     const double* __restrict__ center,
                                     const int4* __restrict__ info, const i64* __restrict__ levels,
                                     int d, int leaves_only, double* __restrict__ nq,
                                     double* __restrict__ nqabs, double* __restrict__ np,
                                     double* __restrict__ ns) {
    const i64 lo = d < 0 ? 0 : levels[d];
    const i64 hi = d < 0 ? levels[66] : levels[d + 1];
    const i64 stride = (i64)gridDim.x * blockDim.x;
    for (i64 node = lo + (i64)blockIdx.x * blockDim.x + threadIdx.x; node < hi; node += stride)
        coulomb_node(spos, scharge, n, center, info, node, leaves_only, nq, nqabs, np, ns);
}

__global__ void permute_kernel(const double* __restrict__ src, const i64* __restrict__ perm, i64 n,
                               int rows, i64 ld_src, double* __restrict__ dst) {
    const i64 k = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const i64 j = perm[k];
    for (int r = 0; r < rows; ++r) dst[r * n + k] = src[r * ld_src + j];
}

}  // namespace

extern "C" int td3_permute(const double* src, int64_t rows, int64_t ld_src, const int64_t* perm,
                           int64_t n, double* dst, void* stream) {
    TD3_ARG(n >= 0 && rows >= 1 && rows <= 8, "sizes");
    if (n == 0) return TD3_OK;
    hipLaunchKernelGGL(permute_kernel, dim3(td3_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream,
                       src, perm, n, (int)rows, ld_src, dst);
    TD3_LAUNCHED();
    return TD3_OK;
}

namespace {
int moments_impl(int kernel, int64_t order, const double* sorted_pos, const double* sorted_charge,
                 const double* sorted_dipole, int64_t n, const double* node_center,
                 const void* node_info, int64_t node_count, const int64_t* levels_host,
                 int64_t n_levels, const double* binomial, int64_t bsz, double* node_q,
                 double* node_qabs, double* node_p, double* node_s, double* moment_re,
                 double* moment_im, void* stream) {
    TD3_ARG(kernel == TD3_COULOMB || kernel == TD3_LOG, "kernel");
    TD3_ARG(n >= 1 && node_count >= 1 && n_levels >= 1, "sizes");
    TD3_ARG(levels_host != nullptr && levels_host[n_levels] == node_count, "levels");
    hipStream_t s = (hipStream_t)stream;
    const int4* info = (const int4*)node_info;
    if (kernel == TD3_LOG) {
        TD3_ARG(order >= 1 && order <= 64 && bsz >= order + 1, "order");
        const int p1 = (int)order + 1;
        if (p1 <= kP2mMax)
            TD3_TIMED(s, "moments_p2m", hipLaunchKernelGGL(log_p2m_leaf_kernel, dim3(td3_blocks(node_count, 256)), dim3(256), 0, s,
                               sorted_pos, sorted_charge, sorted_dipole, n, node_center, info,
                               node_count, p1, node_qabs, moment_re, moment_im));
        else
            TD3_TIMED(s, "moments_p2m", hipLaunchKernelGGL(log_p2m_kernel, dim3(td3_blocks(node_count * p1, 256)), dim3(256), 0, s,
                               sorted_pos, sorted_charge, sorted_dipole, n, node_center, info, node_count,
                               p1, node_qabs, moment_re, moment_im));
        TD3_LAUNCHED();
        for (i64 d = n_levels - 2; d >= 0; --d) {
            const i64 lo = levels_host[d], hi = levels_host[d + 1];
            if (p1 <= 16)
                TD3_TIMED(s, "moments_m2m", hipLaunchKernelGGL(log_m2m64_kernel, dim3(td3_blocks((hi - lo) * 64, 256)), dim3(256), 0, s,
                          node_center, info, lo, hi, p1, binomial, (int)bsz, node_qabs, moment_re,
                          moment_im));
            else
                TD3_TIMED(s, "moments_m2m", hipLaunchKernelGGL(log_m2m_kernel, dim3(td3_blocks((hi - lo) * p1, 256)), dim3(256), 0, s,
                          node_center, info, lo, hi, p1, binomial, (int)bsz, node_qabs, moment_re,
                          moment_im));
            TD3_LAUNCHED();
        }
        return TD3_OK;
    }
    TD3_TIMED(s, "moments_coulomb", hipLaunchKernelGGL(coulomb_kernel, dim3(td3_blocks(node_count, 256)), dim3(256), 0, s, sorted_pos,
                       sorted_charge, n, node_center, info, (i64)0, (i64)node_count, 1, node_q,
                       node_qabs, node_p, node_s));
    TD3_LAUNCHED();
    for (i64 d = n_levels - 2; d >= 0; --d) {
        const i64 lo = levels_host[d], hi = levels_host[d + 1];
        TD3_TIMED(s, "moments_coulomb", hipLaunchKernelGGL(coulomb_kernel, dim3(td3_blocks(hi - lo, 256)), dim3(256), 0, s, sorted_pos,
                           sorted_charge, n, node_center, info, lo, hi, 0, node_q, node_qabs, node_p,
                           node_s));
        TD3_LAUNCHED();
    }
    return TD3_OK;
}
}  // namespace

extern "C" int td3_moments(int kernel, int64_t order, const double* sorted_pos,
                           const double* sorted_charge, int64_t n, const double* node_center,
                           const void* node_info, int64_t node_count, const int64_t* levels_host,
                           int64_t n_levels, const double* binomial, int64_t bsz, double* node_q,
                           double* node_qabs, double* node_p, double* node_s, double* moment_re,
                           double* moment_im, void* stream) {
    return moments_impl(kernel, order, sorted_pos, sorted_charge, nullptr, n, node_center,
                        node_info, node_count, levels_host, n_levels, binomial, bsz, node_q,
                        node_qabs, node_p, node_s, moment_re, moment_im, stream);
}

extern "C" int td3_log_moments_dipole(int64_t order, const double* sorted_pos,
                                      const double* sorted_charge, const double* sorted_dipole,
                                      int64_t n, const double* node_center, const void* node_info,
                                      int64_t node_count, const int64_t* levels_host,
                                      int64_t n_levels, const double* binomial, int64_t bsz,
                                      double* node_qabs, double* moment_re, double* moment_im,
                                      void* stream) {
    TD3_ARG(sorted_dipole != nullptr, "sorted_dipole");
    return moments_impl(TD3_LOG, order, sorted_pos, sorted_charge, sorted_dipole, n, node_center,
                        node_info, node_count, levels_host, n_levels, binomial, bsz, nullptr,
                        node_qabs, nullptr, nullptr, moment_re, moment_im, stream);
}



extern "C" int td3_coulomb_moments_nosync(const double* sorted_pos, const double* sorted_charge,
                                          int64_t n, const double* node_center,
                                          const void* node_info, int64_t capacity,
                                          const int64_t* levels, int64_t max_depth,
                                          double* node_q, double* node_qabs, double* node_p,
                                          double* node_s, void* stream) {
    TD3_ARG(n >= 1 && capacity >= 1, "sizes");
    TD3_ARG(max_depth >= 1 && max_depth <= 64, "max_depth");
    TD3_ARG(levels != nullptr && node_info != nullptr, "levels, node_info");
    hipStream_t s = (hipStream_t)stream;
    const int4* info = (const int4*)node_info;
    const i64 cap_blocks = td3_blocks(capacity, 256);
    const unsigned grid = (unsigned)(cap_blocks < 2048 ? cap_blocks : 2048);
    TD3_TIMED(s, "moments_coulomb", hipLaunchKernelGGL(coulomb_level_kernel, dim3(grid), dim3(256), 0, s, sorted_pos, sorted_charge, n,
                       node_center, info, levels, -1, 1, node_q, node_qabs, node_p, node_s));
    TD3_LAUNCHED();
    // parents sit at depths max_depth - 1 .. 0; levels past the tree's depth are empty
    for (int d = (int)max_depth - 1; d >= 0; --d) {
        TD3_TIMED(s, "moments_coulomb", hipLaunchKernelGGL(coulomb_level_kernel, dim3(grid), dim3(256), 0, s, sorted_pos, sorted_charge,
                           n, node_center, info, levels, d, 0, node_q, node_qabs, node_p, node_s));
        TD3_LAUNCHED();
    }
    return TD3_OK;
}
