// FMM translations and evaluation for the log (line-charge) kernel.
//
// Replaces kernels/fmm_m2l/fmm_m2l_cuda.cu:12-80 (one thread per pair, 2 x 130 local
// doubles, 2(p+1) FP64 atomics per direction), kernels/fmm_downward/fmm_downward_cuda.cu:10-51
// and kernels/fmm_evaluate/fmm_evaluate_cuda.cu:11-87 (6 FP64 atomics per near pair).
// Everything here is owner-computes and deterministic: no float atomics.
//
// Expansions are stored radius-scaled, with r = sqrt(2) h the half-diagonal of a cell:
//   moments  A_k = a_k / r^k,   locals  B_l = b_l r^l,
// so every coefficient of every operator is bounded (|u| + |v| < theta below).  This is
// what lets the deep levels translate in FP32; the FP64 path uses the same code in FP64.
//
// M2L (the reference's factorisation, kernels/fmm_m2l_host/fmm_m2l_host_pyccel.py:16-105):
// for a directed pair, t = c_dst - c_src, u = r_src / t, v = r_dst / t,
//   S_l = sum_k coeff[k, l] A_k u^k,   B_0 += A_0 ln t - S_0,   B_l += (-1)^(l+1) v^l S_l,
// coeff the constant real matrix of flat_tree.py:234-245.  One lane per directed pair over
// the destination-sorted pair list (CSR); lanes of one destination are summed by a
// segmented wave reduction, and destinations whose rows cross a wave boundary are
// finished by a fixed-order pass over per-wave partials.

#include "td3_common.h"

namespace {

template <typename T>
struct Cx {
    T re, im;
};

template <typename T>
__device__ __forceinline__ Cx<T> cmul(Cx<T> a, Cx<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

__device__ __forceinline__ double shfl_up_d(double v, int off) { return __shfl_up(v, off, 64); }
__device__ __forceinline__ float shfl_up_d(float v, int off) { return __shfl_up(v, off, 64); }

// ------------------------------------------------------------------------------ M2L
template <int P, typename T>
__device__ __forceinline__ void m2l_chunk(
    const int* __restrict__ col, const int* __restrict__ dst, const u32* __restrict__ row_start,
    u32 w0, u32 e_hi, u32 chunk, const double* __restrict__ dcenter,
    const double* __restrict__ dhalf, const double* __restrict__ scenter,
    const double* __restrict__ shalf, const T* __restrict__ mre, const T* __restrict__ mim,
    int p1, const T* s_coeff, double* __restrict__ lre, double* __restrict__ lim,
    T* __restrict__ side) {
    const int lane = td3_lane();
    const u32 e = w0 + lane;
    const bool valid = e < e_hi;
    const int d = valid ? dst[e] : -1 - lane;
    Cx<T> b[P + 1];
#pragma unroll
    for (int l = 0; l <= P; ++l) b[l] = {(T)0, (T)0};
    if (valid) {
        const int src = col[e];
        const double tr = dcenter[2 * d] - scenter[2 * src];
        const double ti = dcenter[2 * d + 1] - scenter[2 * src + 1];
        const double t2 = tr * tr + ti * ti;
        const double wr = tr / t2, wi = -ti / t2;
        const double rs = 1.4142135623730951 * shalf[src];
        const double rt = 1.4142135623730951 * dhalf[d];
        const Cx<T> u = {(T)(rs * wr), (T)(rs * wi)};
        const Cx<T> v = {(T)(rt * wr), (T)(rt * wi)};
        // A_k u^k
        Cx<T> a[P + 1];
        Cx<T> pw = {(T)1, (T)0};
#pragma unroll
        for (int k = 0; k <= P; ++k) {
            if (k < p1) {
                const Cx<T> ak = {mre[(i64)src * p1 + k], mim[(i64)src * p1 + k]};
                a[k] = cmul(ak, pw);
            } else {
                a[k] = {(T)0, (T)0};
            }
            pw = cmul(pw, u);
        }
        // S_l = sum_k coeff[k, l] a_k
#pragma unroll
        for (int k = 0; k <= P; ++k) {
#pragma unroll
            for (int l = 0; l <= P; ++l) {
                const T c = s_coeff[k * (P + 1) + l];
                b[l].re += c * a[k].re;
                b[l].im += c * a[k].im;
            }
        }
        // B_0 = A_0 ln t - S_0 ; B_l = (-1)^(l+1) v^l S_l
        T lnr, lni;
        if (sizeof(T) == 8) {
            lnr = (T)(0.5 * log(t2));
            lni = (T)atan2(ti, tr);
        } else {
            lnr = (T)(0.5f * __logf((float)t2));
            lni = (T)atan2f((float)ti, (float)tr);
        }
        b[0].re = a[0].re * lnr - a[0].im * lni - b[0].re;
        b[0].im = a[0].re * lni + a[0].im * lnr - b[0].im;
        Cx<T> vp = v;
        T sign = (T)1;
#pragma unroll
        for (int l = 1; l <= P; ++l) {
            const Cx<T> r = cmul(b[l], vp);
            b[l] = {sign * r.re, sign * r.im};
            sign = -sign;
            vp = cmul(vp, v);
        }
    }
    // segmented inclusive scan over lanes of equal destination
    const int prev_d = __shfl_up(d, 1, 64);
    const bool head = lane == 0 || prev_d != d;
    const u64 heads = __ballot(head);
    const u64 le = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    const int seg_start = 63 - __clzll(heads & le);
#pragma unroll
    for (int l = 0; l <= P; ++l) {
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const T r = shfl_up_d(b[l].re, off);
            const T i = shfl_up_d(b[l].im, off);
            if (lane - off >= seg_start) {
                b[l].re += r;
                b[l].im += i;
            }
        }
    }
    const bool tail = lane == 63 || ((heads >> (lane + 1)) & 1ull);
    if (!valid || !tail) return;
    const u32 rs0 = row_start[d], rs1 = row_start[d + 1];
    const bool before = rs0 < w0;
    const bool after = rs1 > w0 + 64;
    // (loops over the register array with a compile-time bound keep it in registers)
    if (!before && !after) {
#pragma unroll
        for (int l = 0; l <= P; ++l) {
            if (l < p1) {
                lre[(i64)d * p1 + l] = (double)b[l].re;
                lim[(i64)d * p1 + l] = (double)b[l].im;
            }
        }
    } else {
        const i64 slot = ((i64)chunk * 2 + (before ? 0 : 1)) * 2 * p1;
#pragma unroll
        for (int l = 0; l <= P; ++l) {
            if (l < p1) {
                side[slot + 2 * l] = b[l].re;
                side[slot + 2 * l + 1] = b[l].im;
            }
        }
    }
}

// One lane per pair, 64 consecutive pairs (a chunk) per wavefront, over the rows
// [node_lo, node_hi) (their bounds read on the device; the grid covers at most max_chunks
// chunks, and wavefronts past the range exit).  A chunk's rows that end inside it are
// written directly; the partial sums of rows that cross a chunk boundary go to side (two
// slots per chunk) for m2l_fixup_kernel.
template <int P, typename T>
__global__ __launch_bounds__(256) void m2l_kernel(
    const int* __restrict__ col, const int* __restrict__ dst, const u32* __restrict__ row_start,
    i64 node_lo, i64 node_hi, u32 max_chunks, const double* __restrict__ dcenter,
    const double* __restrict__ dhalf, const double* __restrict__ scenter,
    const double* __restrict__ shalf, const T* __restrict__ mre, const T* __restrict__ mim,
    int p1, const double* __restrict__ coeff, double* __restrict__ lre, double* __restrict__ lim,
    T* __restrict__ side) {
    const u32 e_lo = row_start[node_lo], e_hi = row_start[node_hi];
    if (e_lo + (u64)blockIdx.x * (blockDim.x / 64) * 64 >= e_hi) return;  // the whole block
    __shared__ T s_coeff[(P + 1) * (P + 1)];
    for (int i = threadIdx.x; i < (P + 1) * (P + 1); i += blockDim.x) {
        const int k = i / (P + 1), l = i % (P + 1);
        s_coeff[i] = (k < p1 && l < p1) ? (T)coeff[k * p1 + l] : (T)0;
    }
    __syncthreads();
    const u32 chunk = (u32)((blockIdx.x * blockDim.x + threadIdx.x) / 64);
    if (chunk >= max_chunks || e_lo + (u64)chunk * 64 >= e_hi) return;
    m2l_chunk<P, T>(col, dst, row_start, e_lo + chunk * 64, e_hi, chunk, dcenter, dhalf, scenter,
                    shalf, mre, mim, p1, s_coeff, lre, lim, side);
}

// ------------------------------------------------------------------------------ M2L on MFMA
// The FP64 translation with S = coeff^T X as a matrix product on the matrix cores
// (v_mfma_f64_16x16x4f64), for orders up to 15 (p + 1 <= 16 coefficients, zero padded):
// the top levels of the mixed-precision FMM.  One wavefront per destination row; the row's
// pairs go 16 at a time.  Lane (j, q) = (lane % 16, lane / 16) takes pair j of the tile: it
// forms X_k = A_k u^k for k = q, q + 4, q + 8, q + 12 (the B operand of the four K = 4
// products) and holds coeff[4 kb + q, j] (the A operand); the result D[l][j] = S_l of pair j
// comes back in lane (j, l % 4), element l / 4 (v_mfma_f64_16x16x4f64 gives row q + 4 r).
// The lane scales it by (-1)^(l+1) v^l (l >= 1) or forms A_0 ln t - S_0 (l = 0) and adds it
// to its running sum; the 16 lanes of each q are summed at the end in a fixed butterfly, so
// the result is deterministic.
typedef float td3_f32x4 __attribute__((ext_vector_type(4)));
typedef double td3_f64x4 __attribute__((ext_vector_type(4)));

// The 16x16x4 product D += A B of one wavefront (A one element per lane, B one element
// per lane, D four per lane), in FP32 or FP64.
__device__ __forceinline__ td3_f32x4 mfma16(float a, float b, td3_f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ td3_f64x4 mfma16(double a, double b, td3_f64x4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// u^q for q in 0..3, selected without divergence
__device__ __forceinline__ Cx<double> cpow_small(Cx<double> u, Cx<double> u2, Cx<double> u3,
                                                 int q) {
    Cx<double> r = {1.0, 0.0};
    r = q == 1 ? u : r;
    r = q == 2 ? u2 : r;
    r = q == 3 ? u3 : r;
    return r;
}

__global__ __launch_bounds__(256) void m2l_mfma_kernel(
    const int* __restrict__ col, const u32* __restrict__ row_start, i64 node_lo, i64 node_hi,
    const double* __restrict__ dcenter, const double* __restrict__ dhalf,
    const double* __restrict__ scenter, const double* __restrict__ shalf,
    const double* __restrict__ mre, const double* __restrict__ mim, int p1,
    const double* __restrict__ coeff, double* __restrict__ lre, double* __restrict__ lim) {
    const int lane = td3_lane();
    const i64 d = node_lo + ((i64)blockIdx.x * blockDim.x + threadIdx.x) / 64;
    if (d >= node_hi) return;  // uniform over the wavefront
    const u32 r0 = row_start[d], r1 = row_start[d + 1];
    const int j = lane & 15, q = lane >> 4;
    // element r of lane (j, q) of the product holds D[q + 4 r][j]
    auto row_of = [q](int r) { return q + 4 * r; };
    if (r0 == r1) {
        // a row without pairs is written too (zeros), so the caller need not clear it
        if (j == 0)
            for (int r = 0; r < 4; ++r)
                if (row_of(r) < p1) {
                    lre[d * p1 + row_of(r)] = 0.0;
                    lim[d * p1 + row_of(r)] = 0.0;
                }
        return;
    }
    double ca[4];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
        const int k = 4 * kb + q;
        ca[kb] = (k < p1 && j < p1) ? coeff[k * p1 + j] : 0.0;
    }
    const double cx = dcenter[2 * d], cy = dcenter[2 * d + 1];
    const double rt = 1.4142135623730951 * dhalf[d];
    double acc_re[4] = {0.0, 0.0, 0.0, 0.0}, acc_im[4] = {0.0, 0.0, 0.0, 0.0};
    for (u32 e0 = r0; e0 < r1; e0 += 16) {
        const u32 e = e0 + (u32)j;
        const bool valid = e < r1;
        double x_re[4] = {0.0, 0.0, 0.0, 0.0}, x_im[4] = {0.0, 0.0, 0.0, 0.0};
        Cx<double> v = {0.0, 0.0};
        double lnr = 0.0, lni = 0.0;
        if (valid) {
            const u32 src = (u32)col[e];
            const double tr = cx - scenter[2 * src], ti = cy - scenter[2 * src + 1];
            const double t2 = tr * tr + ti * ti;
            const double inv = 1.0 / t2;
            const double wr = tr * inv, wi = -ti * inv;
            const double rs = 1.4142135623730951 * shalf[src];
            const Cx<double> u = {rs * wr, rs * wi};
            v = {rt * wr, rt * wi};
            const Cx<double> u2 = cmul(u, u);
            const Cx<double> u3 = cmul(u2, u);
            const Cx<double> u4 = cmul(u2, u2);
            Cx<double> up = cpow_small(u, u2, u3, q);
            // rows of this lane's k = q + 4 kb: one address, immediate offsets
            const double* ar = mre + (src * (u32)p1 + (u32)q);
            const double* ai = mim + (src * (u32)p1 + (u32)q);
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) {
                const int k = 4 * kb + q;
                if (k < p1) {
                    const Cx<double> x = cmul(Cx<double>{ar[4 * kb], ai[4 * kb]}, up);
                    x_re[kb] = x.re;
                    x_im[kb] = x.im;
                }
                up = cmul(up, u4);
            }
            if (q == 0) {
                lnr = 0.5 * log(t2);
                lni = atan2(ti, tr);
            }
        }
        td3_f64x4 dre = {0.0, 0.0, 0.0, 0.0}, dim = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) {
            dre = mfma16(ca[kb], x_re[kb], dre);
            dim = mfma16(ca[kb], x_im[kb], dim);
        }
        if (valid) {
            // v^l for the first element's row l = q, then the step v^4 to the next
            const Cx<double> v2 = cmul(v, v);
            Cx<double> vp = cpow_small(v, v2, cmul(v2, v), q);
            const Cx<double> step = cmul(v2, v2);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int l = row_of(r);
                const Cx<double> sl = {dre[r], dim[r]};
                if (l == 0) {
                    acc_re[0] += x_re[0] * lnr - x_im[0] * lni - sl.re;
                    acc_im[0] += x_re[0] * lni + x_im[0] * lnr - sl.im;
                } else if (l < p1) {
                    const Cx<double> b = cmul(sl, vp);
                    const double sign = (l & 1) ? 1.0 : -1.0;
                    acc_re[r] += sign * b.re;
                    acc_im[r] += sign * b.im;
                }
                vp = cmul(vp, step);
            }
        }
    }
    // sum the 16 pair lanes of each q
#pragma unroll
    for (int r = 0; r < 4; ++r) {
#pragma unroll
        for (int off = 8; off > 0; off >>= 1) {
            acc_re[r] += __shfl_xor(acc_re[r], off, 64);
            acc_im[r] += __shfl_xor(acc_im[r], off, 64);
        }
    }
    if (j == 0) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int l = row_of(r);
            if (l < p1) {
                lre[d * p1 + l] = acc_re[r];
                lim[d * p1 + l] = acc_im[r];
            }
        }
    }
}

// Destinations whose row crosses a wave boundary: last-segment partial of the first wave
// plus the first-segment partials of the following waves, in order.
template <typename T>
__global__ void m2l_fixup_kernel(const u32* __restrict__ row_start, i64 node_lo, i64 node_hi,
                                 int p1, const T* __restrict__ side, double* __restrict__ lre,
                                 double* __restrict__ lim) {
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    const i64 node = node_lo + t / (2 * p1);
    const int j = (int)(t % (2 * p1));
    if (node >= node_hi) return;
    const u32 e_lo = row_start[node_lo];
    const u32 rs0 = row_start[node], rs1 = row_start[node + 1];
    const int l = j / 2;
    if (rs1 == rs0) {
        // a row without pairs is written too (zeros), so the caller need not clear it
        (j & 1 ? lim : lre)[node * p1 + l] = 0.0;
        return;
    }
    const u32 wa = (rs0 - e_lo) / 64, wb = (rs1 - 1 - e_lo) / 64;
    if (wa == wb) return;
    double v = (double)side[((i64)wa * 2 + 1) * 2 * p1 + j];
    for (u32 w = wa + 1; w <= wb; ++w) v += (double)side[((i64)w * 2) * 2 * p1 + j];
    if (j & 1)
        lim[node * p1 + l] = v;
    else
        lre[node * p1 + l] = v;
}

// ------------------------------------------------------------------------------ FP32 M2L
// Source records (td3_m2l_records), one per cell, m2l_record_floats(p1) floats (a multiple
// of 32 bytes):
//   floats [0, 4)            the centre x, y as FP64
//   floats [4, 4 + 2 p1)     A_k = a_k / r^k (re, im) as FP32, r = sqrt(2) h, except that
//                            the slot of Im A_0 holds r (a_0 is a net charge: real)
//   the rest                 zeros
// so the four coefficients A_{4q} .. A_{4q+3} a lane needs are the two 16-byte words at
// floats 4 + 8 q and 8 + 8 q: inside the record whenever 4 q < p1 (4 q + 2 < p1), zeros
// past the order.
__host__ __device__ __forceinline__ int m2l_record_floats(int p1) {
    return (16 + 8 * p1 + 31) / 32 * 8;
}

// One thread per 16-byte word of a record: scaled as scale_moments_kernel scales.
__global__ void m2l_records_kernel(const double* __restrict__ mre, const double* __restrict__ mim,
                                   const double* __restrict__ center,
                                   const double* __restrict__ half, i64 count, int p1, int rf,
                                   float* __restrict__ rec) {
    const int words = rf / 4;
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count * words) return;
    const i64 node = t / words;
    const int w = (int)(t % words);
    float* out = rec + node * rf + 4 * w;
    if (w == 0) {
        *reinterpret_cast<double2*>(out) = make_double2(center[2 * node], center[2 * node + 1]);
        return;
    }
    const double r = 1.4142135623730951 * half[node];
    const double inv = 1.0 / r;
    float v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int f = 4 * w + i - 4;  // float index in the coefficient area
        const int k = f >> 1;
        float x = 0.0f;
        if (f == 1) {
            x = (float)r;
        } else if (k < p1) {
            double s = 1.0;
            for (int m = 0; m < k; ++m) s *= inv;
            x = (float)(((f & 1) ? mim : mre)[node * p1 + k] * s);
        }
        v[i] = x;
    }
    *reinterpret_cast<float4*>(out) = make_float4(v[0], v[1], v[2], v[3]);
}

// FP32 M2L on the matrix cores from source records, for orders up to 15.  A wavefront sums
// kM2lRows consecutive destination rows, whose pairs are one contiguous run of the rows'
// columns: the run goes through the matrix product 16 pairs at a time regardless of where
// rows begin and end, so only the run's last tile is padded.
//
// Lane (j, q) = (lane % 16, lane / 16) takes pair j of the tile and the coefficients
// k = 4 q + kb (kb = 0..3): it forms X_k = A_k u^k (u = r_src / t, t = c_dst - c_src) as the
// B operand of the four K = 4 products and holds coeff'[4 q + kb, j] as the A operand, where
// coeff' is the reference's coeff (flat_tree.py:234-245) with the signs of the translation
// folded in (-1 for l = 0, (-1)^(l+1) otherwise).  The product S'_l of pair j comes back in
// lane (j, q), element r, for l = 4 q + r: the same four exponents as its inputs, so the
// lane forms u^{4q+r} and v^{4q+r} (v = r_dst / t) alike, from the fourth powers.  The lane
// adds v^l S'_l, and A_0 ln|t| for l = 0 (the imaginary part of
// A_0 ln t reaches neither the potential (-2 Re W), nor the field (W'), nor any other
// coefficient through L2L), to its running sums; when a row ends inside a tile, the lanes
// whose pair belongs to it add theirs, the 16 lanes of each q are summed in a fixed order
// (m2l_row_sums) and the row is written, so every row is summed in one fixed order.
constexpr int kM2lRows = 8;

// The eight running sums of a lane (S_l re, im for its four l) summed over each row of 16
// lanes by recursive halving over DPP mirrors: lane j ends with value m(j) = 4 [j >= 8] +
// 2 [j % 8 >= 4] + [j % 4 >= 2] summed over the row's 16 lanes (lanes j and j ^ 1 alike).
__device__ __forceinline__ float m2l_row_sums(const float (&f)[8], int j) {
    const bool h1 = j >= 8, h2 = (j & 7) >= 4, h3 = (j & 2) != 0;
    float w[4], x[2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
        w[i] = (h1 ? f[4 + i] : f[i]) + td3_dpp_f<0x140>(h1 ? f[i] : f[4 + i]);
#pragma unroll
    for (int i = 0; i < 2; ++i)
        x[i] = (h2 ? w[2 + i] : w[i]) + td3_dpp_f<0x141>(h2 ? w[i] : w[2 + i]);
    const float y = (h3 ? x[1] : x[0]) + td3_dpp_f<0x4e>(h3 ? x[0] : x[1]);
    return y + td3_dpp_f<0xb1>(y);
}

__global__ __launch_bounds__(256) void m2l_f32_kernel(
    const int* __restrict__ col, const u32* __restrict__ row_start, i64 node_lo, i64 node_hi,
    const double* __restrict__ dcenter, const double* __restrict__ dhalf,
    const float* __restrict__ rec, int rf, int p1, const double* __restrict__ coeff,
    double* __restrict__ lre, double* __restrict__ lim) {
    const int lane = td3_lane();
    const i64 d0 = node_lo + (((i64)blockIdx.x * blockDim.x + threadIdx.x) >> 6) * kM2lRows;
    if (d0 >= node_hi) return;  // uniform over the wavefront
    const int nrows = (int)(node_hi - d0 < kM2lRows ? node_hi - d0 : kM2lRows);
    // lane i <= nrows holds the start of row d0 + i (row nrows - 1 ends at lane nrows's)
    const int lane_rs = lane <= nrows ? (int)row_start[d0 + lane] : 0;
    const int j = lane & 15, q = lane >> 4;
    float ca[4];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
        const int k = 4 * q + kb;
        const float sign = j == 0 ? -1.0f : ((j & 1) ? 1.0f : -1.0f);
        ca[kb] = (k < p1 && j < p1) ? sign * (float)coeff[k * p1 + j] : 0.0f;
    }
    // the value this lane writes after a row's sums: l = 4 q + m / 2, re or im
    const int m_out = 4 * (j >= 8) + 2 * ((j & 7) >= 4) + ((j & 2) != 0);
    const int l_out = 4 * q + (m_out >> 1);
    double* const out = (m_out & 1) ? lim : lre;
    const bool writer = (j & 1) == 0 && l_out < p1;
    auto row_at = [&](int i) { return (u32)__builtin_amdgcn_readlane(lane_rs, i); };
    const u32 e_end = row_at(nrows);
    float acc[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    int r = 0;                 // the first row not yet written (uniform)
    u32 r_hi = row_at(1);      // its end
    // A tile's pair for this lane (lanes past the run take its last pair, and add nothing).
    // Each tile works on the source record loaded during the tile before, and loads the
    // column index of the tile after next.
    auto pair_at = [&](u32 base) { return base + (u32)j < e_end ? base + (u32)j : e_end - 1; };
    struct SourceRecord {
        double2 g;
        float rs;
        float4 a1, a2;
    };
    auto load_record = [&](int src) {
        const float* rp = rec + (i64)src * rf;
        SourceRecord s;
        s.g = *reinterpret_cast<const double2*>(rp);
        s.rs = rp[5];
        s.a1 = s.a2 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (4 * q < p1) s.a1 = *reinterpret_cast<const float4*>(rp + 4 + 8 * q);
        if (4 * q + 2 < p1) s.a2 = *reinterpret_cast<const float4*>(rp + 8 + 8 * q);
        return s;
    };
    u32 e0 = row_at(0);
    SourceRecord cur = {};
    int src_next = 0;
    if (e0 < e_end) {
        cur = load_record(col[pair_at(e0)]);
        src_next = col[pair_at(e0 + 16)];
    }
    for (; e0 < e_end; e0 += 16) {
        const bool valid = e0 + (u32)j < e_end;
        const u32 e = pair_at(e0);
        // the row of this lane's pair: r plus the later rows that start at or before e
        int rl = r;
        if (r_hi < e0 + 16)
            for (int t = r + 1; t < nrows && row_at(t) < e0 + 16; ++t) rl += e >= row_at(t) ? 1 : 0;
        const i64 d = d0 + rl;
        const double dcx = dcenter[2 * d], dcy = dcenter[2 * d + 1], dh = dhalf[d];
        const SourceRecord next = load_record(src_next);
        src_next = col[pair_at(e0 + 32)];
        const double2 g = cur.g;
        const float rs = cur.rs;
        float4 a1 = cur.a1;
        const float4 a2 = cur.a2;
        cur = next;
        if (q == 0) a1.y = 0.0f;  // the slot of Im A_0 holds r
        // t = c_dst - c_src and w = 1 / t in FP64 (on gfx950 at the rate of FP32 without
        // packing), u = r_src w and v = r_dst w rounded once to FP32
        const double tr = dcx - g.x, ti = dcy - g.y;
        const double t2 = tr * tr + ti * ti;
        const double inv = __builtin_amdgcn_rcp(t2);
        const double wr = tr * inv, wi = -ti * inv;
        const Cx<float> u = {(float)(rs * wr), (float)(rs * wi)};
        const double rt = 1.4142135623730951 * dh;
        const float rtf = (float)rt;
        const Cx<float> u2 = cmul(u, u);
        const Cx<float> u4 = cmul(u2, u2);
        const Cx<float> u8 = cmul(u4, u4);
        Cx<float> up = {valid ? 1.0f : 0.0f, 0.0f};
        up = q == 1 && valid ? u4 : up;
        up = q == 2 && valid ? u8 : up;
        up = q == 3 && valid ? cmul(u8, u4) : up;
        Cx<float> pw[4];
        pw[0] = up;
        pw[1] = cmul(pw[0], u);
        pw[2] = cmul(pw[1], u);
        pw[3] = cmul(pw[2], u);
        const Cx<float> a[4] = {{a1.x, a1.y}, {a1.z, a1.w}, {a2.x, a2.y}, {a2.z, a2.w}};
        float x_re[4], x_im[4];
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) {
            const Cx<float> x = cmul(a[kb], pw[kb]);
            x_re[kb] = x.re;
            x_im[kb] = x.im;
        }
        // A_0 ln|t| (v_log_f32 is log2)
        const float a0_ln = q == 0 && valid
                                ? a1.x * (0.34657359027997264f * __builtin_amdgcn_logf((float)t2))
                                : 0.0f;
        td3_f32x4 dre = {0.0f, 0.0f, 0.0f, 0.0f}, dim = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) {
            dre = mfma16(ca[kb], x_re[kb], dre);
            dim = mfma16(ca[kb], x_im[kb], dim);
        }
        // v^l S'_l.  Cells of one tree have radii r_root 2^-depth, so r_dst / r_src is a
        // power of two (equal FP32 significands): v = 2^delta u exactly and v^l S'_l is
        // u^l S'_l scaled by 2^(l delta).  Otherwise (cells of two trees) v's own powers.
        float c[8];
        const int sig_mask = 0x7fffff;
        const bool dyadic = (__float_as_int(rtf) & sig_mask) == (__float_as_int(rs) & sig_mask);
        if (__ballot(!dyadic) == 0) {
            const int delta = ((__float_as_int(rtf) >> 23) & 0xff) - ((__float_as_int(rs) >> 23) & 0xff);
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) {
                const int sh = (4 * q + rr) * delta;
                c[2 * rr] = ldexpf(dre[rr] * pw[rr].re - dim[rr] * pw[rr].im, sh);
                c[2 * rr + 1] = ldexpf(dre[rr] * pw[rr].im + dim[rr] * pw[rr].re, sh);
            }
        } else {
            const Cx<float> v = {(float)(rt * wr), (float)(rt * wi)};
            const Cx<float> v2 = cmul(v, v);
            const Cx<float> v4 = cmul(v2, v2);
            const Cx<float> v8 = cmul(v4, v4);
            Cx<float> vp = {1.0f, 0.0f};
            vp = q == 1 ? v4 : vp;
            vp = q == 2 ? v8 : vp;
            vp = q == 3 ? cmul(v8, v4) : vp;
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) {
                c[2 * rr] = dre[rr] * vp.re - dim[rr] * vp.im;
                c[2 * rr + 1] = dre[rr] * vp.im + dim[rr] * vp.re;
                vp = cmul(vp, v);
            }
        }
        c[0] += a0_ln;
        // rows that end inside this tile: their lanes' terms in, summed, written
        while (r < nrows && r_hi <= e0 + 16) {
            const bool mine = rl == r;
            float f[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const float t = mine ? c[i] : 0.0f;
                f[i] = acc[i] + t;
                c[i] -= t;
                acc[i] = 0.0f;
            }
            const float sum = m2l_row_sums(f, j);
            if (writer) out[(d0 + r) * p1 + l_out] = (double)sum;
            ++r;
            r_hi = r < nrows ? row_at(r + 1) : e_end;
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) acc[i] += c[i];
    }
    // rows past the last pair (or every row, without pairs) are empty: zeros
    for (; r < nrows; ++r)
        if (writer) out[(d0 + r) * p1 + l_out] = 0.0;
}

// ------------------------------------------------------------------------------ scaling
// A_k = a_k / r^k (r = sqrt(2) h), as T.
template <typename T>
__global__ void scale_moments_kernel(const double* __restrict__ mre, const double* __restrict__ mim,
                                     const double* __restrict__ half, i64 count, int p1,
                                     T* __restrict__ ore, T* __restrict__ oim) {
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count * p1) return;
    const i64 node = t / p1;
    const int k = (int)(t % p1);
    const double inv = 1.0 / (1.4142135623730951 * half[node]);
    double s = 1.0;
    for (int i = 0; i < k; ++i) s *= inv;
    ore[t] = (T)(mre[t] * s);
    oim[t] = (T)(mim[t] * s);
}

// b_l = B_l / r^l (the reference's unscaled locals), in place.
__global__ void unscale_locals_kernel(double* __restrict__ lre, double* __restrict__ lim,
                                      const double* __restrict__ half, i64 count, int p1) {
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count * p1) return;
    const i64 node = t / p1;
    const int l = (int)(t % p1);
    const double inv = 1.0 / (1.4142135623730951 * half[node]);
    double s = 1.0;
    for (int i = 0; i < l; ++i) s *= inv;
    lre[t] *= s;
    lim[t] *= s;
}

// ------------------------------------------------------------------------------ L2L
// kernels/fmm_downward/fmm_downward_pyccel.py:15-64 in scaled form: with s = c_child - c_parent,
// B^child_m += sum_{l >= m} C(l, m) B_l (r_c / r_p)^m (s / r_p)^(l - m).  One thread per
// (child, m) of the children of the nodes [lo, hi).
__global__ void l2l_kernel(const double* __restrict__ center, const double* __restrict__ half,
                           const int4* __restrict__ info, i64 lo, i64 hi, int p1,
                           const double* __restrict__ binomial, int bsz, double* __restrict__ lre,
                           double* __restrict__ lim) {
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    const i64 parent = lo + t / (4 * p1);
    if (parent >= hi) return;
    const int4 pf = info[parent];
    if (pf.z < 0) return;
    const int rem = (int)(t % (4 * p1));
    const i64 child = pf.z + rem / p1;
    const int m = rem % p1;
    const double rp = 1.4142135623730951 * half[parent];
    const double rc = 1.4142135623730951 * half[child];
    const double sr = (center[2 * child] - center[2 * parent]) / rp;
    const double si = (center[2 * child + 1] - center[2 * parent + 1]) / rp;
    double acc_r = 0.0, acc_i = 0.0, pr = 1.0, pi = 0.0;
    for (int l = m; l < p1; ++l) {
        const double c = binomial[l * bsz + m];
        const double br = lre[parent * p1 + l] * c, bi = lim[parent * p1 + l] * c;
        acc_r += br * pr - bi * pi;
        acc_i += br * pi + bi * pr;
        const double tr = pr * sr - pi * si;
        pi = pr * si + pi * sr;
        pr = tr;
    }
    double f = 1.0;
    const double ratio = rc / rp;
    for (int i = 0; i < m; ++i) f *= ratio;
    lre[child * p1 + m] += acc_r * f;
    lim[child * p1 + m] += acc_i * f;
}

// The same for p + 1 <= 16 coefficients, 16 lanes per (parent, child): lane l loads the
// parent's coefficient l once and the others come by shuffles within the 16.  Same terms
// in the same order as l2l_kernel.
__global__ void l2l16_kernel(const double* __restrict__ center, const double* __restrict__ half,
                             const int4* __restrict__ info, i64 lo, i64 hi, int p1,
                             const double* __restrict__ binomial, int bsz,
                             double* __restrict__ lre, double* __restrict__ lim) {
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    const i64 parent = lo + t / 64;
    const int m = (int)(t % 16);
    if (parent >= hi) return;  // uniform over the 64 lanes of a parent
    const int4 pf = info[parent];
    if (pf.z < 0) return;
    const i64 child = pf.z + (t % 64) / 16;
    const double my_re = m < p1 ? lre[parent * p1 + m] : 0.0;
    const double my_im = m < p1 ? lim[parent * p1 + m] : 0.0;
    const double rp = 1.4142135623730951 * half[parent];
    const double rc = 1.4142135623730951 * half[child];
    const double sr = (center[2 * child] - center[2 * parent]) / rp;
    const double si = (center[2 * child + 1] - center[2 * parent + 1]) / rp;
    double acc_r = 0.0, acc_i = 0.0, pr = 1.0, pi = 0.0;
    for (int j = 0; j < 16; ++j) {
        const int l = m + j;
        const double b_re = __shfl(my_re, l < 16 ? l : 15, 16);
        const double b_im = __shfl(my_im, l < 16 ? l : 15, 16);
        if (l < p1) {
            const double c = binomial[l * bsz + m];
            const double br = b_re * c, bi = b_im * c;
            acc_r += br * pr - bi * pi;
            acc_i += br * pi + bi * pr;
            const double tr = pr * sr - pi * si;
            pi = pr * si + pi * sr;
            pr = tr;
        }
    }
    if (m >= p1) return;
    double f = 1.0;
    const double ratio = rc / rp;
    for (int i = 0; i < m; ++i) f *= ratio;
    lre[child * p1 + m] += acc_r * f;
    lim[child * p1 + m] += acc_i * f;
}

// ------------------------------------------------------------------------------ L2P + P2P
// kernels/fmm_evaluate_host/fmm_evaluate_host_pyccel.py:14-100: L2P W = sum_l b_l u^l,
// phi += -2 Re W, E_x - i E_y += 2 W'; near field phi -= q ln(r^2 + eps^2),
// E += 2 q d / (r^2 + eps^2) over the P2P partner leaves, each particle's own charge
// excluded; a charged coincident pair sets status.
// One lane per target, in tree order (64 consecutive particles per wavefront), each walking its leaf's partner row on its own.  Lanes of one leaf read
// the same sources (one broadcast load), and a row's order fixes the order of every sum,
// so results are deterministic.

// MODE 0 (fmm): targets are the sources themselves, in tree order; a particle's own slot is
// skipped and a coincident pair is an error only when both are charged
// (fmm_evaluate_host_pyccel.py:64-100).  MODE 1 (fmm_at): targets of their own tree;
// uncharged sources and source exclude[i] are skipped; any remaining coincidence is an
// error (log_targets_host_pyccel.py:14-94).
// KERNEL TD3_LOG: scaled complex locals (lre, lim; p1 = order + 1 coefficients);
// TD3_COULOMB: the Cartesian locals T_g of degree <= order in lre (reference layout),
// softened near field q / sqrt(r^2 + eps^2) (cart_evaluate_host_pyccel.py:16-110,
// cart_targets_host_pyccel.py:16-98).
// The exact sum over one partner leaf's sources (sf: the leaf's {start, end}) at a target
// (tx, ty), into the FP64 sums.  MODE 0: the sources themselves (k the target's own index,
// skipped; tq its charge); MODE 1: other points (sources without charge and the source
// tskip skipped).  FP64: every term in FP64.  Mixed: differences in FP64 (exact for nearby
// particles, whatever cells hold them), the kernel in FP32, FP32 sums over the leaf, then
// FP64.  A pair at distance 0 (softening 0) sets bad instead of adding a term.
// One pair's term: the source record sv (index ks) at the target.  FP64: into the FP64 sums;
// mixed: into the FP32 partial sums fp, fx, fy of the partner leaf (the FP64 sums when the
// FP32 distance is zero or below range).  The FP64 sums are explicit fma calls, so every
// kernel built on these helpers rounds each term the same way.
template <int KERNEL, int PREC, int MODE>
__device__ __forceinline__ void p2p_term(double tx, double ty, double tq, i64 k, i64 tskip,
                                         int ks, const double4& sv,
                                         const i64* __restrict__ sperm, double eps2,
                                         int with_potential, double& acc_p, double& acc_x,
                                         double& acc_y, float& fp, float& fx, float& fy,
                                         bool& bad) {
    const double sq = sv.z;
    if (MODE == 0 && ks == k) return;
    if (MODE == 1 && (sq == 0.0 || (tskip >= 0 && sperm[ks] == tskip))) return;
    if (PREC == TD3_FP64) {
        const double dx = tx - sv.x, dy = ty - sv.y;
        const double r2 = fma(dx, dx, fma(dy, dy, eps2));
        if (r2 == 0.0) {
            if (MODE == 1 || (tq != 0.0 && sq != 0.0)) bad = true;
            return;
        }
        if (KERNEL == TD3_COULOMB) {
            const double inv_r = 1.0 / sqrt(r2);
            const double cr3 = sq * inv_r * inv_r * inv_r;
            acc_p = fma(sq, inv_r, acc_p);
            acc_x = fma(cr3, dx, acc_x);
            acc_y = fma(cr3, dy, acc_y);
        } else {
            if (with_potential) acc_p = fma(-sq, log(r2), acc_p);
            const double inv = 2.0 * sq / r2;
            acc_x = fma(inv, dx, acc_x);
            acc_y = fma(inv, dy, acc_y);
        }
    } else {
        const double ddx = tx - sv.x, ddy = ty - sv.y;
        const float dx = (float)ddx, dy = (float)ddy, qf = (float)sq;
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, (float)eps2));
        if (r2 == 0.0f) {
            // zero or below FP32 range: decide on the FP64 pair
            const double dr2 = fma(ddx, ddx, fma(ddy, ddy, eps2));
            if (dr2 == 0.0) {
                if (MODE == 1 || (tq != 0.0 && sq != 0.0)) bad = true;
                return;
            }
            if (KERNEL == TD3_COULOMB) {
                const double inv_r = 1.0 / sqrt(dr2);
                const double cr3 = sq * inv_r * inv_r * inv_r;
                acc_p = fma(sq, inv_r, acc_p);
                acc_x = fma(cr3, ddx, acc_x);
                acc_y = fma(cr3, ddy, acc_y);
            } else {
                if (with_potential) acc_p = fma(-sq, log(dr2), acc_p);
                const double inv = 2.0 * sq / dr2;
                acc_x = fma(inv, ddx, acc_x);
                acc_y = fma(inv, ddy, acc_y);
            }
            return;
        }
        if (KERNEL == TD3_COULOMB) {
            const float inv_r = __builtin_amdgcn_rsqf(r2);
            const float c = qf * inv_r;
            const float cr3 = c * inv_r * inv_r;
            fp += c;
            fx = fmaf(cr3, dx, fx);
            fy = fmaf(cr3, dy, fy);
        } else {
            // -q ln r2 = -ln 2 q log2 r2 and 2 q d / r2; the constants at the end
            if (with_potential) fp = fmaf(qf, __builtin_amdgcn_logf(r2), fp);
            const float inv = qf * __builtin_amdgcn_rcpf(r2);
            fx = fmaf(inv, dx, fx);
            fy = fmaf(inv, dy, fy);
        }
    }
}

// The end of a partner leaf in mixed precision: its FP32 partial sums into the FP64 sums.
template <int KERNEL, int PREC>
__device__ __forceinline__ void p2p_fold(float fp, float fx, float fy, double& acc_p,
                                         double& acc_x, double& acc_y) {
    if (PREC == TD3_FP64) return;
    if (KERNEL == TD3_COULOMB) {
        acc_p += (double)fp;
        acc_x += (double)fx;
        acc_y += (double)fy;
    } else {
        acc_p = fma(-0.6931471805599453, (double)fp, acc_p);
        acc_x += 2.0 * (double)fx;
        acc_y += 2.0 * (double)fy;
    }
}

template <int KERNEL, int PREC, int MODE>
__device__ __forceinline__ void p2p_leaf(double tx, double ty, double tq, i64 k, i64 tskip,
                                         int4 sf, const double4* __restrict__ src4,
                                         const i64* __restrict__ sperm, double eps2,
                                         int with_potential, double& acc_p, double& acc_x,
                                         double& acc_y, bool& bad) {
    float fp = 0.0f, fx = 0.0f, fy = 0.0f;
    for (int ks = sf.x; ks < sf.y; ++ks)
        p2p_term<KERNEL, PREC, MODE>(tx, ty, tq, k, tskip, ks, src4[ks], sperm, eps2,
                                     with_potential, acc_p, acc_x, acc_y, fp, fx, fy, bad);
    p2p_fold<KERNEL, PREC>(fp, fx, fy, acc_p, acc_x, acc_y);
}

// p2p_leaf over a whole P2P row (entries r0 .. r1 - 1 of pcol) as one loop over its sources:
// a lane moves to its next partner leaf when it has run out of the current one, so the
// lanes of a wavefront, which walk different rows, stay busy until the longest row ends
// rather than until each partner position's largest leaf ends.  Same terms, same order and
// same per-leaf partial sums as p2p_leaf called partner by partner.  Some lane of a
// wavefront moves to its next partner at almost every step, so the move is paid on every
// iteration: the FP64 terms outweigh it (the near field of 1M clustered and uniform particles,
// leaf 8, theta 0.5: 0.80 against 0.93 ms), the mixed ones do not (0.44 against 0.39 ms, both
// loading columns ahead), so only FP64 takes this loop.
template <int KERNEL, int PREC>
__device__ __forceinline__ void p2p_row(double tx, double ty, double tq, i64 k, u32 r0, u32 r1,
                                        const int* __restrict__ pcol,
                                        const int4* __restrict__ sinfo,
                                        const double4* __restrict__ src4, double eps2,
                                        int with_potential, double& acc_p, double& acc_x,
                                        double& acc_y, bool& bad) {
    float fp = 0.0f, fx = 0.0f, fy = 0.0f;
    u32 r = r0;
    int ks = 0, ke = 0;
    bool started = false;
    // partner r's range and partner r + 1's column are loaded ahead, so taking a partner
    // never waits on a column before its range can be loaded
    int4 sf_next = r0 < r1 ? sinfo[pcol[r0]] : make_int4(0, 0, 0, 0);
    int pc_next = r0 + 1 < r1 ? pcol[r0 + 1] : 0;
    for (;;) {
        while (ks >= ke) {
            if (started) {
                p2p_fold<KERNEL, PREC>(fp, fx, fy, acc_p, acc_x, acc_y);
                fp = fx = fy = 0.0f;
            }
            if (r >= r1) return;
            started = true;
            ks = sf_next.x;
            ke = sf_next.y;
            ++r;
            if (r < r1) sf_next = sinfo[pc_next];
            if (r + 1 < r1) pc_next = pcol[r + 1];
        }
        p2p_term<KERNEL, PREC, 0>(tx, ty, tq, k, -1, ks, src4[ks], nullptr, eps2,
                                  with_potential, acc_p, acc_x, acc_y, fp, fx, fy, bad);
        ++ks;
    }
}

// The far field at a target of leaf `leaf` from its local expansion, added to the exact
// sums, written at input index i.  Coulomb: phi += sum_g T_g u^g, E -= grad (u = z - c,
// unscaled Cartesian locals); log: Horner on the scaled local expansion at (z - c) / r.
template <int KERNEL>
__device__ __forceinline__ void l2p_write(double tx, double ty, i64 leaf,
                                          const double* __restrict__ tcenter,
                                          const double* __restrict__ thalf,
                                          const double* __restrict__ lre,
                                          const double* __restrict__ lim, int p1, int order,
                                          double acc_p, double acc_x, double acc_y,
                                          int with_potential, i64 i, i64 m,
                                          double* __restrict__ phi, double* __restrict__ field) {
    const double cx = tcenter[2 * leaf], cy = tcenter[2 * leaf + 1];
    if (KERNEL == TD3_COULOMB) {
        const double ux = tx - cx, uy = ty - cy;
        const double* T = lre + leaf * (i64)p1;
        double w = 0.0, gx = 0.0, gy = 0.0, pyb = 1.0, pyb1 = 0.0;
        for (int by = 0; by <= order; ++by) {
            double px = 1.0, px1 = 0.0;
            for (int bx = 0; bx + by <= order; ++bx) {
                const double t = T[(bx + by) * (bx + by + 1) / 2 + by];
                w += t * px * pyb;
                gx += bx * t * px1 * pyb;
                gy += by * t * px * pyb1;
                px1 = px;
                px *= ux;
            }
            pyb1 = pyb;
            pyb *= uy;
        }
        phi[i] = with_potential ? acc_p + w : 0.0;
        field[i] = acc_x - gx;
        field[m + i] = acc_y - gy;
    } else {
        const double inv_r = 1.0 / (1.4142135623730951 * thalf[leaf]);
        const double ur = (tx - cx) * inv_r, ui = (ty - cy) * inv_r;
        const double* bre = lre + leaf * (i64)p1;
        const double* bim = lim + leaf * (i64)p1;
        double w_re = bre[p1 - 1], w_im = bim[p1 - 1];
        double d_re = 0.0, d_im = 0.0;
        for (int l = p1 - 2; l >= 0; --l) {
            const double nd = d_re * ur - d_im * ui + w_re;
            d_im = d_re * ui + d_im * ur + w_im;
            d_re = nd;
            const double nw = w_re * ur - w_im * ui + bre[l];
            w_im = w_re * ui + w_im * ur + bim[l];
            w_re = nw;
        }
        phi[i] = with_potential ? acc_p - 2.0 * w_re : 0.0;
        field[i] = acc_x + 2.0 * d_re * inv_r;
        field[m + i] = acc_y - 2.0 * d_im * inv_r;
    }
}

// L2P + P2P with a lane per target (the FMM at its own sources, MODE 0): the lane sums its
// leaf's P2P row (the next partner leaf's range loaded while this one's sources are summed).
template <int KERNEL, int PREC, int MODE>
__global__ __launch_bounds__(256) void eval_kernel(
    const double* __restrict__ tpos, const i64* __restrict__ tperm, const i64* __restrict__ exclude,
    i64 m, const double* __restrict__ tcenter, const double* __restrict__ thalf,
    const int4* __restrict__ tinfo, i64 t_count, const double* __restrict__ lre,
    const double* __restrict__ lim, int p1, int order, const u32* __restrict__ prow,
    const int* __restrict__ pcol, const double* __restrict__ spos,
    const double* __restrict__ scharge, const i64* __restrict__ sperm, i64 n,
    const int4* __restrict__ sinfo, const int* __restrict__ leaf_of,
    const double4* __restrict__ src4, double eps2, int with_potential,
    double* __restrict__ phi, double* __restrict__ field, i64* __restrict__ status) {
    const i64 k = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const i64 leaf = leaf_of[k];
    const double tx = tpos[k], ty = tpos[m + k];
    const double tq = MODE == 0 ? scharge[k] : 0.0;
    const i64 tskip = MODE == 1 ? exclude[tperm[k]] : -1;
    double acc_p = 0.0, acc_x = 0.0, acc_y = 0.0;
    bool bad = false;
    const u32 r0 = prow[leaf], r1 = prow[leaf + 1];
    if (MODE == 0 && PREC == TD3_FP64) {
        p2p_row<KERNEL, PREC>(tx, ty, tq, k, r0, r1, pcol, sinfo, src4, eps2, with_potential,
                              acc_p, acc_x, acc_y, bad);
    } else {
        // the next partner's range and the one after's column loaded ahead, as in p2p_row
        int4 sf_next = r0 < r1 ? sinfo[pcol[r0]] : make_int4(0, 0, 0, 0);
        int pc_next = r0 + 1 < r1 ? pcol[r0 + 1] : 0;
        for (u32 r = r0; r < r1; ++r) {
            const int4 sf = sf_next;
            if (r + 1 < r1) sf_next = sinfo[pc_next];
            if (r + 2 < r1) pc_next = pcol[r + 2];
            p2p_leaf<KERNEL, PREC, MODE>(tx, ty, tq, k, tskip, sf, src4, sperm, eps2,
                                         with_potential, acc_p, acc_x, acc_y, bad);
        }
    }
    const i64 i = tperm[k];
    if (bad) td3_report(status, i + 1);
    l2p_write<KERNEL>(tx, ty, leaf, tcenter, thalf, lre, lim, p1, order, acc_p, acc_x, acc_y,
                      with_potential, i, m, phi, field);
}

// The two halves of eval_kernel (MODE 0) as separate launches, so that the near field can
// run while the far field is still being translated: eval_near_kernel writes each
// particle's exact sums (potential, field x, field y) in tree order to near (3, N);
// eval_far_kernel adds the local expansion's terms and writes the outputs in input order.
// The sums and their order are eval_kernel's.
template <int KERNEL, int PREC>
__global__ __launch_bounds__(256) void eval_near_kernel(
    const double* __restrict__ tpos, const i64* __restrict__ tperm, i64 m, i64 k0, i64 k1,
    const u32* __restrict__ prow, const int* __restrict__ pcol, const int4* __restrict__ sinfo,
    const int* __restrict__ leaf_of, const double4* __restrict__ src4, double eps2,
    int with_potential, double* __restrict__ near, i64* __restrict__ status) {
    const i64 k = k0 + (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= k1) return;
    const i64 leaf = leaf_of[k];
    const double tx = tpos[k], ty = tpos[m + k];
    const double tq = src4[k].z;
    double acc_p = 0.0, acc_x = 0.0, acc_y = 0.0;
    bool bad = false;
    const u32 r0 = prow[leaf], r1 = prow[leaf + 1];
    if (PREC == TD3_FP64) {
        p2p_row<KERNEL, PREC>(tx, ty, tq, k, r0, r1, pcol, sinfo, src4, eps2, with_potential,
                              acc_p, acc_x, acc_y, bad);
    } else {
        // the next partner's range and the one after's column loaded ahead, as in p2p_row
        int4 sf_next = r0 < r1 ? sinfo[pcol[r0]] : make_int4(0, 0, 0, 0);
        int pc_next = r0 + 1 < r1 ? pcol[r0 + 1] : 0;
        for (u32 r = r0; r < r1; ++r) {
            const int4 sf = sf_next;
            if (r + 1 < r1) sf_next = sinfo[pc_next];
            if (r + 2 < r1) pc_next = pcol[r + 2];
            p2p_leaf<KERNEL, PREC, 0>(tx, ty, tq, k, -1, sf, src4, nullptr, eps2,
                                      with_potential, acc_p, acc_x, acc_y, bad);
        }
    }
    if (bad) td3_report(status, tperm[k] + 1);
    near[k] = acc_p;
    near[m + k] = acc_x;
    near[2 * m + k] = acc_y;
}

template <int KERNEL>
__global__ __launch_bounds__(256) void eval_far_kernel(
    const double* __restrict__ tpos, const i64* __restrict__ tperm, i64 m,
    const double* __restrict__ tcenter, const double* __restrict__ thalf,
    const double* __restrict__ lre, const double* __restrict__ lim, int p1, int order,
    const int* __restrict__ leaf_of, const double* __restrict__ near, int with_potential,
    double* __restrict__ phi, double* __restrict__ field) {
    const i64 k = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    l2p_write<KERNEL>(tpos[k], tpos[m + k], leaf_of[k], tcenter, thalf, lre, lim, p1, order,
                      near[k], near[m + k], near[2 * m + k], with_potential, tperm[k], m, phi,
                      field);
}

// L2P + P2P at other points (fmm_at, MODE 1), a wavefront per target leaf: its targets go
// tl = 4, 8 or 16 at a time (by the leaf's size), each with 64 / tl lanes that split the
// leaf's P2P row between them (lane s of a target takes partner leaves s, s + 64 / tl, ...),
// then sum their partial sums over DPP in a fixed order.  A leaf whose row holds thousands
// of sources (a grid over a dense cluster) is spread over the whole wavefront instead of
// one lane per target.
template <int KERNEL, int PREC>
__global__ __launch_bounds__(256) void eval_rows_kernel(
    const double* __restrict__ tpos, const i64* __restrict__ tperm,
    const i64* __restrict__ exclude, i64 m, const double* __restrict__ tcenter,
    const double* __restrict__ thalf, const int4* __restrict__ tinfo, i64 t_count,
    const double* __restrict__ lre, const double* __restrict__ lim, int p1, int order,
    const u32* __restrict__ prow, const int* __restrict__ pcol, const i64* __restrict__ sperm,
    const int4* __restrict__ sinfo, const double4* __restrict__ src4, double eps2,
    int with_potential, double* __restrict__ phi, double* __restrict__ field,
    i64* __restrict__ status) {
    const int lane = td3_lane();
    const i64 leaf = ((i64)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (leaf >= t_count) return;  // uniform over the wavefront
    const int4 tf = tinfo[leaf];
    if (tf.z >= 0 || tf.y <= tf.x) return;  // an internal or empty cell
    const int nt = tf.y - tf.x;
    const int tl = nt <= 4 ? 4 : (nt <= 8 ? 8 : 16);
    const int lanes = 64 / tl, sub = lane & (lanes - 1);
    const u32 r0 = prow[leaf], r1 = prow[leaf + 1];
    for (int c0 = 0; c0 < nt; c0 += tl) {
        const int tloc = c0 + lane / lanes;
        const bool tvalid = tloc < nt;
        const i64 k = tf.x + (tvalid ? tloc : 0);
        const double tx = tpos[k], ty = tpos[m + k];
        const i64 tskip = exclude[tperm[k]];
        double acc_p = 0.0, acc_x = 0.0, acc_y = 0.0;
        bool bad = false;
        if (tvalid) {
            u32 r = r0 + (u32)sub;
            int4 sf_next = r < r1 ? sinfo[pcol[r]] : make_int4(0, 0, 0, 0);
            for (; r < r1; r += (u32)lanes) {
                const int4 sf = sf_next;
                if (r + lanes < r1) sf_next = sinfo[pcol[r + lanes]];
                p2p_leaf<KERNEL, PREC, 1>(tx, ty, 0.0, k, tskip, sf, src4, sperm, eps2,
                                          with_potential, acc_p, acc_x, acc_y, bad);
            }
        }
        acc_p = td3_group_sum(acc_p, lanes);
        acc_x = td3_group_sum(acc_x, lanes);
        acc_y = td3_group_sum(acc_y, lanes);
        const u64 group = (lanes == 64 ? ~0ull : ((1ull << lanes) - 1ull)) << (lane & ~(lanes - 1));
        const bool group_bad = (__ballot(bad) & group) != 0;
        if (tvalid && sub == 0) {
            const i64 i = tperm[k];
            if (group_bad) td3_report(status, i + 1);
            l2p_write<KERNEL>(tx, ty, leaf, tcenter, thalf, lre, lim, p1, order, acc_p, acc_x,
                              acc_y, with_potential, i, m, phi, field);
        }
    }
}

// The log kernel's evaluation for adjoints, in FP64 (the far field as the locals were
// built).  EXT 1: the sources also carry dipoles D (sdip: x components, then y, in tree
// order), whose potential at d = target - source is -2 D.d / (r^2 + eps^2) and field
// 2 D / (r^2 + eps^2) - 4 (D.d) d / (r^2 + eps^2)^2 (their far field is in the moments).
// EXT 2: the Jacobian dE/dx of the field at each target, rows xx, xy, yy of jac; the far
// field's is 2 W'' (E_x - i E_y = 2 W' is analytic), the near field's
// 2 q (I / (r^2 + eps^2) - 2 d d^T / (r^2 + eps^2)^2).  Targets, exclusions and status as
// in eval_kernel; a dipole counts as a source as a charge does.
// The near sum of log_ext_eval_kernel over one partner leaf (sf: its {start, end}), into
// the FP64 sums of the potential, the field and (EXT 2) the Jacobian.
template <int MODE, int EXT>
__device__ __forceinline__ void ext_p2p_leaf(double tx, double ty, double tq, i64 k, i64 tskip,
                                             int4 sf, const double4* __restrict__ src4,
                                             const double* __restrict__ sdip,
                                             const i64* __restrict__ sperm, i64 n, double eps2,
                                             double (&acc)[6], bool& bad) {
    for (int ks = sf.x; ks < sf.y; ++ks) {
        if (MODE == 0 && ks == k) continue;
        const double4 sv = src4[ks];
        const double sq = sv.z;
        const double px = EXT == 1 ? sdip[ks] : 0.0, py = EXT == 1 ? sdip[n + ks] : 0.0;
        const bool live = sq != 0.0 || px != 0.0 || py != 0.0;
        if (MODE == 1 && (!live || (tskip >= 0 && sperm[ks] == tskip))) continue;
        const double dx = tx - sv.x, dy = ty - sv.y;
        const double r2 = dx * dx + dy * dy + eps2;
        if (r2 == 0.0) {
            if (MODE == 1 || (tq != 0.0 && live)) bad = true;
            continue;
        }
        acc[0] -= sq * log(r2);
        const double c = 2.0 * sq / r2;
        acc[1] += c * dx;
        acc[2] += c * dy;
        if (EXT == 1) {
            const double inv = 1.0 / r2;
            const double pd = px * dx + py * dy;
            const double c4 = 4.0 * pd * inv * inv;
            acc[0] -= 2.0 * pd * inv;
            acc[1] += 2.0 * px * inv - c4 * dx;
            acc[2] += 2.0 * py * inv - c4 * dy;
        }
        if (EXT == 2) {
            const double c2 = 2.0 * c / r2;
            acc[3] += c - c2 * dx * dx;
            acc[4] -= c2 * dx * dy;
            acc[5] += c - c2 * dy * dy;
        }
    }
}

// The far field of log_ext_eval_kernel at a target of leaf `leaf` (Horner on the scaled
// local expansion at u' = (z - c) / r: W, W' and W'' / 2), added to the near sums, written
// at input index i.
template <int EXT>
__device__ __forceinline__ void ext_l2p_write(double tx, double ty, i64 leaf,
                                              const double* __restrict__ tcenter,
                                              const double* __restrict__ thalf,
                                              const double* __restrict__ lre,
                                              const double* __restrict__ lim, int p1,
                                              const double (&acc)[6], i64 i, i64 m,
                                              double* __restrict__ phi,
                                              double* __restrict__ field,
                                              double* __restrict__ jac) {
    const double inv_r = 1.0 / (1.4142135623730951 * thalf[leaf]);
    const double ur = (tx - tcenter[2 * leaf]) * inv_r, ui = (ty - tcenter[2 * leaf + 1]) * inv_r;
    const double* bre = lre + leaf * (i64)p1;
    const double* bim = lim + leaf * (i64)p1;
    double w_re = bre[p1 - 1], w_im = bim[p1 - 1];
    double d_re = 0.0, d_im = 0.0, e_re = 0.0, e_im = 0.0;
    for (int l = p1 - 2; l >= 0; --l) {
        const double ne = e_re * ur - e_im * ui + d_re;
        e_im = e_re * ui + e_im * ur + d_im;
        e_re = ne;
        const double nd = d_re * ur - d_im * ui + w_re;
        d_im = d_re * ui + d_im * ur + w_im;
        d_re = nd;
        const double nw = w_re * ur - w_im * ui + bre[l];
        w_im = w_re * ui + w_im * ur + bim[l];
        w_re = nw;
    }
    phi[i] = acc[0] - 2.0 * w_re;
    field[i] = acc[1] + 2.0 * d_re * inv_r;
    field[m + i] = acc[2] - 2.0 * d_im * inv_r;
    if (EXT == 2) {
        const double s = 4.0 * inv_r * inv_r;
        jac[i] = acc[3] + s * e_re;
        jac[m + i] = acc[4] - s * e_im;
        jac[2 * m + i] = acc[5] - s * e_re;
    }
}

template <int MODE, int EXT>
__global__ __launch_bounds__(256) void log_ext_eval_kernel(
    const double* __restrict__ tpos, const i64* __restrict__ tperm, const i64* __restrict__ exclude,
    i64 m, const double* __restrict__ tcenter, const double* __restrict__ thalf,
    const double* __restrict__ lre, const double* __restrict__ lim, int p1,
    const u32* __restrict__ prow, const int* __restrict__ pcol, const i64* __restrict__ sperm,
    i64 n, const int4* __restrict__ sinfo, const int* __restrict__ leaf_of,
    const double4* __restrict__ src4, const double* __restrict__ sdip, double eps2,
    double* __restrict__ phi, double* __restrict__ field, double* __restrict__ jac,
    i64* __restrict__ status) {
    const i64 k = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const i64 leaf = leaf_of[k];
    const double tx = tpos[k], ty = tpos[m + k];
    const double tq = MODE == 0 ? src4[k].z : 0.0;
    const i64 tskip = MODE == 1 ? exclude[tperm[k]] : -1;
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    bool bad = false;
    const u32 r0 = prow[leaf], r1 = prow[leaf + 1];
    for (u32 r = r0; r < r1; ++r)
        ext_p2p_leaf<MODE, EXT>(tx, ty, tq, k, tskip, sinfo[pcol[r]], src4, sdip, sperm, n, eps2,
                                acc, bad);
    const i64 i = tperm[k];
    if (bad) td3_report(status, i + 1);
    ext_l2p_write<EXT>(tx, ty, leaf, tcenter, thalf, lre, lim, p1, acc, i, m, phi, field, jac);
}

// The same at other points (fmm_at, MODE 1) with a wavefront per target leaf, as
// eval_rows_kernel: 64 / tl lanes per target split the leaf's P2P row, DPP sums.
template <int EXT>
__global__ __launch_bounds__(256) void log_ext_eval_rows_kernel(
    const double* __restrict__ tpos, const i64* __restrict__ tperm, const i64* __restrict__ exclude,
    i64 m, const double* __restrict__ tcenter, const double* __restrict__ thalf,
    const int4* __restrict__ tinfo, i64 t_count, const double* __restrict__ lre,
    const double* __restrict__ lim, int p1, const u32* __restrict__ prow,
    const int* __restrict__ pcol, const i64* __restrict__ sperm, i64 n,
    const int4* __restrict__ sinfo, const double4* __restrict__ src4,
    const double* __restrict__ sdip, double eps2, double* __restrict__ phi,
    double* __restrict__ field, double* __restrict__ jac, i64* __restrict__ status) {
    const int lane = td3_lane();
    const i64 leaf = ((i64)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (leaf >= t_count) return;  // uniform over the wavefront
    const int4 tf = tinfo[leaf];
    if (tf.z >= 0 || tf.y <= tf.x) return;  // an internal or empty cell
    const int nt = tf.y - tf.x;
    const int tl = nt <= 4 ? 4 : (nt <= 8 ? 8 : 16);
    const int lanes = 64 / tl, sub = lane & (lanes - 1);
    const u32 r0 = prow[leaf], r1 = prow[leaf + 1];
    constexpr int kAcc = EXT == 2 ? 6 : 3;
    for (int c0 = 0; c0 < nt; c0 += tl) {
        const int tloc = c0 + lane / lanes;
        const bool tvalid = tloc < nt;
        const i64 k = tf.x + (tvalid ? tloc : 0);
        const double tx = tpos[k], ty = tpos[m + k];
        const i64 tskip = exclude[tperm[k]];
        double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        bool bad = false;
        if (tvalid)
            for (u32 r = r0 + (u32)sub; r < r1; r += (u32)lanes)
                ext_p2p_leaf<1, EXT>(tx, ty, 0.0, k, tskip, sinfo[pcol[r]], src4, sdip, sperm, n,
                                     eps2, acc, bad);
#pragma unroll
        for (int a = 0; a < kAcc; ++a) acc[a] = td3_group_sum(acc[a], lanes);
        const u64 group = (lanes == 64 ? ~0ull : ((1ull << lanes) - 1ull)) << (lane & ~(lanes - 1));
        const bool group_bad = (__ballot(bad) & group) != 0;
        if (tvalid && sub == 0) {
            const i64 i = tperm[k];
            if (group_bad) td3_report(status, i + 1);
            ext_l2p_write<EXT>(tx, ty, leaf, tcenter, thalf, lre, lim, p1, acc, i, m, phi, field,
                               jac);
        }
    }
}

// (x, y, q, 0) of every sorted source: one 32-byte load per source in the near field.
__global__ void pack_sources_kernel(const double* __restrict__ pos, const double* __restrict__ charge,
                                    i64 n, double4* __restrict__ out) {
    const i64 k = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    out[k] = make_double4(pos[k], pos[n + k], charge[k], 0.0);
}

// leaf_of[k] = the leaf holding sorted particle k.
__global__ void leaf_index_kernel(const int4* __restrict__ info, i64 count, int* __restrict__ leaf_of) {
    const i64 node = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= count) return;
    const int4 f = info[node];
    if (f.z >= 0) return;
    for (int k = f.x; k < f.y; ++k) leaf_of[k] = (int)node;
}

template <int KERNEL, int MODE>
int launch_eval(int precision, hipStream_t s, const double* tpos, const i64* tperm,
                const i64* exclude, i64 m, const double* tcenter, const double* thalf,
                const int4* tinfo, i64 t_count, const double* lre, const double* lim, int order,
                const u32* prow, const int* pcol, const double* spos, const double* scharge,
                const i64* sperm, i64 n, const int4* sinfo, const int* leaf_of,
                const double4* src4, double eps2, int with_potential, double* phi,
                double* field, i64* status) {
    TD3_ARG(leaf_of != nullptr, "leaf_of (td3_leaf_index of the target tree)");
    TD3_ARG(src4 != nullptr, "sources4 (td3_pack_sources of the source tree)");
    const unsigned blocks = td3_blocks(m, 256);
    const int p1 = KERNEL == TD3_LOG ? order + 1 : (order + 1) * (order + 2) / 2;
    if (precision == TD3_FP64)
        TD3_TIMED(s, "eval", hipLaunchKernelGGL((eval_kernel<KERNEL, TD3_FP64, MODE>), dim3(blocks), dim3(256), 0, s,
                           tpos, tperm, exclude, m, tcenter, thalf, tinfo, t_count, lre, lim, p1, order,
                           prow, pcol, spos, scharge, sperm, n, sinfo, leaf_of, src4, eps2,
                           with_potential, phi, field, status));
    else
        TD3_TIMED(s, "eval", hipLaunchKernelGGL((eval_kernel<KERNEL, TD3_MIXED, MODE>), dim3(blocks), dim3(256), 0, s,
                           tpos, tperm, exclude, m, tcenter, thalf, tinfo, t_count, lre, lim, p1, order,
                           prow, pcol, spos, scharge, sperm, n, sinfo, leaf_of, src4, eps2,
                           with_potential, phi, field, status));
    TD3_LAUNCHED();
    return TD3_OK;
}

template <int KERNEL>
int launch_eval_rows(int precision, hipStream_t s, const double* tpos, const i64* tperm,
                     const i64* exclude, i64 m, const double* tcenter, const double* thalf,
                     const int4* tinfo, i64 t_count, const double* lre, const double* lim,
                     int order, const u32* prow, const int* pcol, const i64* sperm,
                     const int4* sinfo, const double4* src4, double eps2, int with_potential,
                     double* phi, double* field, i64* status) {
    TD3_ARG(src4 != nullptr, "sources4 (td3_pack_sources of the source tree)");
    const unsigned blocks = td3_blocks(t_count * 64, 256);
    const int p1 = KERNEL == TD3_LOG ? order + 1 : (order + 1) * (order + 2) / 2;
    if (precision == TD3_FP64)
        TD3_TIMED(s, "eval", hipLaunchKernelGGL((eval_rows_kernel<KERNEL, TD3_FP64>), dim3(blocks),
                  dim3(256), 0, s, tpos, tperm, exclude, m, tcenter, thalf, tinfo, t_count, lre,
                  lim, p1, order, prow, pcol, sperm, sinfo, src4, eps2, with_potential, phi,
                  field, status));
    else
        TD3_TIMED(s, "eval", hipLaunchKernelGGL((eval_rows_kernel<KERNEL, TD3_MIXED>), dim3(blocks),
                  dim3(256), 0, s, tpos, tperm, exclude, m, tcenter, thalf, tinfo, t_count, lre,
                  lim, p1, order, prow, pcol, sperm, sinfo, src4, eps2, with_potential, phi,
                  field, status));
    TD3_LAUNCHED();
    return TD3_OK;
}

template <int P, typename T>
int launch_m2l(hipStream_t s, const int* col, const int* dst, const u32* row_start, i64 node_lo,
               i64 node_hi, u32 max_chunks, const double* dcenter, const double* dhalf,
               const double* scenter, const double* shalf, const T* mre, const T* mim, int p1,
               const double* coeff, double* lre, double* lim, T* side) {
    TD3_TIMED(s, "m2l", hipLaunchKernelGGL((m2l_kernel<P, T>), dim3(td3_blocks((i64)max_chunks * 64, 256)), dim3(256), 0, s, col, dst,
                       row_start, node_lo, node_hi, max_chunks, dcenter, dhalf, scenter, shalf, mre,
                       mim, p1, coeff, lre, lim, side));
    TD3_LAUNCHED();
    const i64 items = (node_hi - node_lo) * 2 * p1;
    TD3_TIMED(s, "m2l_fixup", hipLaunchKernelGGL((m2l_fixup_kernel<T>), dim3(td3_blocks(items, 256)), dim3(256), 0, s, row_start,
                       node_lo, node_hi, p1, side, lre, lim));
    TD3_LAUNCHED();
    return TD3_OK;
}

template <typename T>
int dispatch_m2l(hipStream_t s, int order, const int* col, const int* dst, const u32* row_start,
                 i64 node_lo, i64 node_hi, u32 max_chunks, const double* dcenter,
                 const double* dhalf, const double* scenter, const double* shalf, const T* mre,
                 const T* mim, const double* coeff, double* lre, double* lim, T* side) {
    const int p1 = order + 1;
#define TD3_M2L_CASE(PP)                                                                          \
    if (order <= PP)                                                                             \
        return launch_m2l<PP, T>(s, col, dst, row_start, node_lo, node_hi, max_chunks, dcenter,   \
                                 dhalf, scenter, shalf, mre, mim, p1, coeff, lre, lim, side);
    TD3_M2L_CASE(4)
    TD3_M2L_CASE(8)
    TD3_M2L_CASE(10)
    TD3_M2L_CASE(12)
    TD3_M2L_CASE(16)
    TD3_M2L_CASE(20)
    TD3_M2L_CASE(32)
    TD3_M2L_CASE(64)
#undef TD3_M2L_CASE
    td3_set_error("order %d above 64", order);
    return TD3_ERR_ARG;
}

}  // namespace

extern "C" size_t td3_m2l_workspace_bytes(int64_t pair_count, int64_t order, int precision) {
    const i64 waves = (pair_count + 63) / 64 + 1;
    const size_t word = precision == TD3_FP64 ? 8 : 4;
    return (size_t)waves * 2 * 2 * (size_t)(order + 1) * word + 256;
}

extern "C" int td3_scale_moments(int precision, const double* moment_re, const double* moment_im,
                                 const double* node_half, int64_t node_count, int64_t order,
                                 void* out_re, void* out_im, void* stream) {
    TD3_ARG(node_count >= 1 && order >= 1 && order <= 64, "sizes");
    const int p1 = (int)order + 1;
    const i64 items = node_count * p1;
    hipStream_t s = (hipStream_t)stream;
    if (precision == TD3_FP64)
        TD3_TIMED(s, "scale_moments", hipLaunchKernelGGL(scale_moments_kernel<double>, dim3(td3_blocks(items, 256)), dim3(256), 0, s,
                           moment_re, moment_im, node_half, node_count, p1, (double*)out_re,
                           (double*)out_im));
    else
        TD3_TIMED(s, "scale_moments", hipLaunchKernelGGL(scale_moments_kernel<float>, dim3(td3_blocks(items, 256)), dim3(256), 0, s,
                           moment_re, moment_im, node_half, node_count, p1, (float*)out_re,
                           (float*)out_im));
    TD3_LAUNCHED();
    return TD3_OK;
}

extern "C" int td3_m2l(int precision, int64_t order, const int32_t* col, const int32_t* dst,
                       const uint32_t* row_start, int64_t node_lo, int64_t node_hi,
                       const double* dst_center, const double* dst_half, const double* src_center,
                       const double* src_half, const void* scaled_re, const void* scaled_im,
                       const double* coeff, double* local_re, double* local_im, void* workspace,
                       size_t workspace_bytes, void* stream) {
    TD3_ARG(order >= 1 && order <= 64, "order");
    TD3_ARG(node_lo >= 0 && node_hi >= node_lo, "node range");
    hipStream_t s = (hipStream_t)stream;
    if (node_hi == node_lo) return TD3_OK;
    if (order + 1 <= 16 && precision == TD3_FP64 && workspace == nullptr) {
        // FP64 without a workspace: a wavefront per destination row on the matrix cores (a
        // grid by rows, so the pair count need not be known; for a whole tree the
        // lane-per-pair kernel is the faster FP64 one: on gfx950 v_mfma_f64_16x16x4f64 runs
        // at the vector FP64 rate, and that kernel does less work around the product)
        const dim3 grid(td3_blocks((node_hi - node_lo) * 64, 256));
        TD3_TIMED(s, "m2l_mfma64", hipLaunchKernelGGL(m2l_mfma_kernel, grid, dim3(256), 0, s, col,
                  row_start, node_lo, node_hi, dst_center, dst_half, src_center, src_half,
                  (const double*)scaled_re, (const double*)scaled_im, (int)order + 1, coeff,
                  local_re, local_im));
        TD3_LAUNCHED();
        return TD3_OK;
    }
    // the side buffer bounds the chunks: td3_m2l_workspace_bytes(pairs) holds two slots
    // per 64 pairs (+1) of the rows [node_lo, node_hi)
    const size_t word = precision == TD3_FP64 ? 8 : 4;
    const size_t slot_pair = 2 * 2 * (size_t)(order + 1) * word;
    TD3_ARG(workspace != nullptr && workspace_bytes >= slot_pair + 256, "workspace too small");
    const size_t chunks = (workspace_bytes - 256) / slot_pair;
    const u32 max_chunks = chunks > 0xffffffffull ? 0xffffffffu : (u32)chunks;
    if (precision == TD3_FP64)
        return dispatch_m2l<double>(s, (int)order, col, dst, row_start, node_lo, node_hi,
                                    max_chunks, dst_center, dst_half, src_center, src_half,
                                    (const double*)scaled_re, (const double*)scaled_im, coeff,
                                    local_re, local_im, (double*)workspace);
    return dispatch_m2l<float>(s, (int)order, col, dst, row_start, node_lo, node_hi, max_chunks,
                               dst_center, dst_half, src_center, src_half,
                               (const float*)scaled_re, (const float*)scaled_im, coeff, local_re,
                               local_im, (float*)workspace);
}

extern "C" size_t td3_m2l_record_bytes(int64_t order) {
    return order >= 1 && order <= 15 ? 4 * (size_t)m2l_record_floats((int)order + 1) : 0;
}

extern "C" int td3_m2l_records(int64_t order, const double* moment_re, const double* moment_im,
                               const double* node_center, const double* node_half,
                               int64_t node_count, void* records, void* stream) {
    TD3_ARG(order >= 1 && order <= 15, "order (1..15)");
    TD3_ARG(node_count >= 0, "node count");
    if (node_count == 0) return TD3_OK;
    const int p1 = (int)order + 1, rf = m2l_record_floats(p1);
    hipStream_t s = (hipStream_t)stream;
    const i64 items = node_count * (rf / 4);
    TD3_TIMED(s, "m2l_records", hipLaunchKernelGGL(m2l_records_kernel, dim3(td3_blocks(items, 256)),
              dim3(256), 0, s, moment_re, moment_im, node_center, node_half, node_count, p1, rf,
              (float*)records));
    TD3_LAUNCHED();
    return TD3_OK;
}

extern "C" int td3_m2l_f32(int64_t order, const int32_t* col, const uint32_t* row_start,
                           int64_t node_lo, int64_t node_hi, const double* dst_center,
                           const double* dst_half, const void* records, const double* coeff,
                           double* local_re, double* local_im, void* stream) {
    TD3_ARG(order >= 1 && order <= 15, "order (1..15)");
    TD3_ARG(node_lo >= 0 && node_hi >= node_lo, "node range");
    if (node_hi == node_lo) return TD3_OK;
    const int p1 = (int)order + 1;
    hipStream_t s = (hipStream_t)stream;
    const i64 waves = (node_hi - node_lo + kM2lRows - 1) / kM2lRows;
    TD3_TIMED(s, "m2l_f32", hipLaunchKernelGGL(m2l_f32_kernel, dim3(td3_blocks(waves * 64, 256)),
              dim3(256), 0, s, col, row_start, node_lo, node_hi, dst_center, dst_half,
              (const float*)records, m2l_record_floats(p1), p1, coeff, local_re, local_im));
    TD3_LAUNCHED();
    return TD3_OK;
}

extern "C" int td3_l2l(int64_t order, const double* node_center, const double* node_half,
                       const void* node_info, const int64_t* levels_host, int64_t n_levels,
                       const double* binomial, int64_t bsz, double* local_re, double* local_im,
                       void* stream) {
    TD3_ARG(order >= 1 && order <= 64 && bsz >= order + 1, "order");
    hipStream_t s = (hipStream_t)stream;
    const int p1 = (int)order + 1;
    for (i64 d = 0; d + 1 < n_levels; ++d) {
        const i64 lo = levels_host[d], hi = levels_host[d + 1];
        if (p1 <= 16)
            TD3_TIMED(s, "l2l", hipLaunchKernelGGL(l2l16_kernel, dim3(td3_blocks((hi - lo) * 64, 256)), dim3(256), 0, s,
                      node_center, node_half, (const int4*)node_info, lo, hi, p1, binomial, (int)bsz,
                      local_re, local_im));
        else
            TD3_TIMED(s, "l2l", hipLaunchKernelGGL(l2l_kernel, dim3(td3_blocks((hi - lo) * 4 * p1, 256)), dim3(256), 0, s,
                      node_center, node_half, (const int4*)node_info, lo, hi, p1, binomial, (int)bsz,
                      local_re, local_im));
        TD3_LAUNCHED();
    }
    return TD3_OK;
}

extern "C" int td3_unscale_locals(const double* node_half, int64_t node_count, int64_t order,
                                  double* local_re, double* local_im, void* stream) {
    const int p1 = (int)order + 1;
    hipLaunchKernelGGL(unscale_locals_kernel, dim3(td3_blocks(node_count * p1, 256)), dim3(256), 0,
                       (hipStream_t)stream, local_re, local_im, node_half, node_count, p1);
    TD3_LAUNCHED();
    return TD3_OK;
}

extern "C" int td3_fmm_eval(int kernel, int precision, int64_t order, const double* sorted_pos,
                            const double* sorted_charge, const int64_t* perm, int64_t n,
                            const double* node_center, const double* node_half,
                            const void* node_info, int64_t node_count, const double* local_re,
                            const double* local_im, const uint32_t* p2p_row, const int32_t* p2p_col,
                            const int32_t* leaf_of, const void* sources4, double softening,
                            int with_potential, double* phi, double* field, int64_t* status,
                            void* stream) {
    TD3_ARG(kernel == TD3_COULOMB || kernel == TD3_LOG, "kernel");
    TD3_ARG(order >= 1 && order <= 64, "order");
    TD3_ARG(n >= 1 && node_count >= 1, "sizes");
    auto go = kernel == TD3_LOG ? launch_eval<TD3_LOG, 0> : launch_eval<TD3_COULOMB, 0>;
    return go(precision, (hipStream_t)stream, sorted_pos, perm, nullptr, n, node_center, node_half,
              (const int4*)node_info, node_count, local_re, local_im, (int)order, p2p_row, p2p_col,
              sorted_pos, sorted_charge, perm, n, (const int4*)node_info, leaf_of,
              (const double4*)sources4, softening * softening, with_potential, phi, field,
              status);
}

extern "C" int td3_fmm_near(int kernel, int precision, const double* sorted_pos,
                            const int64_t* perm, int64_t n, int64_t first, int64_t count,
                            const void* node_info, const uint32_t* p2p_row,
                            const int32_t* p2p_col, const int32_t* leaf_of, const void* sources4,
                            double softening, int with_potential, double* near, int64_t* status,
                            void* stream) {
    TD3_ARG(kernel == TD3_COULOMB || kernel == TD3_LOG, "kernel");
    TD3_ARG(n >= 1, "sizes");
    TD3_ARG(first >= 0 && count >= 0 && first + count <= n, "target range");
    TD3_ARG(leaf_of != nullptr && sources4 != nullptr, "leaf_of, sources4");
    if (count == 0) return TD3_OK;
    hipStream_t s = (hipStream_t)stream;
    const dim3 grid(td3_blocks(count, 256));
    const double eps2 = softening * softening;
    const int4* info = (const int4*)node_info;
    const double4* src4 = (const double4*)sources4;
#define TD3_NEAR(K, P)                                                                           \
    TD3_TIMED(s, "eval", hipLaunchKernelGGL((eval_near_kernel<K, P>), grid, dim3(256), 0, s,     \
              sorted_pos, perm, n, first, first + count, p2p_row, p2p_col, info, leaf_of, src4,    \
              eps2, with_potential, near, status))
    if (kernel == TD3_LOG) {
        if (precision == TD3_FP64) TD3_NEAR(TD3_LOG, TD3_FP64);
        else TD3_NEAR(TD3_LOG, TD3_MIXED);
    } else {
        if (precision == TD3_FP64) TD3_NEAR(TD3_COULOMB, TD3_FP64);
        else TD3_NEAR(TD3_COULOMB, TD3_MIXED);
    }
#undef TD3_NEAR
    TD3_LAUNCHED();
    return TD3_OK;
}

extern "C" int td3_fmm_far(int kernel, int64_t order, const double* sorted_pos,
                           const int64_t* perm, int64_t n, const double* node_center,
                           const double* node_half, const double* local_re,
                           const double* local_im, const int32_t* leaf_of, const double* near,
                           int with_potential, double* phi, double* field, void* stream) {
    TD3_ARG(kernel == TD3_COULOMB || kernel == TD3_LOG, "kernel");
    TD3_ARG(order >= 1 && order <= 64 && n >= 1, "order, sizes");
    hipStream_t s = (hipStream_t)stream;
    const dim3 grid(td3_blocks(n, 256));
    const int p1 = kernel == TD3_LOG ? (int)order + 1 : (int)((order + 1) * (order + 2) / 2);
    if (kernel == TD3_LOG)
        TD3_TIMED(s, "eval_far", hipLaunchKernelGGL(eval_far_kernel<TD3_LOG>, grid, dim3(256), 0, s,
                  sorted_pos, perm, n, node_center, node_half, local_re, local_im, p1,
                  (int)order, leaf_of, near, with_potential, phi, field));
    else
        TD3_TIMED(s, "eval_far", hipLaunchKernelGGL(eval_far_kernel<TD3_COULOMB>, grid, dim3(256), 0,
                  s, sorted_pos, perm, n, node_center, node_half, local_re, local_im, p1,
                  (int)order, leaf_of, near, with_potential, phi, field));
    TD3_LAUNCHED();
    return TD3_OK;
}

extern "C" int td3_fmm_at_eval(int kernel, int precision, int64_t order,
                               const double* sorted_targets, const int64_t* target_perm,
                               const int64_t* exclude, int64_t m, const double* t_center,
                               const double* t_half, const void* t_info, int64_t t_count,
                               const int32_t* t_leaf_of, const double* local_re,
                               const double* local_im, const uint32_t* p2p_row,
                               const int32_t* p2p_col, const double* sorted_pos,
                               const double* sorted_charge, const int64_t* perm, int64_t n,
                               const void* s_info, const void* s_sources4, double softening,
                               int with_potential, double* phi, double* field, int64_t* status,
                               void* stream) {
    TD3_ARG(kernel == TD3_COULOMB || kernel == TD3_LOG, "kernel");
    TD3_ARG(order >= 1 && order <= 64, "order");
    TD3_ARG(m >= 1 && n >= 1 && t_count >= 1, "sizes");
    auto go = kernel == TD3_LOG ? launch_eval_rows<TD3_LOG> : launch_eval_rows<TD3_COULOMB>;
    return go(precision, (hipStream_t)stream, sorted_targets, target_perm, exclude, m, t_center,
              t_half, (const int4*)t_info, t_count, local_re, local_im, (int)order, p2p_row,
              p2p_col, perm, (const int4*)s_info, (const double4*)s_sources4,
              softening * softening, with_potential, phi, field, status);
}

namespace {
template <int MODE>
int launch_ext(hipStream_t s, const double* tpos, const i64* tperm, const i64* exclude, i64 m,
               const double* tcenter, const double* thalf, const double* lre, const double* lim,
               int order, const u32* prow, const int* pcol, const i64* sperm, i64 n,
               const int4* sinfo, const int* leaf_of, const double4* src4, const double* sdip,
               double eps2, double* phi, double* field, double* jac, i64* status) {
    TD3_ARG(order >= 1 && order <= 64, "order");
    TD3_ARG(leaf_of != nullptr, "leaf_of (td3_leaf_index of the target tree)");
    TD3_ARG(src4 != nullptr, "sources4 (td3_pack_sources of the source tree)");
    TD3_ARG(sdip == nullptr || jac == nullptr, "dipole sources with a Jacobian");
    const unsigned blocks = td3_blocks(m, 256);
    const int p1 = (int)order + 1;
    if (sdip != nullptr)
        TD3_TIMED(s, "eval_ext", hipLaunchKernelGGL((log_ext_eval_kernel<MODE, 1>), dim3(blocks), dim3(256), 0, s,
                  tpos, tperm, exclude, m, tcenter, thalf, lre, lim, p1, prow, pcol, sperm, n,
                  sinfo, leaf_of, src4, sdip, eps2, phi, field, jac, status));
    else if (jac != nullptr)
        TD3_TIMED(s, "eval_ext", hipLaunchKernelGGL((log_ext_eval_kernel<MODE, 2>), dim3(blocks), dim3(256), 0, s,
                  tpos, tperm, exclude, m, tcenter, thalf, lre, lim, p1, prow, pcol, sperm, n,
                  sinfo, leaf_of, src4, sdip, eps2, phi, field, jac, status));
    else
        TD3_TIMED(s, "eval_ext", hipLaunchKernelGGL((log_ext_eval_kernel<MODE, 0>), dim3(blocks), dim3(256), 0, s,
                  tpos, tperm, exclude, m, tcenter, thalf, lre, lim, p1, prow, pcol, sperm, n,
                  sinfo, leaf_of, src4, sdip, eps2, phi, field, jac, status));
    TD3_LAUNCHED();
    return TD3_OK;
}
}  // namespace

extern "C" int td3_log_fmm_eval_ext(int64_t order, const double* sorted_pos, const int64_t* perm,
                                    int64_t n, const double* node_center,
                                    const double* node_half, const void* node_info,
                                    const double* local_re, const double* local_im,
                                    const uint32_t* p2p_row, const int32_t* p2p_col,
                                    const int32_t* leaf_of, const void* sources4,
                                    const double* sorted_dipole, double softening, double* phi,
                                    double* field, double* jacobian, int64_t* status,
                                    void* stream) {
    TD3_ARG(n >= 1, "sizes");
    return launch_ext<0>((hipStream_t)stream, sorted_pos, perm, nullptr, n, node_center,
                         node_half, local_re, local_im, (int)order, p2p_row, p2p_col, perm, n,
                         (const int4*)node_info, leaf_of, (const double4*)sources4,
                         sorted_dipole, softening * softening, phi, field, jacobian, status);
}

extern "C" int td3_log_fmm_at_eval_ext(int64_t order, const double* sorted_targets,
                                       const int64_t* target_perm, const int64_t* exclude,
                                       int64_t m, const double* t_center, const double* t_half,
                                       const void* t_info, int64_t t_count,
                                       const double* local_re, const double* local_im,
                                       const uint32_t* p2p_row, const int32_t* p2p_col,
                                       const int64_t* perm, int64_t n, const void* s_info,
                                       const void* s_sources4, const double* sorted_dipole,
                                       double softening, double* phi, double* field,
                                       double* jacobian, int64_t* status, void* stream) {
    TD3_ARG(m >= 1 && n >= 1 && t_count >= 1, "sizes");
    TD3_ARG(order >= 1 && order <= 64, "order");
    TD3_ARG(exclude != nullptr && t_info != nullptr, "exclude, t_info");
    TD3_ARG(s_sources4 != nullptr, "sources4 (td3_pack_sources of the source tree)");
    TD3_ARG(sorted_dipole == nullptr || jacobian == nullptr, "dipole sources with a Jacobian");
    hipStream_t s = (hipStream_t)stream;
    const unsigned blocks = td3_blocks(t_count * 64, 256);
    const int p1 = (int)order + 1;
    const double eps2 = softening * softening;
    const int4* ti = (const int4*)t_info;
    const int4* si = (const int4*)s_info;
    const double4* src4 = (const double4*)s_sources4;
#define TD3_EXT_ROWS(EXT)                                                                       \
    TD3_TIMED(s, "eval_ext", hipLaunchKernelGGL((log_ext_eval_rows_kernel<EXT>), dim3(blocks),  \
              dim3(256), 0, s, sorted_targets, target_perm, exclude, m, t_center, t_half, ti,    \
              t_count, local_re, local_im, p1, p2p_row, p2p_col, perm, n, si, src4,              \
              sorted_dipole, eps2, phi, field, jacobian, status))
    if (sorted_dipole != nullptr)
        TD3_EXT_ROWS(1);
    else if (jacobian != nullptr)
        TD3_EXT_ROWS(2);
    else
        TD3_EXT_ROWS(0);
#undef TD3_EXT_ROWS
    TD3_LAUNCHED();
    return TD3_OK;
}

extern "C" int td3_leaf_index(const void* node_info, int64_t node_count, int32_t* leaf_of,
                              void* stream) {
    TD3_ARG(node_count >= 1 && node_count < 0x7fffffff && leaf_of != nullptr, "sizes");
    hipLaunchKernelGGL(leaf_index_kernel, dim3(td3_blocks(node_count, 256)), dim3(256), 0,
                       (hipStream_t)stream, (const int4*)node_info, node_count, leaf_of);
    TD3_LAUNCHED();
    return TD3_OK;
}

extern "C" int td3_pack_sources(const double* sorted_pos, const double* sorted_charge, int64_t n,
                                void* sources4, void* stream) {
    TD3_ARG(n >= 0 && sources4 != nullptr, "arguments");
    if (n == 0) return TD3_OK;
    hipLaunchKernelGGL(pack_sources_kernel, dim3(td3_blocks(n, 256)), dim3(256), 0,
                       (hipStream_t)stream, sorted_pos, sorted_charge, n, (double4*)sources4);
    TD3_LAUNCHED();
    return TD3_OK;
}
