// Particle push of the electrostatic simulation (the reference's simulation.py:800-835):
// electric half-kick, exact drift and rotation in a uniform B_z, electric half-kick.
// One thread per particle, FP64.

#include "td3_common.h"

namespace {

// sin(x) / x, exact at 0
__device__ __forceinline__ double sinc(double x) { return x == 0.0 ? 1.0 : sin(x) / x; }

__global__ void push_kernel(int first_half, i64 n, double* __restrict__ pos,
                            double* __restrict__ vel, double* __restrict__ force,
                            const double* __restrict__ electric, const double* __restrict__ mass,
                            const double* __restrict__ charge, const uint8_t* __restrict__ fixed,
                            double dt, double bz) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool mobile = !fixed[i];
    double vx = vel[i], vy = vel[n + i];
    if (mobile) {
        // (0.5 dt E) (1/m), the rounding order of the host reference
        const double inv_mass = 1.0 / mass[i];
        vx += 0.5 * dt * electric[i] * inv_mass;
        vy += 0.5 * dt * electric[n + i] * inv_mass;
    }
    if (first_half) {
        if (mobile) {
            const double angle = charge[i] / mass[i] * bz * dt;
            const double c = cos(angle), s = sin(angle);
            const double a = sinc(angle);
            const double half = sinc(0.5 * angle);
            const double b = 0.5 * angle * half * half;
            pos[i] += dt * (a * vx + b * vy);
            pos[n + i] += dt * (a * vy - b * vx);
            const double nvx = c * vx + s * vy;
            vy = c * vy - s * vx;
            vx = nvx;
        }
    } else {
        force[i] = electric[i] + charge[i] * bz * vy;
        force[n + i] = electric[n + i] - charge[i] * bz * vx;
    }
    vel[i] = vx;
    vel[n + i] = vy;
}

// One block per system: the bounding box of its particles and, when they leave its root
// square, the new root (simulation.py:625-639): the square around the box and the old
// root, centre lo/2 + hi/2, half max(hi/2 - lo/2) (1 + 1e-12).
__global__ __launch_bounds__(256) void system_bounds_kernel(const double* __restrict__ pos, i64 n,
                                                            const i64* __restrict__ start,
                                                            double* __restrict__ roots,
                                                            u32* __restrict__ grown) {
    __shared__ double s_v[4][256 / TD3_WAVE];
    const i64 sys = blockIdx.x;
    const i64 a = start[sys], b = start[sys + 1];
    double lo_x = INFINITY, lo_y = INFINITY, hi_x = -INFINITY, hi_y = -INFINITY;
    for (i64 i = a + threadIdx.x; i < b; i += blockDim.x) {
        const double x = pos[i], y = pos[n + i];
        lo_x = fmin(lo_x, x);
        hi_x = fmax(hi_x, x);
        lo_y = fmin(lo_y, y);
        hi_y = fmax(hi_y, y);
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo_x = fmin(lo_x, __shfl_xor(lo_x, off, 64));
        hi_x = fmax(hi_x, __shfl_xor(hi_x, off, 64));
        lo_y = fmin(lo_y, __shfl_xor(lo_y, off, 64));
        hi_y = fmax(hi_y, __shfl_xor(hi_y, off, 64));
    }
    const int wave = threadIdx.x / TD3_WAVE;
    if (td3_lane() == 0) {
        s_v[0][wave] = lo_x;
        s_v[1][wave] = hi_x;
        s_v[2][wave] = lo_y;
        s_v[3][wave] = hi_y;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int w = 1; w < (int)(blockDim.x / TD3_WAVE); ++w) {
        lo_x = fmin(lo_x, s_v[0][w]);
        hi_x = fmax(hi_x, s_v[1][w]);
        lo_y = fmin(lo_y, s_v[2][w]);
        hi_y = fmax(hi_y, s_v[3][w]);
    }
    const double cx = roots[3 * sys], cy = roots[3 * sys + 1], h = roots[3 * sys + 2];
    const double rlo_x = cx - h, rlo_y = cy - h, rhi_x = cx + h, rhi_y = cy + h;
    if (!(lo_x < rlo_x || lo_y < rlo_y || hi_x > rhi_x || hi_y > rhi_y)) return;
    lo_x = fmin(lo_x, rlo_x);
    lo_y = fmin(lo_y, rlo_y);
    hi_x = fmax(hi_x, rhi_x);
    hi_y = fmax(hi_y, rhi_y);
    const double wx = hi_x / 2 - lo_x / 2, wy = hi_y / 2 - lo_y / 2;
    roots[3 * sys] = lo_x / 2 + hi_x / 2;
    roots[3 * sys + 1] = lo_y / 2 + hi_y / 2;
    roots[3 * sys + 2] = (wx > wy ? wx : wy) * (1 + 1e-12);
    atomicAdd(grown, 1u);
}

}  // namespace

extern "C" int td3_system_bounds(const double* positions, int64_t n, int64_t systems,
                                 const int64_t* system_start, double* roots, uint32_t* grown,
                                 void* stream) {
    TD3_ARG(n >= 0 && systems >= 1 && system_start != nullptr && roots != nullptr &&
                grown != nullptr, "arguments");
    hipLaunchKernelGGL(system_bounds_kernel, dim3((unsigned)systems), dim3(256), 0,
                       (hipStream_t)stream, positions, n, system_start, roots, grown);
    TD3_LAUNCHED();
    return TD3_OK;
}

extern "C" int td3_push(int first_half, int64_t n, double* positions, double* velocities,
                        double* forces, const double* electric, const double* masses,
                        const double* charges, const uint8_t* fixed, double dt, double b_z,
                        void* stream) {
    TD3_ARG(n >= 0, "n");
    if (n == 0) return TD3_OK;
    hipLaunchKernelGGL(push_kernel, dim3(td3_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream,
                       first_half, n, positions, velocities, forces, electric, masses, charges,
                       fixed, dt, b_z);
    TD3_LAUNCHED();
    return TD3_OK;
}

// ------------------------------------------------------------------ energies per system
namespace {

constexpr int kEnergyThreads = 256;

// One workgroup per system sums its contiguous particles: thread t takes particles
// start + t, start + t + 256, ... in order, then a fixed tree over the 256 partial sums,
// so every run adds the same numbers in the same order (no float atomics).
// out[s] = sum 1/2 m |v|^2, out[S + s] = sum 1/2 q phi.
__global__ __launch_bounds__(kEnergyThreads) void system_energies_kernel(
    const double* __restrict__ mass, const double* __restrict__ vel,
    const double* __restrict__ charge, const double* __restrict__ potential, i64 n,
    i64 systems, const i64* __restrict__ system_start, double* __restrict__ out) {
    __shared__ double s_kin[kEnergyThreads], s_pot[kEnergyThreads];
    const i64 s = blockIdx.x;
    const i64 lo = system_start[s], hi = system_start[s + 1];
    double kin = 0.0, pot = 0.0;
    for (i64 i = lo + threadIdx.x; i < hi; i += kEnergyThreads) {
        const double vx = vel[i], vy = vel[n + i];
        kin += 0.5 * mass[i] * (vx * vx + vy * vy);
        if (potential != nullptr) pot += 0.5 * charge[i] * potential[i];
    }
    s_kin[threadIdx.x] = kin;
    s_pot[threadIdx.x] = pot;
    __syncthreads();
    for (int half = kEnergyThreads / 2; half > 0; half >>= 1) {
        if ((int)threadIdx.x < half) {
            s_kin[threadIdx.x] += s_kin[threadIdx.x + half];
            s_pot[threadIdx.x] += s_pot[threadIdx.x + half];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[s] = s_kin[0];
        out[systems + s] = s_pot[0];
    }
}

}  // namespace

extern "C" int td3_system_energies(const double* masses, const double* velocities,
                                   const double* charges, const double* potential, int64_t n,
                                   int64_t systems, const int64_t* system_start, double* out,
                                   void* stream) {
    TD3_ARG(n >= 0 && systems > 0, "sizes");
    hipStream_t s = (hipStream_t)stream;
    TD3_TIMED(s, "system_energies",
              hipLaunchKernelGGL(system_energies_kernel, dim3((unsigned)systems),
                                 dim3(kEnergyThreads), 0, s, masses, velocities, charges,
                                 potential, n, systems, system_start, out));
    TD3_LAUNCHED();
    return TD3_OK;
}
