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

}  // namespace

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
