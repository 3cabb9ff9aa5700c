// Close-pair softening of line charges in the physical plane (the grounded strip's
// correction, strip.py:168-193 and :288-323).
//
// The reference finds the pairs closer than cutoff * softening with a k-d tree on the host
// and then adds, per pair, the softened-minus-bare line-charge kernel to both partners
// (kernels/soften_pairs/soften_pairs_cuda.cu:9-42, one thread per pair, four float
// atomics).  Here each target owns its sum: one lane per target walks a quadtree of the
// sources, stacklessly (child links down, parent links back up), visiting only cells whose
// box comes within the radius, and adds for every source within it
//   phi += -q ln((r^2 + eps^2) / r^2),   E += 2 q d (1 / (r^2 + eps^2) - 1 / r^2),
// d = target - source.  No pair list, no atomics, a fixed order per target.

#include "td3_common.h"

namespace {

__global__ void strip_parent_kernel(const int4* __restrict__ info, i64 count, int* __restrict__ parent) {
    const i64 node = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= count) return;
    if (node == 0) parent[0] = -1;
    const int4 f = info[node];
    if (f.z >= 0)
        for (int k = 0; k < 4; ++k) parent[f.z + k] = (int)node;
}

__global__ __launch_bounds__(256) void soften_near_kernel(
    const double* __restrict__ targets, i64 m, const i64* __restrict__ order,
    const i64* __restrict__ exclude, const double* __restrict__ spos,
    const double* __restrict__ scharge, const i64* __restrict__ perm, i64 n,
    const double* __restrict__ center, const double* __restrict__ half,
    const int4* __restrict__ info, const int* __restrict__ parent, double radius, double eps2,
    int with_potential, double* __restrict__ phi, double* __restrict__ field,
    i64* __restrict__ status) {
    const i64 w = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= m) return;
    const i64 t = order != nullptr ? order[w] : w;
    const double tx = targets[t], ty = targets[m + t];
    const i64 skip = exclude != nullptr ? exclude[t] : -1;
    const double r2max = radius * radius;
    // cells are only a pruning device (a small margin keeps every source at distance
    // <= radius); each source is decided by its own r^2 <= radius^2
    const double cell_r2 = r2max * (1.0 + 1e-12);
    double dp = 0.0, ex = 0.0, ey = 0.0;
    bool bad = false;
    int node = 0;
    while (true) {
        const int4 f = info[node];
        bool descend = false;
        if (f.y > f.x) {
            const double h = half[node];
            const double gx = fmax(fabs(tx - center[2 * node]) - h, 0.0);
            const double gy = fmax(fabs(ty - center[2 * node + 1]) - h, 0.0);
            if (gx * gx + gy * gy <= cell_r2) {
                if (f.z >= 0) {
                    descend = true;
                } else {
                    for (int k = f.x; k < f.y; ++k) {
                        if (perm[k] == skip) continue;
                        const double dx = tx - spos[k], dy = ty - spos[n + k];
                        const double r2 = dx * dx + dy * dy;
                        if (r2 > r2max) continue;
                        if (r2 == 0.0) {
                            bad = true;
                            continue;
                        }
                        const double q = scharge[k];
                        const double soft = r2 + eps2;
                        if (with_potential) dp += q * -log(soft / r2);
                        const double scale = 2.0 / soft - 2.0 / r2;
                        ex += q * scale * dx;
                        ey += q * scale * dy;
                    }
                }
            }
        }
        if (descend) {
            node = f.z;
            continue;
        }
        while (node != 0 && ((node - 1) & 3) == 3) node = parent[node];
        if (node == 0) break;
        ++node;
    }
    if (bad) td3_report(status, t + 1);
    if (with_potential) phi[t] += dp;
    field[t] += ex;
    field[m + t] += ey;
}

}  // namespace

extern "C" size_t td3_soften_near_workspace_bytes(int64_t node_count) {
    return sizeof(int) * (size_t)(node_count > 0 ? node_count : 1);
}

extern "C" int td3_soften_near(const double* targets, int64_t m, const int64_t* order,
                               const int64_t* exclude, const double* sorted_pos,
                               const double* sorted_charge, const int64_t* perm, int64_t n,
                               const double* node_center, const double* node_half,
                               const void* node_info, int64_t node_count, double radius,
                               double softening, int with_potential, double* phi, double* field,
                               int64_t* status, void* workspace, size_t workspace_bytes,
                               void* stream) {
    TD3_ARG(m >= 0 && n >= 0 && node_count >= 1, "sizes");
    TD3_ARG(radius >= 0.0 && softening >= 0.0, "radius and softening");
    TD3_ARG(workspace_bytes >= td3_soften_near_workspace_bytes(node_count) && workspace != nullptr,
            "workspace too small");
    if (m == 0 || n == 0 || radius == 0.0 || softening == 0.0) return TD3_OK;
    hipStream_t s = (hipStream_t)stream;
    int* parent = static_cast<int*>(workspace);
    hipLaunchKernelGGL(strip_parent_kernel, dim3(td3_blocks(node_count, 256)), dim3(256), 0, s,
                       (const int4*)node_info, node_count, parent);
    TD3_LAUNCHED();
    TD3_TIMED(s, "soften_near",
              hipLaunchKernelGGL(soften_near_kernel, dim3(td3_blocks(m, 256)), dim3(256), 0, s,
                                 targets, m, order, exclude, sorted_pos, sorted_charge, perm, n,
                                 node_center, node_half, (const int4*)node_info, parent, radius,
                                 softening * softening, with_potential, phi, field, status));
    TD3_LAUNCHED();
    return TD3_OK;
}
