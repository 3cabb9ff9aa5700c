// Direct all-pairs sums (the FP64 oracle of every benchmark configuration and the
// small-system path of the simulation).
//
// Replaces kernels/direct_sum/direct_sum_cuda.cu:7-44 and
// kernels/direct_log_sum/direct_log_sum_cuda.cu:7-42 (one thread per target, each
// streaming every source from global memory).  Here a 256-thread workgroup owns 256
// targets and one contiguous chunk of the sources; the chunk is streamed through LDS in
// tiles loaded with coalesced component-major reads, so each source is read from HBM
// once per workgroup rather than once per target.  When the targets alone cannot fill
// the 256 CUs, the sources are split into chunks along grid.y and a second kernel adds
// the chunk partials in a fixed order (deterministic, no float atomics).

#include <mutex>

#include "td3_common.h"

// Scratch of td3_direct_noalloc: the per-chunk partial sums (phi, E_x, E_y of every target
// and source chunk) that td3_direct keeps in its caller's workspace.  One copy per GPU,
// held until the process ends; every call on a GPU shares its copy, so calls that overlap
// on one GPU race.  Its host-side name is not a device address: kernels get the address
// that hipGetSymbolAddress returns for the current device.
constexpr size_t kDirectScratchBytes = size_t(16) << 20;
static __device__ __align__(256) double scratch[kDirectScratchBytes / sizeof(double)];

namespace {

constexpr int kThreads = 256;

template <int KERNEL, int PREC>
__global__ __launch_bounds__(kThreads) void direct_tile_kernel(
    const double* __restrict__ targets, i64 m, const i64* __restrict__ exclude,
    const double* __restrict__ pos, i64 n, const double* __restrict__ charge, double eps2,
    i64 chunk, double* __restrict__ out_phi, double* __restrict__ out_field, i64 ld_out,
    i64* __restrict__ status) {
    __shared__ double sx[kThreads], sy[kThreads], sq[kThreads];
    const i64 i = (i64)blockIdx.x * kThreads + threadIdx.x;
    const bool active = i < m;
    const double tx = active ? targets[i] : 0.0;
    const double ty = active ? targets[m + i] : 0.0;
    const i64 skip = active ? exclude[i] : -1;
    const i64 j0 = (i64)blockIdx.y * chunk;
    const i64 j1 = (j0 + chunk < n) ? j0 + chunk : n;
    double potential = 0.0, ex = 0.0, ey = 0.0;
    bool bad = false;
    for (i64 base = j0; base < j1; base += kThreads) {
        const i64 j = base + threadIdx.x;
        __syncthreads();
        if (j < j1) {
            sx[threadIdx.x] = pos[j];
            sy[threadIdx.x] = pos[n + j];
            sq[threadIdx.x] = charge[j];
        } else {
            sx[threadIdx.x] = 0.0;
            sy[threadIdx.x] = 0.0;
            sq[threadIdx.x] = 0.0;
        }
        __syncthreads();
        const int count = (int)((j1 - base) < kThreads ? (j1 - base) : kThreads);
        for (int t = 0; t < count; ++t) {
            const double c = sq[t];
            if (base + t == skip || c == 0.0) continue;
            const double dx = tx - sx[t];
            const double dy = ty - sy[t];
            const double r2 = dx * dx + dy * dy + eps2;
            if (r2 == 0.0) {
                bad = true;
                continue;
            }
            if (PREC == TD3_FP64) {
                if (KERNEL == TD3_LOG) {
                    potential -= c * log(r2);
                    const double inv = 2.0 * c / r2;
                    ex += inv * dx;
                    ey += inv * dy;
                } else {
                    const double inv_r = 1.0 / sqrt(r2);
                    const double cr3 = c * inv_r * inv_r * inv_r;
                    potential += c * inv_r;
                    ex += cr3 * dx;
                    ey += cr3 * dy;
                }
            } else {
                // FP32 transcendentals and reciprocals on FP64 differences, FP64 sums.
                const float r2f = (float)r2;
                const float cf = (float)c;
                if (KERNEL == TD3_LOG) {
                    potential -= (double)(cf * __logf(r2f));
                    const float inv = 2.0f * cf * __frcp_rn(r2f);
                    ex += (double)(inv * (float)dx);
                    ey += (double)(inv * (float)dy);
                } else {
                    const float inv_r = __frsqrt_rn(r2f);
                    const float cr3 = cf * inv_r * inv_r * inv_r;
                    potential += (double)(cf * inv_r);
                    ex += (double)(cr3 * (float)dx);
                    ey += (double)(cr3 * (float)dy);
                }
            }
        }
    }
    if (!active) return;
    if (bad) td3_report(status, i + 1);
    const i64 o = (i64)blockIdx.y * ld_out;
    out_phi[o + i] = potential;
    out_field[2 * o + i] = ex;
    out_field[2 * o + m + i] = ey;
}

// Adds the chunk partials of each target in chunk order.
__global__ void direct_reduce_kernel(const double* __restrict__ part_phi,
                                     const double* __restrict__ part_field, i64 m, int chunks,
                                     double* __restrict__ phi, double* __restrict__ field) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double p = 0.0, ex = 0.0, ey = 0.0;
    for (int c = 0; c < chunks; ++c) {
        p += part_phi[(i64)c * m + i];
        ex += part_field[2 * (i64)c * m + i];
        ey += part_field[2 * (i64)c * m + m + i];
    }
    phi[i] = p;
    field[i] = ex;
    field[m + i] = ey;
}

int direct_chunks(i64 m, i64 n) {
    const i64 target_blocks = (m + kThreads - 1) / kThreads;
    i64 chunks = (2048 + target_blocks - 1) / target_blocks;
    const i64 max_chunks = (n + 4 * kThreads - 1) / (4 * kThreads);
    if (chunks > max_chunks) chunks = max_chunks;
    if (chunks > 1024) chunks = 1024;
    if (chunks < 1) chunks = 1;
    return (int)chunks;
}

template <int KERNEL, int PREC>
void launch_direct(dim3 grid, hipStream_t s, const double* targets, i64 m, const i64* exclude,
                   const double* pos, i64 n, const double* charge, double eps2, i64 chunk,
                   double* out_phi, double* out_field, i64 ld, i64* status) {
    TD3_TIMED(s, "direct", hipLaunchKernelGGL((direct_tile_kernel<KERNEL, PREC>), grid, dim3(kThreads), 0, s, targets, m,
                       exclude, pos, n, charge, eps2, chunk, out_phi, out_field, ld, status));
}

}  // namespace

extern "C" size_t td3_direct_workspace_bytes(int64_t m, int64_t n) {
    if (m <= 0 || n <= 0) return 0;
    const int chunks = direct_chunks(m, n);
    if (chunks <= 1) return 0;
    return (size_t)chunks * (size_t)m * 3 * sizeof(double);
}

extern "C" int td3_direct(int kernel, int precision, const double* targets, int64_t m,
                          const int64_t* exclude, const double* pos, int64_t n,
                          const double* charge, double softening, double* phi, double* field,
                          int64_t* status, void* workspace, size_t workspace_bytes,
                          void* stream) {
    TD3_ARG(kernel == TD3_COULOMB || kernel == TD3_LOG, "kernel");
    TD3_ARG(precision == TD3_FP64 || precision == TD3_MIXED, "precision");
    TD3_ARG(m >= 0 && n >= 0, "sizes");
    hipStream_t s = (hipStream_t)stream;
    if (m == 0) return TD3_OK;
    if (n == 0) {
        TD3_HIP(hipMemsetAsync(phi, 0, sizeof(double) * (size_t)m, s));
        TD3_HIP(hipMemsetAsync(field, 0, 2 * sizeof(double) * (size_t)m, s));
        return TD3_OK;
    }
    const int chunks = direct_chunks(m, n);
    const size_t need = td3_direct_workspace_bytes(m, n);
    TD3_ARG(workspace_bytes >= need && (need == 0 || workspace != nullptr), "workspace too small");
    i64 chunk = (n + chunks - 1) / chunks;
    chunk = ((chunk + kThreads - 1) / kThreads) * kThreads;
    const double eps2 = softening * softening;
    dim3 grid(td3_blocks(m, kThreads), (unsigned)chunks);
    double* out_phi = phi;
    double* out_field = field;
    if (chunks > 1) {
        out_phi = static_cast<double*>(workspace);
        out_field = out_phi + (size_t)chunks * (size_t)m;
    }
    const i64 ld = chunks > 1 ? m : 0;
    if (kernel == TD3_LOG) {
        if (precision == TD3_FP64)
            launch_direct<TD3_LOG, TD3_FP64>(grid, s, targets, m, exclude, pos, n, charge, eps2, chunk,
                                             out_phi, out_field, ld, status);
        else
            launch_direct<TD3_LOG, TD3_MIXED>(grid, s, targets, m, exclude, pos, n, charge, eps2,
                                              chunk, out_phi, out_field, ld, status);
    } else {
        if (precision == TD3_FP64)
            launch_direct<TD3_COULOMB, TD3_FP64>(grid, s, targets, m, exclude, pos, n, charge, eps2,
                                                 chunk, out_phi, out_field, ld, status);
        else
            launch_direct<TD3_COULOMB, TD3_MIXED>(grid, s, targets, m, exclude, pos, n, charge,
                                                  eps2, chunk, out_phi, out_field, ld, status);
    }
    TD3_LAUNCHED();
    if (chunks > 1) {
        TD3_TIMED(s, "direct_reduce", hipLaunchKernelGGL(direct_reduce_kernel, dim3(td3_blocks(m, 256)), dim3(256), 0, s, out_phi,
                           out_field, m, chunks, phi, field));
        TD3_LAUNCHED();
    }
    return TD3_OK;
}

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per device.
int direct_scratch(void** out) {
    int dev = 0;
    TD3_HIP(hipGetDevice(&dev));
    TD3_ARG(dev >= 0 && dev < kMaxDevices, "device index");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        TD3_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return TD3_OK;
}
}  // namespace

extern "C" size_t td3_direct_noalloc_scratch_bytes(void) { return kDirectScratchBytes; }

extern "C" int td3_direct_noalloc(int kernel, int precision, const double* targets, int64_t m,
                                  const int64_t* exclude, const double* pos, int64_t n,
                                  const double* charge, double softening, double* phi,
                                  double* field, int64_t* status, void* stream) {
    TD3_ARG(m >= 0 && n >= 0, "sizes");
    const size_t need = td3_direct_workspace_bytes(m, n);
    if (need > kDirectScratchBytes) {
        td3_set_error("td3_direct_noalloc: %lld targets x %lld sources need %zu bytes of "
                      "scratch, more than its %zu (call td3_direct with a workspace)",
                      (long long)m, (long long)n, need, kDirectScratchBytes);
        return TD3_ERR_CAPACITY;
    }
    void* ws = nullptr;
    if (need > 0) {
        const int code = direct_scratch(&ws);
        if (code != TD3_OK) return code;
    }
    // the chunk partials are written in full before they are read: nothing to clear
    return td3_direct(kernel, precision, targets, m, exclude, pos, n, charge, softening, phi,
                      field, status, ws, need, stream);
}
