// On-device reductions for run observers: the field energy and the block-mean
// snapshot of a field (etalon/studio/runner.py:258-311), so that a progress report
// copies three doubles or a 256 x 256 snapshot to the host instead of whole fields.
//
// Field energy: the three float64 sums the reference forms with xp.sum —
//   S_ex = sum over Ex nodes of ex^2 / cb_ex,  S_ey = ... ey^2 / cb_ey,  S_hz = sum hz^2
// (each term (double(v) * double(v)) / double(cb), the reference's operation order) —
// reduced in a fixed order, so a call's result depends on its inputs only:
//   pass 1: block b of array a sums its elements [4096 b, 4096 (b + 1)); thread t adds
//           elements 4096 b + t + 256 k for k = 0..15 in that order, then the block's
//           256 thread sums are added in a binary tree (stride 128, 64, ..., 1);
//   pass 2: per array, thread t adds partials t, t + 256, ... in that order, then the
//           same tree over the 256 threads.
// The NumPy reference sums pairwise in another order: the results agree to float64
// rounding.  The partial sums are the scratch buffer: etl_field_energy takes them from
// the library's allocator, etl_field_energy_noalloc from a fixed device array.
#include <algorithm>

#include "common.h"

namespace etl {
namespace {

constexpr int kEnergyThreads = 256;
constexpr int kEnergyPerThread = 16;
constexpr int64_t kEnergyPerBlock = int64_t(kEnergyThreads) * kEnergyPerThread;
// The no-alloc variant's scratch: 3 arrays x this many block partials (1.5 MiB per
// device, held until the process ends), so each array may hold up to
// kEnergyNoallocBlocks * 4096 = 268435456 nodes.
constexpr int kEnergyNoallocBlocks = 65536;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[3 * kEnergyNoallocBlocks];
// Per device: the device address of `scratch` (the host-side name is only a stand-in).
double* g_scratch_ptr[kMaxDevices] = {nullptr};

template <typename R>
struct EnergyTerm {
    const R* f;
    const R* w;  // null: f^2, else f^2 / w
    int64_t n;
};

__device__ __forceinline__ double block_tree(double acc, double* lds) {
    lds[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kEnergyThreads / 2; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) lds[threadIdx.x] = lds[threadIdx.x] + lds[threadIdx.x + s];
        __syncthreads();
    }
    return lds[0];
}

template <typename R>
__global__ __launch_bounds__(kEnergyThreads) void energy_partial_kernel(
    EnergyTerm<R> t0, EnergyTerm<R> t1, EnergyTerm<R> t2, double* __restrict__ partials,
    int stride) {
    const EnergyTerm<R> t = blockIdx.y == 0 ? t0 : (blockIdx.y == 1 ? t1 : t2);
    const int64_t base = int64_t(blockIdx.x) * kEnergyPerBlock;
    if (base >= t.n) return;  // block-uniform: this array has fewer blocks
    double acc = 0.0;
    for (int k = 0; k < kEnergyPerThread; ++k) {
        const int64_t i = base + threadIdx.x + int64_t(k) * kEnergyThreads;
        if (i < t.n) {
            const double v = double(t.f[i]);
            double term = v * v;
            if (t.w != nullptr) term = term / double(t.w[i]);
            acc = acc + term;
        }
    }
    __shared__ double lds[kEnergyThreads];
    const double total = block_tree(acc, lds);
    if (threadIdx.x == 0) partials[int64_t(blockIdx.y) * stride + blockIdx.x] = total;
}

__global__ __launch_bounds__(kEnergyThreads) void energy_final_kernel(
    const double* __restrict__ partials, int stride, int n0, int n1, int n2,
    double* __restrict__ out) {
    const int n = blockIdx.x == 0 ? n0 : (blockIdx.x == 1 ? n1 : n2);
    const double* p = partials + int64_t(blockIdx.x) * stride;
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += kEnergyThreads) acc = acc + p[i];
    __shared__ double lds[kEnergyThreads];
    const double total = block_tree(acc, lds);
    if (threadIdx.x == 0) out[blockIdx.x] = total;
}

int blocks_of(int64_t n) { return int((n + kEnergyPerBlock - 1) / kEnergyPerBlock); }

// Both passes on the library's stream; `partials` holds 3 * (the largest block count).
template <typename R>
int launch_energy(const void* ex, const void* ey, const void* hz, const void* cb_ex,
                  const void* cb_ey, int nx, int ny, double* partials, double* out) {
    const EnergyTerm<R> t[3] = {
        {static_cast<const R*>(ex), static_cast<const R*>(cb_ex), int64_t(nx) * (ny + 1)},
        {static_cast<const R*>(ey), static_cast<const R*>(cb_ey), int64_t(nx + 1) * ny},
        {static_cast<const R*>(hz), nullptr, int64_t(nx) * ny}};
    const int nb[3] = {blocks_of(t[0].n), blocks_of(t[1].n), blocks_of(t[2].n)};
    const int stride = std::max(nb[0], std::max(nb[1], nb[2]));
    hipStream_t s = ctx().stream;
    hipLaunchKernelGGL(energy_partial_kernel<R>, dim3(unsigned(stride), 3), dim3(kEnergyThreads),
                       0, s, t[0], t[1], t[2], partials, stride);
    ETL_LAUNCHED();
    hipLaunchKernelGGL(energy_final_kernel, dim3(3), dim3(kEnergyThreads), 0, s, partials, stride,
                       nb[0], nb[1], nb[2], out);
    ETL_LAUNCHED();
    return ETL_OK;
}

int check_energy_args(int precision, const void* ex, const void* ey, const void* hz,
                      const void* cb_ex, const void* cb_ey, int nx, int ny, double* out,
                      const char* who) {
    ETL_REQUIRE(precision == ETL_F32 || precision == ETL_F64, "%s: precision must be 0 or 1", who);
    ETL_REQUIRE(ex && ey && hz && cb_ex && cb_ey && out, "%s: null pointer", who);
    ETL_REQUIRE(nx >= 1 && ny >= 1 && nx < (1 << 24) && ny < (1 << 24), "%s: grid %d x %d", who,
                nx, ny);
    return ETL_OK;
}

template <typename R>
__global__ void downsample_kernel(const R* __restrict__ f, int ny, int sx, int sy, int mx, int my,
                                  float* __restrict__ out) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= int64_t(mx) * my) return;
    const int a = int(k / my);
    const int b = int(k - int64_t(a) * my);
    double acc = 0.0;
    for (int p = 0; p < sx; ++p) {
        const R* row = f + (int64_t(a) * sx + p) * ny + int64_t(b) * sy;
        for (int q = 0; q < sy; ++q) acc = acc + double(row[q]);
    }
    out[k] = float(acc / double(sx * sy));
}

}  // namespace

// Called by etl_init with `device` current: the device address of this file's scratch
// on that device (each device holds its own copy of the array).
int observe_setup(int device) {
    if (device < 0 || device >= kMaxDevices) {
        set_error("etl_init: device %d is beyond the %d the observer scratch supports", device,
                  kMaxDevices);
        return ETL_EARG;
    }
    if (g_scratch_ptr[device] != nullptr) return ETL_OK;
    void* p = nullptr;
    ETL_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = static_cast<double*>(p);
    return ETL_OK;
}

}  // namespace etl

using namespace etl;

extern "C" {

int etl_field_energy(int precision, const void* ex, const void* ey, const void* hz,
                     const void* cb_ex, const void* cb_ey, int nx, int ny, double* out) {
    ETL_CHECK_INIT();
    if (int rc = check_energy_args(precision, ex, ey, hz, cb_ex, cb_ey, nx, ny, out,
                                   "etl_field_energy"))
        return rc;
    const int stride = blocks_of(int64_t(nx + 1) * (ny + 1));  // >= every array's blocks
    void* partials = nullptr;
    if (int rc = etl_malloc(&partials, size_t(3) * stride * sizeof(double))) return rc;
    const int rc = precision == ETL_F32
                       ? launch_energy<float>(ex, ey, hz, cb_ex, cb_ey, nx, ny,
                                              static_cast<double*>(partials), out)
                       : launch_energy<double>(ex, ey, hz, cb_ex, cb_ey, nx, ny,
                                               static_cast<double*>(partials), out);
    // stream-ordered: a later allocation of the block is used after these kernels
    const int rc_free = etl_free(partials);
    return rc ? rc : rc_free;
}

int etl_field_energy_noalloc(int precision, const void* ex, const void* ey, const void* hz,
                             const void* cb_ex, const void* cb_ey, int nx, int ny, double* out) {
    ETL_CHECK_INIT();
    if (int rc = check_energy_args(precision, ex, ey, hz, cb_ex, cb_ey, nx, ny, out,
                                   "etl_field_energy_noalloc"))
        return rc;
    const int64_t largest = std::max(int64_t(nx) * (ny + 1), int64_t(nx + 1) * ny);
    ETL_REQUIRE(blocks_of(largest) <= kEnergyNoallocBlocks,
                "etl_field_energy_noalloc: %lld nodes in one array, at most %lld (its fixed "
                "scratch holds %d block sums per array); use etl_field_energy",
                (long long)largest, (long long)kEnergyNoallocBlocks * kEnergyPerBlock,
                kEnergyNoallocBlocks);
    double* partials = g_scratch_ptr[ctx().device];
    ETL_REQUIRE(partials != nullptr, "etl_field_energy_noalloc: no scratch address for device %d",
                ctx().device);
    // Every partial that pass 2 reads is written by pass 1 of the same call, so the
    // previous call's contents need no clearing.
    return precision == ETL_F32
               ? launch_energy<float>(ex, ey, hz, cb_ex, cb_ey, nx, ny, partials, out)
               : launch_energy<double>(ex, ey, hz, cb_ex, cb_ey, nx, ny, partials, out);
}

int etl_downsample(int precision, const void* field, int nx, int ny, int sx, int sy, float* out) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(precision == ETL_F32 || precision == ETL_F64, "etl_downsample: precision");
    ETL_REQUIRE(field && out, "etl_downsample: null pointer");
    ETL_REQUIRE(nx >= 1 && ny >= 1 && sx >= 1 && sy >= 1 && sx <= nx && sy <= ny,
                "etl_downsample: field %d x %d, blocks %d x %d", nx, ny, sx, sy);
    const int mx = nx / sx, my = ny / sy;
    const int64_t n = int64_t(mx) * my;
    const unsigned blocks = unsigned((n + 255) / 256);
    if (precision == ETL_F32)
        hipLaunchKernelGGL(downsample_kernel<float>, dim3(blocks), dim3(256), 0, ctx().stream,
                           static_cast<const float*>(field), ny, sx, sy, mx, my, out);
    else
        hipLaunchKernelGGL(downsample_kernel<double>, dim3(blocks), dim3(256), 0, ctx().stream,
                           static_cast<const double*>(field), ny, sx, sy, mx, my, out);
    ETL_LAUNCHED();
    return ETL_OK;
}

}  // extern "C"
