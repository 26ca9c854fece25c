// Device selection, errors and memory of libetalon_hip.so.
//
// Memory: a small caching allocator keyed by exact (rounded) size, so that the
// per-step host uploads of the backend protocol (one tiny delta array per source
// per step, etalon/fdtd/solver.py:87) do not call hipMalloc/hipFree each step
// (hipFree synchronises the device).  Every block carries kGuard bytes of valid
// device memory before and after the array, so the fused kernels can issue whole
// vector loads that overhang a row by a few elements without branching.
#include <algorithm>
#include <cstdarg>
#include <map>
#include <mutex>
#include <unordered_map>

#include "common.h"

namespace etl {

static thread_local std::string g_error;

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
}

Context& ctx() {
    static Context c;
    return c;
}

int sync_lanes() {
    for (int k = 0; k < kLanes; ++k)
        if (ctx().lane[k]) ETL_HIP(hipStreamSynchronize(ctx().lane[k]));
    return ETL_OK;
}

int join_lanes() {
    for (int k = 0; k < kLanes; ++k) ETL_HIP(hipEventRecord(ctx().lane_ev[k], ctx().lane[k]));
    for (int k = 0; k < kLanes; ++k)
        for (int m = 0; m < kLanes; ++m)
            if (m != k) ETL_HIP(hipStreamWaitEvent(ctx().lane[k], ctx().lane_ev[m], 0));
    return ETL_OK;
}

int wait_other_lanes(int k) {
    for (int m = 0; m < kLanes; ++m)
        if (m != k) {
            ETL_HIP(hipEventRecord(ctx().lane_ev[m], ctx().lane[m]));
            ETL_HIP(hipStreamWaitEvent(ctx().lane[k], ctx().lane_ev[m], 0));
        }
    return ETL_OK;
}

int require_init() {
    if (ctx().stream == nullptr) {
        set_error("libetalon_hip: etl_init() has not been called");
        return ETL_ESTATE;
    }
    return ETL_OK;
}

namespace {
std::mutex g_pool_mutex;
struct Block {
    void* base;    // start of the hipMalloc block
    size_t size;   // size class (bytes usable between the guards)
    size_t bytes;  // bytes asked for
};
std::multimap<size_t, void*> g_free;           // size class -> base, cached for reuse
std::unordered_map<void*, Block> g_live;       // user pointer (base + kGuard) -> block
size_t g_cached_bytes = 0;
constexpr size_t kCacheLimit = size_t(1) << 30;  // keep at most 1 GiB of freed blocks

size_t round_size(size_t bytes) {
    if (bytes <= 256) return 256;
    if (bytes <= (size_t(1) << 20)) {  // powers of two up to 1 MiB
        size_t s = 256;
        while (s < bytes) s <<= 1;
        return s;
    }
    return (bytes + 65535) & ~size_t(65535);  // 64 KiB granules above
}

__global__ void fill_zero_kernel(uint32_t* __restrict__ p, size_t n_words) {
    size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (; k < n_words; k += stride) p[k] = 0u;
}

// Streaming copy for measuring the device's copy rate: 16-byte elements, grid-stride.
__global__ void stream_copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst,
                                   size_t n) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n; k += stride)
        dst[k] = src[k];
}

__global__ void fill_zero_bytes_kernel(uint8_t* __restrict__ p, size_t n) {
    size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0;
}
}  // namespace

size_t guarded_bytes(const void* p) {
    std::lock_guard<std::mutex> lock(g_pool_mutex);
    auto it = g_live.find(const_cast<void*>(p));
    return it == g_live.end() ? 0 : it->second.bytes;
}

int zero_async(void* dst, size_t bytes) {
    if (bytes == 0) return ETL_OK;
    uint8_t* p = static_cast<uint8_t*>(dst);
    // Word part (blocks are 256-byte aligned; a tail of < 4 bytes is cleared bytewise).
    size_t words = bytes / 4;
    if (words) {
        size_t blocks = (words + 255) / 256;
        if (blocks > 65536) blocks = 65536;
        hipLaunchKernelGGL(fill_zero_kernel, dim3(unsigned(blocks)), dim3(256), 0, ctx().stream,
                           reinterpret_cast<uint32_t*>(p), words);
        ETL_LAUNCHED();
    }
    size_t tail = bytes - words * 4;
    if (tail) {
        hipLaunchKernelGGL(fill_zero_bytes_kernel, dim3(1), dim3(64), 0, ctx().stream,
                           p + words * 4, tail);
        ETL_LAUNCHED();
    }
    return ETL_OK;
}

}  // namespace etl

using namespace etl;

extern "C" {

int etl_last_error(char* buf, size_t n) {
    if (buf == nullptr || n == 0) return ETL_EARG;
    snprintf(buf, n, "%s", g_error.c_str());
    return ETL_OK;
}

int etl_device_count(int* count) {
    ETL_REQUIRE(count != nullptr, "etl_device_count: null pointer");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) c = 0;
    *count = c;
    return ETL_OK;
}

int etl_init(int device) {
    Context& c = ctx();
    if (c.stream != nullptr) {
        ETL_REQUIRE(device == c.device, "etl_init: already initialised on device %d, asked for %d",
                    c.device, device);
        return ETL_OK;
    }
    int n = 0;
    ETL_HIP(hipGetDeviceCount(&n));
    ETL_REQUIRE(device >= 0 && device < n, "etl_init: device %d does not exist (%d visible)",
                device, n);
    ETL_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    ETL_HIP(hipGetDeviceProperties(&prop, device));
    snprintf(c.name, sizeof c.name, "%s", prop.name);
    snprintf(c.arch, sizeof c.arch, "%s", prop.gcnArchName);
    c.cu_count = prop.multiProcessorCount;
    c.mem = prop.totalGlobalMem;
    if (int rc = observe_setup(device)) return rc;  // with `device` current
    ETL_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    c.lane[0] = c.stream;
    for (int k = 1; k < kLanes; ++k)
        ETL_HIP(hipStreamCreateWithFlags(&c.lane[k], hipStreamNonBlocking));
    for (int k = 0; k < kLanes; ++k)
        ETL_HIP(hipEventCreateWithFlags(&c.lane_ev[k], hipEventDisableTiming));
    c.device = device;
    // every lane's queue gets the scratch of the largest kernel now, not at the first
    // dispatch of a spilling kernel in a run
    for (int k = 0; k < kLanes; ++k)
        if (int rc = warm_scratch(c.lane[k])) return rc;
    return sync_lanes();
}

int etl_device_info(char* buf, size_t n) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(buf != nullptr && n > 0, "etl_device_info: null buffer");
    snprintf(buf, n, "%s %s CUs=%d mem=%zu", ctx().name, ctx().arch, ctx().cu_count, ctx().mem);
    return ETL_OK;
}

int etl_malloc(void** ptr, size_t bytes) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(ptr != nullptr, "etl_malloc: null pointer");
    size_t size = round_size(bytes == 0 ? 1 : bytes);
    {
        std::lock_guard<std::mutex> lock(g_pool_mutex);
        auto it = g_free.find(size);
        if (it != g_free.end()) {
            void* base = it->second;
            g_free.erase(it);
            g_cached_bytes -= size;
            *ptr = static_cast<char*>(base) + kGuard;
            g_live[*ptr] = Block{base, size, bytes};
            return ETL_OK;
        }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, size + 2 * kGuard);
    if (e != hipSuccess) {
        // Release the cache and retry once.
        std::lock_guard<std::mutex> lock(g_pool_mutex);
        if (int rc = sync_lanes()) return rc;
        for (auto& kv : g_free) (void)hipFree(kv.second);
        g_free.clear();
        g_cached_bytes = 0;
        e = hipMalloc(&p, size + 2 * kGuard);
        if (e != hipSuccess) {
            set_error("etl_malloc: hipMalloc(%zu) failed: %s", size, hipGetErrorString(e));
            return ETL_ENOMEM;
        }
    }
    std::lock_guard<std::mutex> lock(g_pool_mutex);
    *ptr = static_cast<char*>(p) + kGuard;
    g_live[*ptr] = Block{p, size, bytes};
    return ETL_OK;
}

int etl_free(void* ptr) {
    if (ptr == nullptr) return ETL_OK;
    ETL_CHECK_INIT();
    std::lock_guard<std::mutex> lock(g_pool_mutex);
    auto it = g_live.find(ptr);
    ETL_REQUIRE(it != g_live.end(), "etl_free: %p was not allocated by etl_malloc", ptr);
    const Block blk = it->second;
    g_live.erase(it);
    if (g_cached_bytes + blk.size <= kCacheLimit) {
        // Stream-ordered reuse: every later use of this block is issued, on either
        // lane, after the work that used it before (the lanes are joined here).
        if (int rc = join_lanes()) return rc;
        g_free.emplace(blk.size, blk.base);
        g_cached_bytes += blk.size;
        return ETL_OK;
    }
    if (int rc = sync_lanes()) return rc;
    ETL_HIP(hipFree(blk.base));
    return ETL_OK;
}

int etl_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    ETL_CHECK_INIT();
    if (bytes == 0) return ETL_OK;
    ETL_REQUIRE(dst && src, "etl_memcpy_h2d: null pointer");
    if (int rc = wait_other_lanes(0)) return rc;  // after the work of every lane
    ETL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx().stream));
    ETL_HIP(hipStreamSynchronize(ctx().stream));
    return ETL_OK;
}

int etl_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    ETL_CHECK_INIT();
    if (bytes == 0) return ETL_OK;
    ETL_REQUIRE(dst && src, "etl_memcpy_d2h: null pointer");
    if (int rc = wait_other_lanes(0)) return rc;  // after the work of every lane
    ETL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx().stream));
    ETL_HIP(hipStreamSynchronize(ctx().stream));
    return ETL_OK;
}

int etl_memcpy_d2d(void* dst, const void* src, size_t bytes) {
    ETL_CHECK_INIT();
    if (bytes == 0) return ETL_OK;
    ETL_REQUIRE(dst && src, "etl_memcpy_d2d: null pointer");
    if (int rc = wait_other_lanes(0)) return rc;  // after the work of every lane
    ETL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx().stream));
    return ETL_OK;
}

int etl_copy_rate(size_t bytes, int reps, float* ms_per_copy) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(ms_per_copy != nullptr && reps >= 1 && bytes >= 16 && bytes % 16 == 0,
                "etl_copy_rate: bytes must be a positive multiple of 16, reps >= 1");
    void* a = nullptr;
    void* b = nullptr;
    int rc = etl_malloc(&a, bytes);
    if (rc == ETL_OK) rc = etl_malloc(&b, bytes);
    if (rc == ETL_OK) rc = zero_async(a, bytes);
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (rc == ETL_OK && (hipEventCreate(&t0) != hipSuccess || hipEventCreate(&t1) != hipSuccess)) {
        set_error("etl_copy_rate: event creation failed");
        rc = ETL_EHIP;
    }
    if (rc == ETL_OK) {
        const size_t n = bytes / 16;
        hipStream_t s = ctx().stream;
        const unsigned blocks = unsigned(std::min<size_t>((n + 255) / 256, 8192));
        for (int r = 0; r <= reps && rc == ETL_OK; ++r) {  // launch 0 untimed
            if (r == 1 && hipEventRecord(t0, s) != hipSuccess) rc = ETL_EHIP;
            hipLaunchKernelGGL(stream_copy_kernel, dim3(blocks), dim3(256), 0, s,
                               static_cast<const float4*>(a), static_cast<float4*>(b), n);
            if (hipGetLastError() != hipSuccess) rc = ETL_EHIP;
        }
        float ms = 0.f;
        if (rc == ETL_OK && (hipEventRecord(t1, s) != hipSuccess ||
                             hipEventSynchronize(t1) != hipSuccess ||
                             hipEventElapsedTime(&ms, t0, t1) != hipSuccess))
            rc = ETL_EHIP;
        if (rc == ETL_OK) *ms_per_copy = ms / float(reps);
        else set_error("etl_copy_rate: copy or timing failed");
    }
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    if (a) etl_free(a);
    if (b) etl_free(b);
    return rc;
}

int etl_zero(void* dst, size_t bytes) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(dst != nullptr || bytes == 0, "etl_zero: null pointer");
    return zero_async(dst, bytes);
}

int etl_synchronize(void) {
    ETL_CHECK_INIT();
    return sync_lanes();
}

int etl_lane_wait(int waiter, int signaler) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(waiter >= 0 && waiter < kLanes && signaler >= 0 && signaler < kLanes,
                "etl_lane_wait: lanes are 0 .. %d", kLanes - 1);
    if (waiter == signaler) return ETL_OK;
    ETL_HIP(hipEventRecord(ctx().lane_ev[signaler], ctx().lane[signaler]));
    ETL_HIP(hipStreamWaitEvent(ctx().lane[waiter], ctx().lane_ev[signaler], 0));
    return ETL_OK;
}

int etl_memcpy_d2d_lane(void* dst, const void* src, size_t bytes, int lane) {
    ETL_CHECK_INIT();
    ETL_REQUIRE(lane >= 0 && lane < kLanes, "etl_memcpy_d2d_lane: lanes are 0 .. %d", kLanes - 1);
    if (bytes == 0) return ETL_OK;
    ETL_REQUIRE(dst && src, "etl_memcpy_d2d_lane: null pointer");
    ETL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx().lane[lane]));
    return ETL_OK;
}

}  // extern "C"
