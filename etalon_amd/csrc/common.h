// Shared runtime state and helpers of libetalon_hip.so (gfx950 only).
//
// Floating-point rules for every kernel in this library: the build uses
// -ffp-contract=off (no fused multiply-add), keeps f32 denormals, and never uses
// fast-math, so each + - * rounds once in IEEE binary32/binary64, exactly as the
// NumPy reference backend does (etalon/fdtd/backends/numpy_backend.py).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/etalon_hip.h"

namespace etl {

// Last error message of the calling thread.
void set_error(const char* fmt, ...);

// Streams of the library: lane 0 (`stream`) carries every call that names no lane;
// plans may be moved to lane 1 so that two plans run concurrently
// (etl_plan_set_lane, etl_adjplan_set_lane), ordered by etl_lane_wait.
constexpr int kLanes = 2;

struct Context {
    int device = -1;
    hipStream_t stream = nullptr;
    hipStream_t lane[kLanes] = {nullptr};
    hipEvent_t lane_ev[kLanes] = {nullptr};
    int cu_count = 0;
    char name[256] = {0};
    char arch[64] = {0};
    size_t mem = 0;
};
Context& ctx();
int require_init();
// The stream of lane `k` (0 <= k < kLanes).
inline hipStream_t lane_stream(int k) { return ctx().lane[k]; }
// Wait on the host for the work of every lane.
int sync_lanes();
// Make every lane's later work wait for every lane's earlier work (no host wait).
int join_lanes();
// Make lane k's later work wait for the earlier work of the other lanes.
int wait_other_lanes(int k);
// Stream-ordered zero fill with a kernel (safe inside hipGraph capture).
int zero_async(void* dst, size_t bytes);
// Valid device memory that etl_malloc places before and after every array.
constexpr size_t kGuard = 256;
// Bytes asked for when `p` is the start of a live etl_malloc array, else 0.
size_t guarded_bytes(const void* p);
// Per-device set-up of the observer reductions (observe.hip), called by etl_init with
// the device current.
int observe_setup(int device);
// Size the scratch of stream s's queue for the library's largest kernel private segment
// (joint.h, warm_scratch_fwd / warm_scratch_pair); called by etl_init for each lane.
int warm_scratch(hipStream_t s);

}  // namespace etl

#define ETL_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t etl_err_ = (call);                                                   \
        if (etl_err_ != hipSuccess) {                                                   \
            etl::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(etl_err_), \
                           __FILE__, __LINE__);                                         \
            return ETL_EHIP;                                                            \
        }                                                                               \
    } while (0)

#define ETL_REQUIRE(cond, ...)          \
    do {                                \
        if (!(cond)) {                  \
            etl::set_error(__VA_ARGS__); \
            return ETL_EARG;            \
        }                               \
    } while (0)

#define ETL_CHECK_INIT()                     \
    do {                                     \
        int etl_rc_ = etl::require_init();   \
        if (etl_rc_ != ETL_OK) return etl_rc_; \
    } while (0)

// Launch check: the launch itself (configuration errors) is reported here; faults
// inside kernels surface at the next synchronising call.
#define ETL_LAUNCHED() ETL_HIP(hipGetLastError())
