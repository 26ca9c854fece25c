// A host-only stand-in for the HIP runtime, for checking libetalon_hip's host code
// (plans, graphs, tables, the caching allocator, teardown) under AddressSanitizer on a
// machine without a GPU.
//
// The library's sources are compiled for the host only (no device code) and linked
// against these definitions instead of libamdhip64: device memory is host heap memory,
// copies are memcpy, kernel launches do nothing but the zero fills (memset).  Streams, events, graphs and graph
// execs are heap objects that every call dereferences and checks, and destroying one
// frees it, so a use after destruction — a stale graph exec replayed, an event recorded
// after its plan was torn down, a stream used after its plan destroyed it — is a
// heap-use-after-free that AddressSanitizer reports with both stacks.  Launches check
// their stream the same way and that no stream they touch is being torn down.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>

namespace {

constexpr unsigned kStreamMagic = 0x5354524du, kEventMagic = 0x45564e54u,
                   kGraphMagic = 0x47524150u, kExecMagic = 0x45584543u;

struct Obj {
    unsigned magic;
};
struct Stream : Obj {
    int capturing = 0;
};
struct Event : Obj {
    int recorded = 0;
};
struct Graph : Obj {};
struct Exec : Obj {};

[[noreturn]] void fail(const char* what, const void* p) {
    std::fprintf(stderr, "hip_stub: %s (%p)\n", what, p);
    std::fflush(stderr);
    std::abort();
}

// A null stream is the legacy default stream; anything else must be a live Stream.
Stream* stream_of(hipStream_t s) {
    if (s == nullptr) return nullptr;
    auto* st = reinterpret_cast<Stream*>(s);
    if (st->magic != kStreamMagic) fail("not a live stream", s);
    return st;
}
Event* event_of(hipEvent_t e) {
    auto* ev = reinterpret_cast<Event*>(e);
    if (ev == nullptr || ev->magic != kEventMagic) fail("not a live event", e);
    return ev;
}

struct LaunchConfig {
    dim3 grid, block;
    size_t shm;
    hipStream_t stream;
};
thread_local LaunchConfig g_config;

hipDeviceProp_t g_prop;

// Kernel names by host stub, from the registration calls of each object's constructor.
std::map<const void*, std::string>& kernel_names() {
    static std::map<const void*, std::string> names;
    return names;
}

}  // namespace

extern "C" {

hipError_t hipGetDeviceCount(int* n) {
    *n = 1;
    return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetDevicePropertiesR0600(hipDeviceProp_t* prop, int) {
    std::memset(&g_prop, 0, sizeof g_prop);
    std::snprintf(g_prop.name, sizeof g_prop.name, "host stub");
    std::snprintf(g_prop.gcnArchName, sizeof g_prop.gcnArchName, "gfx950 (host stub)");
    g_prop.multiProcessorCount = 256;
    g_prop.totalGlobalMem = size_t(288) << 30;
    *prop = g_prop;
    return hipSuccess;
}
const char* hipGetErrorString(hipError_t) { return "hip_stub error"; }
hipError_t hipGetLastError(void) { return hipSuccess; }

hipError_t hipMalloc(void** p, size_t bytes) {
    *p = std::calloc(1, bytes ? bytes : 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* p) {
    std::free(p);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind, hipStream_t s) {
    stream_of(s);
    if (bytes) std::memcpy(dst, src, bytes);
    return hipSuccess;
}
hipError_t hipGetSymbolAddress(void** p, const void* symbol) {
    *p = const_cast<void*>(symbol);  // the host shadow of the __device__ variable
    return hipSuccess;
}

hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
    auto* st = new Stream;
    st->magic = kStreamMagic;
    *s = reinterpret_cast<hipStream_t>(st);
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
    Stream* st = stream_of(s);
    if (st == nullptr) fail("destroying the null stream", s);
    if (st->capturing) fail("destroying a stream under capture", s);
    st->magic = 0;
    delete st;
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s) {
    Stream* st = stream_of(s);
    if (st && st->capturing) fail("synchronising a stream under capture", s);
    return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned) {
    stream_of(s);
    event_of(e);
    return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
    Stream* st = stream_of(s);
    if (st == nullptr || st->capturing) fail("capture on the null or a capturing stream", s);
    st->capturing = 1;
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g) {
    Stream* st = stream_of(s);
    if (st == nullptr || !st->capturing) fail("end of a capture that never began", s);
    st->capturing = 0;
    auto* gr = new Graph;
    gr->magic = kGraphMagic;
    *g = reinterpret_cast<hipGraph_t>(gr);
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* ex, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
    auto* gr = reinterpret_cast<Graph*>(g);
    if (gr == nullptr || gr->magic != kGraphMagic) fail("instantiating a dead graph", g);
    auto* e = new Exec;
    e->magic = kExecMagic;
    *ex = reinterpret_cast<hipGraphExec_t>(e);
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t ex, hipStream_t s) {
    auto* e = reinterpret_cast<Exec*>(ex);
    if (e == nullptr || e->magic != kExecMagic) fail("launching a dead graph exec", ex);
    stream_of(s);
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t ex) {
    auto* e = reinterpret_cast<Exec*>(ex);
    if (e == nullptr || e->magic != kExecMagic) fail("destroying a dead graph exec", ex);
    e->magic = 0;
    delete e;
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
    auto* gr = reinterpret_cast<Graph*>(g);
    if (gr == nullptr || gr->magic != kGraphMagic) fail("destroying a dead graph", g);
    gr->magic = 0;
    delete gr;
    return hipSuccess;
}

hipError_t hipEventCreate(hipEvent_t* e) {
    auto* ev = new Event;
    ev->magic = kEventMagic;
    *e = reinterpret_cast<hipEvent_t>(ev);
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) {
    Event* ev = event_of(e);
    ev->magic = 0;
    delete ev;
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
    event_of(e)->recorded = 1;
    stream_of(s);
    return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t e) {
    event_of(e);
    return hipSuccess;
}
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
    if (!event_of(a)->recorded || !event_of(b)->recorded) fail("timing an unrecorded event", a);
    *ms = 0.01f;
    return hipSuccess;
}

// Kernel launches: <<<>>> pushes the configuration, the host stub pops it and calls
// hipLaunchKernel, which checks the stream and runs nothing.
hipError_t __hipPushCallConfiguration(dim3 grid, dim3 block, size_t shm, hipStream_t s) {
    g_config = {grid, block, shm, s};
    return hipSuccess;
}
hipError_t __hipPopCallConfiguration(dim3* grid, dim3* block, size_t* shm, hipStream_t* s) {
    *grid = g_config.grid;
    *block = g_config.block;
    *shm = g_config.shm;
    *s = g_config.stream;
    return hipSuccess;
}
hipError_t hipLaunchKernel(const void* fn, dim3 grid, dim3 block, void** args, size_t,
                           hipStream_t s) {
    stream_of(s);
    if (grid.x == 0 || block.x == 0 || block.x > 1024) fail("launch shape", nullptr);
    // The zero fills are the one kernel whose effect host code reads back (the flags of
    // the plan's state checks start at zero): they run here, as memset — which also
    // checks that their target is a live allocation of at least that size.
    auto it = kernel_names().find(fn);
    if (it != kernel_names().end()) {
        const std::string& name = it->second;
        if (name.find("fill_zero_kernel") != std::string::npos)
            std::memset(*static_cast<uint32_t**>(args[0]), 0, *static_cast<size_t*>(args[1]) * 4);
        else if (name.find("fill_zero_bytes_kernel") != std::string::npos)
            std::memset(*static_cast<char**>(args[0]), 0, *static_cast<size_t*>(args[1]));
    }
    return hipSuccess;
}

// Code-object registration of the host-only build: nothing to register.
void** __hipRegisterFatBinary(const void*) {
    static void* handle = nullptr;
    return &handle;
}
void __hipUnregisterFatBinary(void**) {}
void __hipRegisterFunction(void**, const void* host_fn, char*, const char* device_name,
                           unsigned int, void*, void*, void*, void*, int*) {
    kernel_names()[host_fn] = device_name;
}
void __hipRegisterVar(void**, void*, char*, const char*, int, size_t, int, int) {}

}  // extern "C"
