// trex_amd C ABI: module loading, device memory plumbing, pattern handles and
// kernel launches for the HIP matching kernels (gfx950).
//
// Every entry point returns 0 on success or a negative TREX_ERR_* code and
// never exits the process; trex_last_error() returns the message of the last
// failure on the calling thread. Kernels live in kernels/*.hsaco next to this
// library (one gfx950 code object per kernel source file) and are launched
// with hipModuleLaunchKernel.

#include "trex_hip.h"

#include <hip/hip_runtime_api.h>

#include <dirent.h>
#include <dlfcn.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "runtime_internal.h"

namespace trex {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    return fail(TREX_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

// ---------------------------------------------------------------------------
// Module
// ---------------------------------------------------------------------------
namespace {
std::once_flag g_module_once;
int g_module_status = TREX_ERR_NOT_INIT;
std::string g_module_error;
std::vector<hipModule_t> g_modules;
std::vector<std::pair<std::string, hipFunction_t>> g_functions;

std::string library_dir() {
    Dl_info info;
    if (dladdr(reinterpret_cast<void*>(&trex_device_check), &info) == 0 || !info.dli_fname) return ".";
    std::string path = info.dli_fname;
    size_t slash = path.rfind('/');
    return slash == std::string::npos ? std::string(".") : path.substr(0, slash);
}

void load_module() {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        g_module_status = TREX_ERR_NO_DEVICE;
        g_module_error = "no HIP device";
        return;
    }
    const std::string dir = library_dir() + "/kernels";
    std::vector<std::string> files;
    if (DIR* d = opendir(dir.c_str())) {
        while (dirent* ent = readdir(d)) {
            std::string name = ent->d_name;
            if (name.size() > 6 && name.compare(name.size() - 6, 6, ".hsaco") == 0) files.push_back(dir + "/" + name);
        }
        closedir(d);
    }
    std::sort(files.begin(), files.end());
    if (files.empty()) {
        g_module_status = TREX_ERR_NOT_INIT;
        g_module_error = "no kernel code objects in " + dir + " (run python -m trex_amd.build)";
        return;
    }
    for (const auto& path : files) {
        hipModule_t m = nullptr;
        hipError_t e = hipModuleLoad(&m, path.c_str());
        if (e != hipSuccess) {
            g_module_status = TREX_ERR_HIP;
            g_module_error = "hipModuleLoad(" + path + "): " + hipGetErrorString(e);
            return;
        }
        g_modules.push_back(m);
    }
    // Stream-ordered allocations (hipMallocAsync: the chunked probabilistic
    // path's workspace, mixed-queue plan buffers) come from the device's
    // default pool; by default it returns memory to the system at every
    // synchronisation and maps it again on the next allocation. Keep it.
    int dev = 0;
    hipMemPool_t pool = nullptr;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess && pool) {
        uint64_t keep = UINT64_MAX;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    g_module_status = TREX_OK;
}
}  // namespace

int ensure_module() {
    std::call_once(g_module_once, load_module);
    if (g_module_status != TREX_OK) return fail(g_module_status, "%s", g_module_error.c_str());
    return TREX_OK;
}

int get_function(const char* name, hipFunction_t* out) {
    static std::mutex mu;
    int rc = ensure_module();
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(mu);
    for (auto& kv : g_functions)
        if (kv.first == name) { *out = kv.second; return TREX_OK; }
    for (hipModule_t m : g_modules) {
        hipFunction_t f = nullptr;
        if (hipModuleGetFunction(&f, m, name) == hipSuccess && f) {
            g_functions.emplace_back(name, f);
            *out = f;
            return TREX_OK;
        }
    }
    return fail(TREX_ERR_HIP, "kernel %s not found in any loaded code object", name);
}

int launch(const char* name, unsigned grid, unsigned block, unsigned lds, hipStream_t stream, void** args) {
    hipFunction_t f;
    int rc = get_function(name, &f);
    if (rc) return rc;
    if (grid == 0) return TREX_OK;
    hipError_t e = hipModuleLaunchKernel(f, grid, 1, 1, block, 1, 1, lds, stream, args, nullptr);
    if (e != hipSuccess) return fail(TREX_ERR_HIP, "launch %s: %s", name, hipGetErrorString(e));
    return TREX_OK;
}

// ---------------------------------------------------------------------------
// Growable device buffer
// ---------------------------------------------------------------------------
int DevBuf::reserve(size_t n) {
    if (n <= cap) return TREX_OK;
    if (ptr) { hipFree(ptr); ptr = nullptr; cap = 0; }
    size_t want = std::max<size_t>(n, 256);
    hipError_t e = hipMalloc(&ptr, want);
    if (e != hipSuccess) { ptr = nullptr; return hip_fail(e, "hipMalloc"); }
    cap = want;
    return TREX_OK;
}

void DevBuf::release() {
    if (ptr) hipFree(ptr);
    ptr = nullptr;
    cap = 0;
}

int device_cu_count() {
    static int cached = 0;
    if (cached > 0) return cached;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    cached = cus;
    return cached;
}

static std::atomic<uint32_t> g_wait_bound_us{kWaitBoundUs};

uint64_t device_wait_ticks() {
    static std::atomic<int> cached_khz{0};
    int khz = cached_khz.load();
    if (khz <= 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
            khz = 100000;   // the 100 MHz counter of gfx9 parts
        cached_khz.store(khz);
    }
    return (uint64_t)khz * g_wait_bound_us.load() / 1000u;
}

int ensure_host_dispatch(hipStream_t& stream, hipEvent_t (&ev)[4]) {
    if (stream) return TREX_OK;
    hipStream_t s = nullptr;
    TREX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    for (int i = 0; i < 4; ++i) {
        const hipError_t e = hipEventCreate(&ev[i]);
        if (e != hipSuccess) {
            for (int j = 0; j < i; ++j) { hipEventDestroy(ev[j]); ev[j] = nullptr; }
            ev[i] = nullptr;
            hipStreamDestroy(s);
            return hip_fail(e, "hipEventCreate");
        }
    }
    stream = s;
    return TREX_OK;
}

int upload(void** dptr, const void* src, size_t n) {
    hipError_t e = hipMalloc(dptr, std::max<size_t>(n, 16));
    if (e != hipSuccess) { *dptr = nullptr; return hip_fail(e, "hipMalloc"); }
    e = hipMemcpy(*dptr, src, n, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy H2D");
    return TREX_OK;
}

}  // namespace trex

using namespace trex;

// ---------------------------------------------------------------------------
// Device, errors, memory, streams
// ---------------------------------------------------------------------------
extern "C" {

int trex_device_check(void) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(TREX_ERR_NO_DEVICE, "no HIP device");
    hipDeviceProp_t prop;
    int dev = 0;
    hipGetDevice(&dev);
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return fail(TREX_ERR_NO_DEVICE, "hipGetDeviceProperties failed");
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(TREX_ERR_ARCH, "device is %s; the kernels are built for gfx950", prop.gcnArchName);
    return TREX_OK;
}

int trex_set_wait_bound_us(uint32_t us) {
    g_wait_bound_us.store(us ? us : kWaitBoundUs);
    return TREX_OK;
}

int trex_last_error(char* buf, int n) {
    if (!buf || n <= 0) return (int)g_last_error.size();
    snprintf(buf, (size_t)n, "%s", g_last_error.c_str());
    return (int)g_last_error.size();
}

int trex_malloc(void** ptr, size_t bytes) {
    if (!ptr) return fail(TREX_ERR_INVALID, "null out pointer");
    hipError_t e = hipMalloc(ptr, std::max<size_t>(bytes, 1));
    return e == hipSuccess ? TREX_OK : hip_fail(e, "hipMalloc");
}

int trex_free(void* ptr) {
    hipError_t e = hipFree(ptr);
    return e == hipSuccess ? TREX_OK : hip_fail(e, "hipFree");
}

int trex_host_alloc(void** ptr, size_t bytes) {
    if (!ptr) return fail(TREX_ERR_INVALID, "null out pointer");
    hipError_t e = hipHostMalloc(ptr, std::max<size_t>(bytes, 1), hipHostMallocDefault);
    return e == hipSuccess ? TREX_OK : hip_fail(e, "hipHostMalloc");
}

int trex_host_free(void* ptr) {
    hipError_t e = hipHostFree(ptr);
    return e == hipSuccess ? TREX_OK : hip_fail(e, "hipHostFree");
}

int trex_memcpy_h2d(void* dst, const void* src, size_t n, void* stream) {
    hipError_t e = hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, (hipStream_t)stream);
    return e == hipSuccess ? TREX_OK : hip_fail(e, "hipMemcpyAsync H2D");
}

int trex_memcpy_d2h(void* dst, const void* src, size_t n, void* stream) {
    hipError_t e = hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, (hipStream_t)stream);
    return e == hipSuccess ? TREX_OK : hip_fail(e, "hipMemcpyAsync D2H");
}

int trex_memcpy_d2d(void* dst, const void* src, size_t n, void* stream) {
    hipError_t e = hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToDevice, (hipStream_t)stream);
    return e == hipSuccess ? TREX_OK : hip_fail(e, "hipMemcpyAsync D2D");
}

int trex_stream_create(void** stream) {
    if (!stream) return fail(TREX_ERR_INVALID, "null out pointer");
    hipStream_t s;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
    *stream = s;
    return TREX_OK;
}

int trex_stream_destroy(void* stream) {
    hipError_t e = hipStreamDestroy((hipStream_t)stream);
    return e == hipSuccess ? TREX_OK : hip_fail(e, "hipStreamDestroy");
}

int trex_stream_sync(void* stream) {
    hipError_t e = hipStreamSynchronize((hipStream_t)stream);
    return e == hipSuccess ? TREX_OK : hip_fail(e, "hipStreamSynchronize");
}

int trex_device_sync(void) {
    hipError_t e = hipDeviceSynchronize();
    return e == hipSuccess ? TREX_OK : hip_fail(e, "hipDeviceSynchronize");
}

int trex_event_create(void** ev) {
    if (!ev) return fail(TREX_ERR_INVALID, "null out pointer");
    hipEvent_t e0;
    hipError_t e = hipEventCreate(&e0);
    if (e != hipSuccess) return hip_fail(e, "hipEventCreate");
    *ev = e0;
    return TREX_OK;
}

int trex_event_destroy(void* ev) {
    hipError_t e = hipEventDestroy((hipEvent_t)ev);
    return e == hipSuccess ? TREX_OK : hip_fail(e, "hipEventDestroy");
}

int trex_event_record(void* ev, void* stream) {
    hipError_t e = hipEventRecord((hipEvent_t)ev, (hipStream_t)stream);
    return e == hipSuccess ? TREX_OK : hip_fail(e, "hipEventRecord");
}

int trex_event_elapsed_ms(void* start, void* stop, float* ms) {
    hipError_t e = hipEventSynchronize((hipEvent_t)stop);
    if (e != hipSuccess) return hip_fail(e, "hipEventSynchronize");
    e = hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop);
    return e == hipSuccess ? TREX_OK : hip_fail(e, "hipEventElapsedTime");
}

int trex_event_query(void* ev) {
    const hipError_t e = hipEventQuery((hipEvent_t)ev);
    if (e == hipSuccess) return 0;
    if (e == hipErrorNotReady) return 1;
    return hip_fail(e, "hipEventQuery");
}

int trex_event_sync(void* ev) {
    TREX_HIP(hipEventSynchronize((hipEvent_t)ev), "hipEventSynchronize");
    return TREX_OK;
}

int trex_stream_wait_event(void* stream, void* ev) {
    TREX_HIP(hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)ev, 0), "hipStreamWaitEvent");
    return TREX_OK;
}

// ---------------------------------------------------------------------------
// Graphs of captured launches
// ---------------------------------------------------------------------------
}  // extern "C"

struct trex_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" {

int trex_graph_begin(void* stream) {
    if (!stream) return fail(TREX_ERR_INVALID, "graph capture needs a non-default stream");
    TREX_HIP(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    return TREX_OK;
}

int trex_graph_end(void* stream, trex_graph** out) {
    if (!stream || !out) return fail(TREX_ERR_INVALID, "null argument");
    *out = nullptr;
    hipGraph_t g = nullptr;
    TREX_HIP(hipStreamEndCapture((hipStream_t)stream, &g), "hipStreamEndCapture");
    trex_graph* h = new trex_graph();
    h->graph = g;
    const hipError_t e = hipGraphInstantiate(&h->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        trex_graph_destroy(h);
        return hip_fail(e, "hipGraphInstantiate");
    }
    *out = h;
    return TREX_OK;
}

int trex_graph_abort(void* stream) {
    if (!stream) return TREX_OK;
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing((hipStream_t)stream, &st) == hipSuccess && st != hipStreamCaptureStatusNone) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture((hipStream_t)stream, &g);
        if (g) hipGraphDestroy(g);
    }
    return TREX_OK;
}

int trex_graph_launch(trex_graph* g, void* stream) {
    if (!g || !g->exec) return fail(TREX_ERR_INVALID, "null graph");
    TREX_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream), "hipGraphLaunch");
    return TREX_OK;
}

int trex_graph_destroy(trex_graph* g) {
    if (!g) return TREX_OK;
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    delete g;
    return TREX_OK;
}

int trex_device_info(char* name, int n, int* cu_count, int64_t* hbm_bytes) {
    int dev = 0;
    TREX_HIP(hipGetDevice(&dev), "hipGetDevice");
    hipDeviceProp_t prop;
    TREX_HIP(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    if (name && n > 0) snprintf(name, (size_t)n, "%s", prop.gcnArchName);
    if (cu_count) *cu_count = prop.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = (int64_t)prop.totalGlobalMem;
    return TREX_OK;
}

int trex_memcpy_kernel(void* dst, const void* src, int64_t n, void* stream) {
    if (n < 0) return fail(TREX_ERR_INVALID, "negative size");
    if (n == 0) return TREX_OK;
    if (!dst || !src || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15))
        return fail(TREX_ERR_INVALID, "copy_bytes needs 16-byte aligned, non-null pointers");
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, src) != hipSuccess) {
        (void)hipGetLastError();
        return fail(TREX_ERR_INVALID, "source is neither device memory nor pinned host memory");
    }
    if (a.type != hipMemoryTypeHost && a.type != hipMemoryTypeDevice)
        return fail(TREX_ERR_INVALID, "source is neither device memory nor pinned host memory (type %d)", (int)a.type);
    // Pinned host memory is mapped into the device's address space at its host address.
    const int64_t n16 = n >> 4;
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n16 + 255) / 256, 1024));
    const uint8_t* s = static_cast<const uint8_t*>(src);
    uint8_t* d = static_cast<uint8_t*>(dst);
    void* args[] = {&d, &s, &n};
    return launch("copy_bytes", grid, 256, 0, (hipStream_t)stream, args);
}

int trex_gen_symbols(void* d_out, int64_t n_bytes, uint64_t seed, uint32_t lo, uint32_t span, void* stream) {
    if (!d_out || n_bytes < 0 || n_bytes % 16 || ((uintptr_t)d_out & 15))
        return fail(TREX_ERR_INVALID, "gen_symbols needs a 16-byte aligned buffer and a multiple of 16 bytes");
    if (span < 1 || span > 256 || (span & (span - 1)) || lo + span > 256)
        return fail(TREX_ERR_INVALID, "span must be a power of two with lo + span <= 256");
    int64_t n16 = n_bytes / 16;
    if (n16 == 0) return TREX_OK;
    const unsigned grid = (unsigned)std::min<int64_t>((n16 + 255) / 256, 65536);
    void* args[] = {&d_out, &n16, &seed, &lo, &span};
    return launch("gen_symbols", grid, 256, 0, (hipStream_t)stream, args);
}

}  // extern "C"
