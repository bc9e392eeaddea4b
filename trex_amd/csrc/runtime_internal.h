// Internal helpers shared by the trex_amd runtime translation units.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#include <string>

namespace trex {

extern thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int hip_fail(hipError_t e, const char* what);
int ensure_module();
int get_function(const char* name, hipFunction_t* out);
// Launch a 1-D grid of `name` from the kernel module.
int launch(const char* name, unsigned grid, unsigned block, unsigned lds, hipStream_t stream, void** args);
// hipMalloc + synchronous copy of a host table.
int upload(void** dptr, const void* src, size_t n);

struct DevBuf {
    void* ptr = nullptr;
    size_t cap = 0;
    int reserve(size_t n);
    void release();
    template <class T> T* as() const { return static_cast<T*>(ptr); }
};

#define TREX_TRY(expr) do { int _rc = (expr); if (_rc) return _rc; } while (0)
#define TREX_HIP(expr, what) do { hipError_t _e = (expr); if (_e != hipSuccess) return ::trex::hip_fail(_e, what); } while (0)

constexpr unsigned kBlock = 256;   // threads per workgroup of the string kernels

}  // namespace trex
