// Internal helpers shared by the trex_amd runtime translation units.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#include <mutex>
#include <string>

#include "trex_hip.h"

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

// Creates the stream and four timing events of a handle's synchronous host
// entry point on its first use (call with the handle's mutex held). A process
// that keeps many patterns resident and drives them on its own streams never
// creates them, so they take none of the runtime's hardware queues.
int ensure_host_dispatch(hipStream_t& stream, hipEvent_t (&ev)[4]);

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

// Number of compute units of the current device (cached).
int device_cu_count();

// Time bound of the kernels' cross-wave waits in ticks of the device counter
// wall_clock64 reads, on the current device: trex_set_wait_bound_us, by
// default kWaitBoundUs.
uint64_t device_wait_ticks();
// Default time bound of a cross-wave wait: far above any legitimate wait (a
// predecessor's phase 1 over one chunk takes microseconds), far below a test's
// time limit.
constexpr uint32_t kWaitBoundUs = 100000;

// Dependent chains per lane in phase 1 of the look-back scan by default: 1
// (each chunk walked front to back) or 2 (its two halves side by side).
constexpr int kScanChains = 1;

}  // namespace trex

// A resident DFA byte table (the handle behind trex_dfa_*).
struct trex_dfa {
    int n_states = 0;
    int stride = 16;          // row stride of t1 (states, rounded up to 16)
    int start = 0;
    int dead = -1;            // absorbing non-accepting state, -1 if none
    int window_base = -1;     // aligned 32-byte window holding every live byte, -1 if none
    int auto_variant = 1;
    int variant = 1;
    uint8_t* d_t1 = nullptr;  // [256][stride]
    uint8_t* d_acc = nullptr; // [stride]
    uint8_t* d_t2 = nullptr;  // [1024][16] pair table t2[pair * 16 + s] (window variants only)
    uint8_t* d_f2 = nullptr;  // [1024] uint64 pair functions, nibble s = successor of s
    int prefetch = 8;         // 16-byte vectors in flight per lane (8 or 16)
    int rotation = -1;        // rotated read order of contiguous strings: -1 auto, 0 off, 1 on
    bool synchronizes = false;  // some window word sends every state that can occur to one state
    uint32_t rot_init[4] = {0, 0, 0, 0};  // starting function of the rotated kernel's composition
    // staging for trex_dfa_match_host
    trex::DevBuf in_bytes, in_offsets, out_state, out_accept;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    std::mutex mu;            // serialises host dispatches on this handle
};

// A resident Boolean NFA-64 (the handle behind trex_nfa_*).
struct trex_nfa {
    int n_symbols = 0;
    uint64_t start = 0, accept = 0;
    int window_base = -1;
    int variant = TREX_NFA_BYTES;
    int n_classes = 0, n_classes_win = 0;
    uint64_t* d_tab = nullptr;      // [n_symbols + 2][16][16]: symbols, reject, skip (bank-swizzled)
    uint8_t* d_cls = nullptr;       // [256] byte -> class
    uint64_t* d_tab_win = nullptr;  // [34][16][16]: window bytes base+0..31, reject, skip
    uint8_t* d_cls_win = nullptr;   // [256]
    uint8_t* d_tab5 = nullptr;      // 32 window classes of 5-bit slice tables, class k at byte k * 3336
    uint8_t* d_tab6 = nullptr;      // 32 window classes of 6/5-bit slice tables, class k at byte k * 4104
    trex::DevBuf in_bytes, in_offsets, out_mask, out_accept;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    std::mutex mu;
};
