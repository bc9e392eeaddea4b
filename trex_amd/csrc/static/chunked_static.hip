// trex_dfa_match_chunked_static: the decoupled look-back over long strings
// (trex_plan_create + trex_dfa_match_plan) in one call that allocates no
// device memory, for callers that cannot allocate (code under HIP graph
// capture, tight serving loops).
//
// Every buffer of the launch beyond its inputs and outputs (the chunk table,
// the look-back status, aggregates and jumps, the ticket) lives in `scratch`,
// a static device array of kStaticScratchBytes that each GPU holds for the
// life of the process. Its device address is looked up once per device with
// hipGetSymbolAddress and passed to the kernels as a plain pointer. Every call
// on one GPU shares that GPU's array: calls that overlap on one GPU (on
// different streams, or from several threads) race and give wrong results.
//
// This file is compiled by hipcc as a HIP source (host code plus the array's
// device definition) and linked into libtrexhip.so; the kernels it launches
// live in the gfx950 code objects (scan_kernels.hip).

#include <hip/hip_runtime.h>

#include <mutex>

#include "runtime_internal.h"
#include "trex_hip.h"

using namespace trex;

namespace {

constexpr int64_t kMaxStrings = TREX_STATIC_MAX_STRINGS;
constexpr int64_t kMaxChunks = TREX_STATIC_MAX_CHUNKS;

// Layout of the array (byte offsets, each 256-aligned).
constexpr size_t align256(size_t x) { return (x + 255) / 256 * 256; }
constexpr size_t kOffTicket = 0;                                          // uint32
constexpr size_t kOffFirst = 256;                                         // int64 [kMaxStrings + 1]
constexpr size_t kOffBegin = kOffFirst + align256((kMaxStrings + 1) * 8);  // int64 [kMaxChunks]
constexpr size_t kOffEnd = kOffBegin + kMaxChunks * 8;                     // int64
constexpr size_t kOffAgg = kOffEnd + kMaxChunks * 8;                       // uint64
constexpr size_t kOffSid = kOffAgg + kMaxChunks * 8;                       // int32
constexpr size_t kOffPred = kOffSid + kMaxChunks * 4;                      // int32
constexpr size_t kOffFlags = kOffPred + kMaxChunks * 4;                    // uint32
constexpr size_t kOffStatus = kOffFlags + kMaxChunks * 4;                  // uint32
constexpr size_t kOffJump = kOffStatus + kMaxChunks * 4;                   // int32
constexpr size_t kStaticScratchBytes = kOffJump + kMaxChunks * 4;
static_assert(kStaticScratchBytes == TREX_STATIC_SCRATCH_BYTES, "header and layout disagree");

constexpr int kMaxDevices = 64;

}  // namespace

// The array itself. In host code `scratch` names a host-side stand-in, never
// the device memory: host code uses the address from scratch_ptr() only.
static __device__ __align__(256) uint8_t scratch[kStaticScratchBytes];

namespace {

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up on the first
// call on each device (with that device current) and kept.
int scratch_ptr(uint8_t** out) {
    int dev = 0;
    TREX_HIP(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(TREX_ERR_INVALID, "device %d beyond %d devices", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!p) return fail(TREX_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
        g_scratch_ptr[dev] = p;
    }
    *out = static_cast<uint8_t*>(g_scratch_ptr[dev]);
    return TREX_OK;
}

}  // namespace

extern "C" {

int trex_dfa_match_chunked_static(trex_dfa* h, const uint8_t* d_bytes, int64_t bytes_len, const int64_t* d_begin,
                                  const int64_t* d_end, int64_t n_strings, int64_t chunk_len,
                                  uint8_t* d_final_state, uint8_t* d_accept, uint32_t* d_error, uint32_t max_spins,
                                  void* stream) {
    if (!h) return fail(TREX_ERR_INVALID, "null handle");
    if (h->n_states > 16) return fail(TREX_ERR_INVALID, "chunked matching needs a table of at most 16 states");
    if (chunk_len < 64 || chunk_len % 16 != 0) return fail(TREX_ERR_INVALID, "chunk_len must be a multiple of 16, >= 64");
    if (n_strings < 0 || bytes_len < 0) return fail(TREX_ERR_INVALID, "negative size");
    if (!d_error) return fail(TREX_ERR_INVALID, "null error word");
    if (n_strings > kMaxStrings)
        return fail(TREX_ERR_CAPACITY, "%lld strings; the allocation-free form takes at most %lld",
                    (long long)n_strings, (long long)kMaxStrings);
    // chunks of ranges that lie in [0, bytes_len) without overlapping: at most
    // one per string plus one per chunk_len bytes
    int64_t n_bound = n_strings + (bytes_len + chunk_len - 1) / chunk_len;
    if (n_bound > kMaxChunks)
        return fail(TREX_ERR_CAPACITY,
                    "%lld strings over %lld bytes in chunks of %lld may need %lld chunks; the allocation-free "
                    "form takes at most %lld (use longer chunks or trex_plan_create)",
                    (long long)n_strings, (long long)bytes_len, (long long)chunk_len, (long long)n_bound,
                    (long long)kMaxChunks);
    hipStream_t s = (hipStream_t)stream;
    if (n_strings == 0) {
        int64_t one = 1;
        uint32_t zero = 0;
        void* a[] = {&d_error, &one, &zero};
        return launch("fill_u32", 1, kBlock, 0, s, a);
    }
    if (!d_bytes || !d_begin || !d_end || !d_final_state || !d_accept) return fail(TREX_ERR_INVALID, "null buffer");

    uint8_t* base = nullptr;
    TREX_TRY(scratch_ptr(&base));
    uint32_t* ticket = reinterpret_cast<uint32_t*>(base + kOffTicket);
    int64_t* first = reinterpret_cast<int64_t*>(base + kOffFirst);
    int64_t* cbegin = reinterpret_cast<int64_t*>(base + kOffBegin);
    int64_t* cend = reinterpret_cast<int64_t*>(base + kOffEnd);
    uint64_t* agg = reinterpret_cast<uint64_t*>(base + kOffAgg);
    int32_t* csid = reinterpret_cast<int32_t*>(base + kOffSid);
    int32_t* cpred = reinterpret_cast<int32_t*>(base + kOffPred);
    uint32_t* cflags = reinterpret_cast<uint32_t*>(base + kOffFlags);
    uint32_t* status = reinterpret_cast<uint32_t*>(base + kOffStatus);
    int32_t* jump = reinterpret_cast<int32_t*>(base + kOffJump);

    // 1. chunk counts and first chunks per string; ticket and error word cleared
    {
        void* a[] = {&d_begin, &d_end, &n_strings, &bytes_len, &chunk_len, &n_bound, &first, &ticket, &d_error};
        TREX_TRY(launch("chunk_table_count", 1, 1024, 0, s, a));
    }
    // 2. the chunk table and cleared status words
    {
        void* a[] = {&d_begin, &d_end, &n_strings, &chunk_len, &n_bound, &first, &d_error,
                     &cbegin, &cend, &csid, &cpred, &cflags, &status};
        const int64_t blocks = std::min<int64_t>((n_bound + kBlock - 1) / kBlock, 1024);
        TREX_TRY(launch("chunk_table_fill", (unsigned)blocks, kBlock, 0, s, a));
    }
    // 3. the look-back launch over n_bound chunk slots (as trex_dfa_match_plan)
    uint64_t max_ticks = device_wait_ticks();
    const bool window = h->window_base >= 0;
    uint32_t win4 = window ? (uint32_t)h->window_base * 0x01010101u : 0u;
    uint32_t start = (uint32_t)h->start;
    uint32_t dead = h->dead >= 0 ? (uint32_t)h->dead : 0u;
    uint32_t n_states = (uint32_t)h->n_states;
    uint8_t* t2 = window ? h->d_t2 : h->d_t1;
    uint8_t* f2 = window ? h->d_f2 : h->d_t1;
    const int64_t* tile_off = nullptr;   // unused without an image
    uint32_t hold = 0;
    int64_t n_chunks = n_bound;
    void* args[] = {&d_bytes, &tile_off, &cbegin, &cend, &csid, &cpred, &cflags, &n_chunks, &t2, &h->d_t1, &f2,
                    &h->d_acc, &win4, &start, &dead, &n_states, &status, &agg, &jump, &ticket, &d_error,
                    &max_spins, &max_ticks, &hold, &d_final_state, &d_accept};
    const int64_t waves_needed = (n_chunks + 63) / 64;
    const int64_t blocks = std::min<int64_t>((waves_needed + 3) / 4, (int64_t)device_cu_count() * 8);
    const char* kernel = window ? (kScanChains == 2 ? "dfa16_chunks_window_c2" : "dfa16_chunks_window")
                                : (kScanChains == 2 ? "dfa16_chunks_bytes_c2" : "dfa16_chunks_bytes");
    return launch(kernel, (unsigned)std::max<int64_t>(blocks, 1), kBlock, 0, s, args);
}

}  // extern "C"
