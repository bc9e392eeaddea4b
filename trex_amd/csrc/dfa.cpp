// DFA byte-table engine: handle creation (table analysis and expansion),
// dispatch, and the reference-compatible prefix_engine_* / monoid_batch_engine_*
// entry points.

#include <hip/hip_runtime_api.h>
#include <string.h>

#include <algorithm>
#include <cstdio>
#include <mutex>
#include <vector>

#include "runtime_internal.h"
#include "trex_hip.h"

using namespace trex;

namespace {

// Finds the absorbing non-accepting state, or -1.
int find_dead(const uint8_t* delta, const uint8_t* accept, int n) {
    for (int d = 0; d < n; ++d) {
        if (accept[d]) continue;
        bool absorbing = true;
        for (int b = 0; b < 256 && absorbing; ++b) absorbing = delta[d * 256 + b] == d;
        if (absorbing) return d;
    }
    return -1;
}

// Aligned 32-byte window containing every byte that does not send all states
// to `dead`; -1 if the live bytes span more than one window.
int find_window(const uint8_t* delta, int n, int dead) {
    if (dead < 0) return -1;
    int window = -1;
    for (int b = 0; b < 256; ++b) {
        bool kills = true;
        for (int s = 0; s < n && kills; ++s) kills = delta[s * 256 + b] == dead;
        if (kills) continue;
        int w = b & ~31;
        if (window >= 0 && window != w) return -1;
        window = w;
    }
    return window < 0 ? 0 : window;
}

// States a window walk can be in, and whether they synchronise: the states
// other than `dead`, plus `dead` when a window byte sends another state to
// it. Bytes of the window that send every such state to `dead` are left out
// (they synchronise anything but do not occur in matched text). Sets
// h->synchronizes, and h->rot_init: the identity on those states, every
// other slot (absent states, and `dead` when no window byte reaches it and
// it is not the start) started at the start state, so that the rotated
// kernel's composed function turns constant once those states merge.
void analyse_window(trex_dfa* h, const uint8_t* delta) {
    const int n = h->n_states, dead = h->dead, base = h->window_base;
    std::vector<int> live_bytes;
    for (int b = base; b < base + 32; ++b) {
        bool kills = true;
        for (int s = 0; s < n && kills; ++s) kills = s == dead || delta[s * 256 + b] == dead;
        if (!kills) live_bytes.push_back(b);
    }
    // a byte that kills every state makes the composed function constant
    // (dead) whatever its slots hold, so only live bytes decide whether the
    // dead state's slot is needed
    bool dead_reachable = false;
    for (int b : live_bytes)
        for (int s = 0; s < n; ++s) dead_reachable = dead_reachable || (s != dead && delta[s * 256 + b] == dead);
    const bool dead_absent = dead >= 0 && !dead_reachable && h->start != dead;
    std::vector<int> states;
    for (int s = 0; s < n; ++s)
        if (!(s == dead && dead_absent)) states.push_back(s);
    uint8_t init[16];
    for (int s = 0; s < 16; ++s) init[s] = (uint8_t)h->start;
    for (int s : states) init[s] = (uint8_t)s;
    memcpy(h->rot_init, init, 16);
    // pairs that some word sends to one state: backward closure from the diagonal
    bool good[16][16] = {};
    for (int s = 0; s < 16; ++s) good[s][s] = true;
    for (bool changed = true; changed;) {
        changed = false;
        for (int p : states)
            for (int q : states) {
                if (good[p][q]) continue;
                for (int b : live_bytes)
                    if (good[delta[p * 256 + b]][delta[q * 256 + b]]) {
                        good[p][q] = good[q][p] = true;
                        changed = true;
                        break;
                    }
            }
    }
    bool all = true;
    for (int p : states)
        for (int q : states) all = all && good[p][q];
    h->synchronizes = all || live_bytes.empty();
}

bool variant_allowed(const trex_dfa* h, int v) {
    switch (v) {
        case TREX_DFA_K1_GENERIC: return true;
        case TREX_DFA_K1_16: return h->n_states <= 16;
        case TREX_DFA_K2_WINDOW:
        case TREX_DFA_K2_WINDOW_F:
        case TREX_DFA_K2_WINDOW_FR: return h->n_states <= 16 && h->window_base >= 0 && h->d_t2 != nullptr;
        default: return false;
    }
}

}  // namespace

extern "C" {

int trex_dfa_create(const uint8_t* delta, const uint8_t* accept, int n_states, int start, trex_dfa** out) {
    if (!delta || !accept || !out) return fail(TREX_ERR_INVALID, "null argument");
    *out = nullptr;
    if (n_states < 1 || n_states > 256) return fail(TREX_ERR_INVALID, "n_states=%d outside 1..256", n_states);
    if (start < 0 || start >= n_states) return fail(TREX_ERR_INVALID, "start=%d outside 0..%d", start, n_states - 1);
    for (int i = 0; i < n_states * 256; ++i)
        if (delta[i] >= n_states) return fail(TREX_ERR_INVALID, "delta[%d]=%d names a state >= %d", i, delta[i], n_states);
    TREX_TRY(ensure_module());

    trex_dfa* h = new trex_dfa();
    h->n_states = n_states;
    h->start = start;
    h->stride = std::max(16, (n_states + 15) / 16 * 16);
    h->dead = find_dead(delta, accept, n_states);
    h->window_base = n_states <= 16 ? find_window(delta, n_states, h->dead) : -1;

    const int stride = h->stride;
    std::vector<uint8_t> t1((size_t)256 * stride), acc((size_t)stride, 0);
    for (int b = 0; b < 256; ++b)
        for (int s = 0; s < stride; ++s)
            t1[(size_t)b * stride + s] = s < n_states ? delta[s * 256 + b] : (uint8_t)s;
    for (int s = 0; s < n_states; ++s) acc[s] = accept[s] ? 1 : 0;

    int rc = upload((void**)&h->d_t1, t1.data(), t1.size());
    if (!rc) rc = upload((void**)&h->d_acc, acc.data(), acc.size());
    if (!rc && h->window_base >= 0) {
        std::vector<uint8_t> t2(1024 * 16);
        std::vector<uint64_t> f2(1024, 0);
        const int base = h->window_base;
        for (int i = 0; i < 1024; ++i) {
            const int b0 = base + (i & 31), b1 = base + (i >> 5);
            for (int s = 0; s < 16; ++s) {
                const uint8_t nxt = s < n_states ? delta[delta[s * 256 + b0] * 256 + b1] : (uint8_t)s;
                t2[i * 16 + s] = nxt;
                f2[i] |= (uint64_t)nxt << (4 * s);
            }
        }
        rc = upload((void**)&h->d_t2, t2.data(), t2.size());
        if (!rc) rc = upload((void**)&h->d_f2, f2.data(), f2.size() * sizeof(uint64_t));
    }
    if (rc) { trex_dfa_destroy(h); return rc; }
    if (h->window_base >= 0) analyse_window(h, delta);
    if (h->window_base >= 0) h->auto_variant = TREX_DFA_K2_WINDOW_F;
    else if (n_states <= 16) h->auto_variant = TREX_DFA_K1_16;
    else h->auto_variant = TREX_DFA_K1_GENERIC;
    h->variant = h->auto_variant;
    *out = h;
    return TREX_OK;
}

int trex_dfa_destroy(trex_dfa* h) {
    if (!h) return TREX_OK;
    if (h->stream) hipStreamSynchronize(h->stream);
    hipFree(h->d_t1);
    hipFree(h->d_acc);
    hipFree(h->d_t2);
    hipFree(h->d_f2);
    h->in_bytes.release();
    h->in_offsets.release();
    h->out_state.release();
    h->out_accept.release();
    for (auto& e : h->ev) if (e) hipEventDestroy(e);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
    return TREX_OK;
}

int trex_dfa_info(const trex_dfa* h, int* variant, int* window_base, int* dead_state) {
    if (!h) return fail(TREX_ERR_INVALID, "null handle");
    if (variant) *variant = h->variant;
    if (window_base) *window_base = h->window_base;
    if (dead_state) *dead_state = h->dead;
    return TREX_OK;
}

int trex_dfa_set_variant(trex_dfa* h, int variant) {
    if (!h) return fail(TREX_ERR_INVALID, "null handle");
    if (variant == TREX_DFA_AUTO) { h->variant = h->auto_variant; return TREX_OK; }
    if (!variant_allowed(h, variant)) return fail(TREX_ERR_INVALID, "variant %d not usable for this table", variant);
    h->variant = variant;
    return TREX_OK;
}

int trex_dfa_set_rotation(trex_dfa* h, int mode) {
    if (!h) return fail(TREX_ERR_INVALID, "null handle");
    if (mode < -1 || mode > 1) return fail(TREX_ERR_INVALID, "rotation mode %d: -1 (auto), 0 or 1", mode);
    h->rotation = mode;
    return TREX_OK;
}

int trex_dfa_rotation(const trex_dfa* h, int* active, int* synchronizes) {
    if (!h) return fail(TREX_ERR_INVALID, "null handle");
    const bool usable = h->window_base >= 0 && h->n_states <= 16;
    if (active) *active = usable && (h->rotation == 1 || (h->rotation == -1 && h->synchronizes)) ? 1 : 0;
    if (synchronizes) *synchronizes = h->synchronizes ? 1 : 0;
    return TREX_OK;
}

int trex_dfa_set_prefetch(trex_dfa* h, int vectors) {
    if (!h) return fail(TREX_ERR_INVALID, "null handle");
    if (vectors != 8 && vectors != 16) return fail(TREX_ERR_INVALID, "prefetch must be 8 or 16 vectors");
    h->prefetch = vectors;
    return TREX_OK;
}

namespace {

// One launch of the handle's variant over begin/end ranges; `il` selects the
// kernels of the interleaved layout.
int dfa_dispatch(trex_dfa* h, const uint8_t* d_bytes, const int64_t* d_begin, const int64_t* d_end,
                 int64_t n_strings, uint8_t* d_final_state, uint8_t* d_accept, void* stream, bool il) {
    if (!h) return fail(TREX_ERR_INVALID, "null handle");
    if (n_strings < 0) return fail(TREX_ERR_INVALID, "n_strings < 0");
    if (n_strings == 0) return TREX_OK;
    if (!d_bytes || !d_begin || !d_end || !d_final_state || !d_accept) return fail(TREX_ERR_INVALID, "null buffer");
    const int64_t grid64 = (n_strings + kBlock - 1) / kBlock;
    if (grid64 > 0x7fffffff) return fail(TREX_ERR_CAPACITY, "too many strings");
    const unsigned grid = (unsigned)grid64;
    hipStream_t s = (hipStream_t)stream;
    uint32_t start = (uint32_t)h->start;
    const char* il_tag = il ? "_il" : "";
    char name[48];
    if (h->variant == TREX_DFA_K2_WINDOW || h->variant == TREX_DFA_K2_WINDOW_F ||
        h->variant == TREX_DFA_K2_WINDOW_FR) {
        uint32_t win4 = (uint32_t)h->window_base * 0x01010101u;
        uint32_t dead = (uint32_t)h->dead;
        uint8_t* table = h->variant == TREX_DFA_K2_WINDOW ? h->d_t2 : h->d_f2;
        const char* form = h->variant == TREX_DFA_K2_WINDOW ? "k2w" : h->variant == TREX_DFA_K2_WINDOW_F ? "k2f" : "k2fr";
        int rot = 0;
        trex_dfa_rotation(h, &rot, nullptr);
        if (!il && rot && h->variant == TREX_DFA_K2_WINDOW_F && h->prefetch == 8) {
            // contiguous strings, read order rotated per wave (spreads HBM channels)
            void* args[] = {&d_bytes, &d_begin, &d_end, &n_strings, &h->d_f2, &h->d_t2, &h->d_t1, &h->d_acc,
                            &win4, &start, &dead, &h->rot_init[0], &h->rot_init[1], &h->rot_init[2],
                            &h->rot_init[3], &d_final_state, &d_accept};
            return launch("dfa16_k2f_rot_pf8", grid, kBlock, 0, s, args);
        }
        std::snprintf(name, sizeof(name), "dfa16_%s%s_pf%d", form, il_tag, h->prefetch == 16 ? 16 : 8);
        void* args[] = {&d_bytes, &d_begin, &d_end, &n_strings, &table, &h->d_t1, &h->d_acc,
                        &win4, &start, &dead, &d_final_state, &d_accept};
        return launch(name, grid, kBlock, 0, s, args);
    }
    uint32_t stride = (uint32_t)h->stride;
    void* args[] = {&d_bytes, &d_begin, &d_end, &n_strings, &h->d_t1, &h->d_acc, &stride, &start,
                    &d_final_state, &d_accept};
    const unsigned lds = 256u * stride + stride;
    std::snprintf(name, sizeof(name), "%s%s", h->variant == TREX_DFA_K1_16 ? "dfa16_k1" : "dfa_k1", il_tag);
    return launch(name, grid, kBlock, lds, s, args);
}

}  // namespace

int trex_dfa_match_ranges(trex_dfa* h, const uint8_t* d_bytes, const int64_t* d_begin, const int64_t* d_end,
                          int64_t n_strings, uint8_t* d_final_state, uint8_t* d_accept, void* stream) {
    return dfa_dispatch(h, d_bytes, d_begin, d_end, n_strings, d_final_state, d_accept, stream, false);
}

int trex_dfa_match_interleaved(trex_dfa* h, const uint8_t* d_bytes, const int64_t* d_begin, const int64_t* d_end,
                               int64_t n_strings, uint8_t* d_final_state, uint8_t* d_accept, void* stream) {
    return dfa_dispatch(h, d_bytes, d_begin, d_end, n_strings, d_final_state, d_accept, stream, true);
}

int trex_dfa_match(trex_dfa* h, const uint8_t* d_bytes, const int64_t* d_offsets, int64_t n_strings,
                   uint8_t* d_final_state, uint8_t* d_accept, void* stream) {
    if (!d_offsets) return fail(TREX_ERR_INVALID, "null offsets");
    return trex_dfa_match_ranges(h, d_bytes, d_offsets, d_offsets + 1, n_strings, d_final_state, d_accept, stream);
}

int trex_dfa_match_host(trex_dfa* h, const uint8_t* bytes, const int64_t* offsets, int64_t n_strings,
                        uint8_t* final_state, uint8_t* accept, float* kernel_ms, float* total_ms) {
    if (!h) return fail(TREX_ERR_INVALID, "null handle");
    if (kernel_ms) *kernel_ms = 0.f;
    if (total_ms) *total_ms = 0.f;
    if (n_strings <= 0) return n_strings < 0 ? fail(TREX_ERR_INVALID, "n_strings < 0") : TREX_OK;
    if (!offsets || !final_state || !accept) return fail(TREX_ERR_INVALID, "null buffer");
    const int64_t base = offsets[0], total = offsets[n_strings] - offsets[0];
    if (total < 0) return fail(TREX_ERR_INVALID, "offsets decrease");
    if (total > 0 && !bytes) return fail(TREX_ERR_INVALID, "null bytes");
    for (int64_t i = 0; i < n_strings; ++i)
        if (offsets[i + 1] < offsets[i]) return fail(TREX_ERR_INVALID, "offsets decrease at %lld", (long long)i);

    std::lock_guard<std::mutex> lock(h->mu);
    TREX_TRY(h->in_bytes.reserve((size_t)total + 16));
    TREX_TRY(h->in_offsets.reserve(sizeof(int64_t) * (size_t)(n_strings + 1)));
    TREX_TRY(h->out_state.reserve((size_t)n_strings));
    TREX_TRY(h->out_accept.reserve((size_t)n_strings));
    // Offsets are rebased so the device batch starts at byte 0.
    std::vector<int64_t> rebased((size_t)n_strings + 1);
    for (int64_t i = 0; i <= n_strings; ++i) rebased[i] = offsets[i] - base;

    TREX_TRY(ensure_host_dispatch(h->stream, h->ev));
    hipStream_t s = h->stream;
    TREX_HIP(hipEventRecord(h->ev[0], s), "hipEventRecord");
    if (total > 0) TREX_HIP(hipMemcpyAsync(h->in_bytes.ptr, bytes + base, (size_t)total, hipMemcpyHostToDevice, s), "H2D bytes");
    TREX_HIP(hipMemcpyAsync(h->in_offsets.ptr, rebased.data(), rebased.size() * sizeof(int64_t), hipMemcpyHostToDevice, s), "H2D offsets");
    TREX_HIP(hipEventRecord(h->ev[1], s), "hipEventRecord");
    TREX_TRY(trex_dfa_match(h, h->in_bytes.as<uint8_t>(), h->in_offsets.as<int64_t>(), n_strings,
                            h->out_state.as<uint8_t>(), h->out_accept.as<uint8_t>(), s));
    TREX_HIP(hipEventRecord(h->ev[2], s), "hipEventRecord");
    TREX_HIP(hipMemcpyAsync(final_state, h->out_state.ptr, (size_t)n_strings, hipMemcpyDeviceToHost, s), "D2H state");
    TREX_HIP(hipMemcpyAsync(accept, h->out_accept.ptr, (size_t)n_strings, hipMemcpyDeviceToHost, s), "D2H accept");
    TREX_HIP(hipEventRecord(h->ev[3], s), "hipEventRecord");
    TREX_HIP(hipEventSynchronize(h->ev[3]), "hipEventSynchronize");
    float k = 0.f, t = 0.f;
    hipEventElapsedTime(&k, h->ev[1], h->ev[2]);
    hipEventElapsedTime(&t, h->ev[0], h->ev[3]);
    if (kernel_ms) *kernel_ms = k;
    if (total_ms) *total_ms = t;
    return TREX_OK;
}

int trex_copy(void* d_dst, const void* d_src, size_t n_bytes, void* stream) {
    if (n_bytes % 16 != 0) return fail(TREX_ERR_INVALID, "copy size must be a multiple of 16");
    if (((uintptr_t)d_dst | (uintptr_t)d_src) & 15) return fail(TREX_ERR_INVALID, "copy buffers must be 16-byte aligned");
    int64_t n16 = (int64_t)(n_bytes / 16);
    const int64_t want = (n16 + kBlock - 1) / kBlock;
    const unsigned grid = (unsigned)std::min<int64_t>(want, 8192);
    void* args[] = {&d_src, &d_dst, &n16};
    return launch("copy_u4", grid, kBlock, 0, (hipStream_t)stream, args);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Reference-compatible singletons
// ---------------------------------------------------------------------------
namespace {

struct CompatEngine {
    trex_dfa* h = nullptr;
    int max_total_chars = 0;
    int max_batch = 0;
    int empty_accept = 0;      // accept bit of the start state, for empty strings
    std::vector<int64_t> offs;
    std::vector<uint8_t> st, acc;
    void reset() { trex_dfa_destroy(h); h = nullptr; }
};

std::mutex g_compat_mu;
CompatEngine g_prefix, g_monoid;

int compat_init(CompatEngine& eng, const std::vector<uint8_t>& delta, const std::vector<uint8_t>& accept, int n,
                int start, int max_total_chars, int max_batch) {
    eng.reset();
    TREX_TRY(trex_dfa_create(delta.data(), accept.data(), n, start, &eng.h));
    eng.max_total_chars = max_total_chars;
    eng.max_batch = max_batch;
    eng.empty_accept = accept[start] ? 1 : 0;
    return TREX_OK;
}

int compat_dispatch(CompatEngine& eng, const uint8_t* raw_concat, const int* offsets, int* results, int B,
                    int total_chars, float* kernel_ms, float* total_ms) {
    if (!eng.h) return TREX_ERR_NO_DEVICE;  // reference: -1 when not initialised
    if (B < 0 || total_chars < 0) return fail(TREX_ERR_INVALID, "negative size");
    if (B > eng.max_batch || total_chars > eng.max_total_chars)
        return fail(TREX_ERR_CAPACITY, "batch of %d strings / %d chars exceeds init capacity %d / %d", B, total_chars,
                    eng.max_batch, eng.max_total_chars);
    if (B == 0) { if (kernel_ms) *kernel_ms = 0.f; if (total_ms) *total_ms = 0.f; return TREX_OK; }
    eng.offs.resize((size_t)B + 1);
    for (int i = 0; i <= B; ++i) eng.offs[i] = offsets[i];
    if (eng.offs[B] > total_chars) return fail(TREX_ERR_INVALID, "offsets[B] > total_chars");
    eng.st.resize((size_t)B);
    eng.acc.resize((size_t)B);
    TREX_TRY(trex_dfa_match_host(eng.h, raw_concat, eng.offs.data(), B, eng.st.data(), eng.acc.data(), kernel_ms,
                                 total_ms));
    for (int i = 0; i < B; ++i) results[i] = eng.acc[i];
    return TREX_OK;
}

// The prefix (long-string) dispatch: for tables of at most 16 states every
// string is cut into chunks chained by the decoupled look-back (one launch;
// chunk length the power of two nearest total / 65,536 within 1-64 KiB, about
// one wave of 64 chunks per SIMD); wider tables walk each string on one lane.
int compat_dispatch_prefix(CompatEngine& eng, const uint8_t* raw_concat, const int* offsets, int* results, int B,
                           int total_chars, float* kernel_ms, float* total_ms) {
    if (!eng.h) return TREX_ERR_NO_DEVICE;
    if (eng.h->n_states > 16 || B <= 0) return compat_dispatch(eng, raw_concat, offsets, results, B, total_chars,
                                                               kernel_ms, total_ms);
    if (total_chars < 0) return fail(TREX_ERR_INVALID, "negative size");
    if (B > eng.max_batch || total_chars > eng.max_total_chars)
        return fail(TREX_ERR_CAPACITY, "batch of %d strings / %d chars exceeds init capacity %d / %d", B, total_chars,
                    eng.max_batch, eng.max_total_chars);
    if (!offsets || !results) return fail(TREX_ERR_INVALID, "null buffer");
    const int64_t base = offsets[0], total = (int64_t)offsets[B] - base;
    if (offsets[B] > total_chars || total < 0) return fail(TREX_ERR_INVALID, "offsets out of range");
    if (total > 0 && !raw_concat) return fail(TREX_ERR_INVALID, "null bytes");
    std::vector<int64_t> begin((size_t)B), end((size_t)B);
    for (int i = 0; i < B; ++i) {
        if (offsets[i + 1] < offsets[i]) return fail(TREX_ERR_INVALID, "offsets decrease at %d", i);
        begin[i] = offsets[i] - base;
        end[i] = offsets[i + 1] - base;
    }
    int64_t chunk = 1024;
    while (chunk < 65536 && (double)total / 65536.0 >= (double)chunk * 1.41421356) chunk <<= 1;
    trex_plan* plan = nullptr;
    TREX_TRY(trex_plan_create(begin.data(), end.data(), B, chunk, &plan));
    trex_dfa* h = eng.h;
    int rc = TREX_OK;
    {
        std::lock_guard<std::mutex> lock(h->mu);
        rc = h->in_bytes.reserve((size_t)total + 16);
        if (!rc) rc = h->out_state.reserve((size_t)B);
        if (!rc) rc = h->out_accept.reserve((size_t)B);
        if (!rc) rc = ensure_host_dispatch(h->stream, h->ev);
        hipStream_t s = h->stream;
        eng.st.resize((size_t)B);
        eng.acc.resize((size_t)B);
        hipError_t e = hipSuccess;
        if (!rc) e = hipEventRecord(h->ev[0], s);
        if (!rc && e == hipSuccess && total > 0)
            e = hipMemcpyAsync(h->in_bytes.ptr, raw_concat + base, (size_t)total, hipMemcpyHostToDevice, s);
        if (!rc && e == hipSuccess) e = hipEventRecord(h->ev[1], s);
        if (!rc && e != hipSuccess) rc = hip_fail(e, "prefix dispatch upload");
        if (!rc) rc = trex_dfa_match_plan(h, plan, h->in_bytes.as<uint8_t>(), h->out_state.as<uint8_t>(),
                                          h->out_accept.as<uint8_t>(), 0, s);
        if (!rc) e = hipEventRecord(h->ev[2], s);
        if (!rc && e == hipSuccess)
            e = hipMemcpyAsync(eng.acc.data(), h->out_accept.ptr, (size_t)B, hipMemcpyDeviceToHost, s);
        if (!rc && e == hipSuccess) e = hipEventRecord(h->ev[3], s);
        if (!rc && e == hipSuccess) e = hipEventSynchronize(h->ev[3]);
        if (!rc && e != hipSuccess) rc = hip_fail(e, "prefix dispatch");
        if (!rc) {
            float k = 0.f, t = 0.f;
            hipEventElapsedTime(&k, h->ev[1], h->ev[2]);
            hipEventElapsedTime(&t, h->ev[0], h->ev[3]);
            if (kernel_ms) *kernel_ms = k;
            if (total_ms) *total_ms = t;
        }
    }
    uint32_t err = 0;
    if (!rc) rc = trex_plan_error(plan, &err, h->stream);
    trex_plan_destroy(plan);
    if (rc) return rc;
    if (err) return fail(TREX_ERR_TIMEOUT, "look-back launch reported error word %u", err);
    for (int i = 0; i < B; ++i) results[i] = eng.acc[i];
    return TREX_OK;
}

}  // namespace

extern "C" {

int prefix_engine_device_check(void) { return trex_device_check(); }

int prefix_engine_init(const uint8_t* tmap, const uint8_t* accept, int start_state, int N, int K,
                       int max_total_chars, int max_batch) {
    (void)K;
    if (!tmap || !accept) return fail(TREX_ERR_INVALID, "null table");
    if (N < 1 || N > 256) return fail(TREX_ERR_INVALID, "N=%d outside 1..256", N);
    std::vector<uint8_t> delta((size_t)N * 256), acc(accept, accept + N);
    for (int b = 0; b < 256; ++b)
        for (int s = 0; s < N; ++s) delta[(size_t)s * 256 + b] = tmap[b * N + s];
    std::lock_guard<std::mutex> lock(g_compat_mu);
    return compat_init(g_prefix, delta, acc, N, start_state, max_total_chars, max_batch);
}

int prefix_engine_dispatch(const uint8_t* raw_concat, const int* offsets, int* results, int B, int total_chars,
                           float* kernel_ms, float* total_ms) {
    std::lock_guard<std::mutex> lock(g_compat_mu);
    return compat_dispatch(g_prefix, raw_concat, offsets, results, B, total_chars, kernel_ms, total_ms);
}

void prefix_engine_destroy(void) {
    std::lock_guard<std::mutex> lock(g_compat_mu);
    g_prefix.reset();
}

int monoid_batch_engine_device_check(void) { return trex_device_check(); }

int monoid_batch_engine_init(int M, int sigma_ext, int identity, const uint8_t* char_compose,
                             const uint8_t* raw_char_map, const uint8_t* accept, const uint8_t* monoid_compose,
                             int max_total_chars, int max_batch) {
    (void)monoid_compose;
    if (!char_compose || !raw_char_map || !accept) return fail(TREX_ERR_INVALID, "null table");
    if (M < 1 || M > 256 || sigma_ext < 1) return fail(TREX_ERR_INVALID, "M=%d sigma_ext=%d", M, sigma_ext);
    std::vector<uint8_t> delta((size_t)M * 256), acc(accept, accept + M);
    for (int m = 0; m < M; ++m)
        for (int b = 0; b < 256; ++b) {
            int c = raw_char_map[b];
            if (c >= sigma_ext) return fail(TREX_ERR_INVALID, "raw_char_map[%d]=%d >= sigma_ext", b, c);
            delta[(size_t)m * 256 + b] = char_compose[m * sigma_ext + c];
        }
    std::lock_guard<std::mutex> lock(g_compat_mu);
    return compat_init(g_monoid, delta, acc, M, identity, max_total_chars, max_batch);
}

int monoid_batch_engine_set_kgram(const uint8_t* fused_compose, const uint8_t* kgram_compose, int kgram_k,
                                  int sigma_k, int sigma) {
    (void)fused_compose; (void)kgram_compose; (void)kgram_k; (void)sigma_k; (void)sigma;
    std::lock_guard<std::mutex> lock(g_compat_mu);
    return g_monoid.h ? TREX_OK : TREX_ERR_NO_DEVICE;
}

int monoid_batch_engine_dispatch(const uint8_t* raw_concat, const int* offsets, int* results, int B,
                                 int total_chars, float* kernel_ms, float* total_ms) {
    std::lock_guard<std::mutex> lock(g_compat_mu);
    return compat_dispatch(g_monoid, raw_concat, offsets, results, B, total_chars, kernel_ms, total_ms);
}

int monoid_batch_engine_dispatch_prefix(const uint8_t* raw_concat, const int* offsets, int* results, int B,
                                        int total_chars, float* kernel_ms, float* total_ms) {
    std::lock_guard<std::mutex> lock(g_compat_mu);
    return compat_dispatch_prefix(g_monoid, raw_concat, offsets, results, B, total_chars, kernel_ms, total_ms);
}

void monoid_batch_engine_destroy(void) {
    std::lock_guard<std::mutex> lock(g_compat_mu);
    g_monoid.reset();
}

}  // extern "C"
