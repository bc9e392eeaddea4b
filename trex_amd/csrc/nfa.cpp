// Boolean NFA-64 engine: handle creation (nibble tables), dispatch, and the
// reference-compatible nfa_tc_engine_* entry points.

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "runtime_internal.h"
#include "trex_hip.h"

using namespace trex;

namespace {
constexpr int kNfaBlock = 512;
constexpr int kNfa5Block = 1024;   // nfa64_window5: one 1024-thread workgroup per CU holds the 104 KB of tables
constexpr int kMaxSymbols = 62;   // (62 + 2) classes x 2 KB of nibble tables fit one workgroup's LDS
}

namespace {

// Nibble tables of one class from its 64 successor masks.
void nibble_tables(const uint64_t* succ64, uint64_t* out /*[16][16]*/) {
    for (int p = 0; p < 16; ++p)
        for (int v = 0; v < 16; ++v) {
            uint64_t m = 0;
            for (int i = 0; i < 4; ++i)
                if (v & (1 << i)) m |= succ64[4 * p + i];
            out[p * 16 + v] = m;
        }
}

// LDS placement of the tables: class k stores entry (p, v) at row p ^ h and
// slot v ^ x with h = (k >> 4) & 1 and x = k & 15, so lanes that look up the
// same nibble value in different classes hit different banks (a dense state
// set makes v = 15 far more frequent than the other values).
std::vector<uint64_t> swizzled(const std::vector<uint64_t>& tab) {
    std::vector<uint64_t> out(tab.size());
    const size_t n_classes = tab.size() / 256;
    for (size_t k = 0; k < n_classes; ++k) {
        const unsigned h = (unsigned)(k >> 4) & 1u, x = (unsigned)k & 15u;
        for (unsigned p = 0; p < 16; ++p)
            for (unsigned v = 0; v < 16; ++v) out[k * 256 + (p ^ h) * 16 + (v ^ x)] = tab[k * 256 + p * 16 + v];
    }
    return out;
}

// 5-bit slice tables of the 32 window classes (layout of nfa64_window5):
// class k at byte k * kClass5, entry (r, v) = union of succ[5r + i] over the
// bits i of v at byte r * 256 + v * 8 of it; row 12 holds states 60..63 (its
// entries v >= 16 repeat v & 15 and are never read).
constexpr size_t kClass5 = 13 * 256 + 8;

std::vector<uint8_t> slice5_tables(const std::vector<const uint64_t*>& succ_of_class) {
    std::vector<uint8_t> out(32 * kClass5, 0);
    for (int k = 0; k < 32; ++k) {
        const uint64_t* succ64 = succ_of_class[k];
        for (int r = 0; r < 13; ++r)
            for (int v = 0; v < 32; ++v) {
                uint64_t m = 0;
                for (int i = 0; i < 5; ++i)
                    if ((v & (1 << i)) && 5 * r + i < 64) m |= succ64[5 * r + i];
                std::memcpy(out.data() + k * kClass5 + r * 256 + v * 8, &m, 8);
            }
    }
    return out;
}

// 6/5-bit slice tables (layout of nfa64_window6): each 32-bit half of the
// state cut into fields of 6, 6, 5, 5, 5, 5 bits; class k at byte k * kClass6,
// field r's row at byte kRow6[r] of it, entry v = union of succ[b_r + i] over
// the bits i of v.
constexpr size_t kClass6 = 4096 + 8;
constexpr int kBits6[12] = {6, 6, 5, 5, 5, 5, 6, 6, 5, 5, 5, 5};
constexpr int kFirst6[12] = {0, 6, 12, 17, 22, 27, 32, 38, 44, 49, 54, 59};
constexpr size_t kRow6[12] = {0, 512, 1024, 1280, 1536, 1792, 2048, 2560, 3072, 3328, 3584, 3840};

std::vector<uint8_t> slice6_tables(const std::vector<const uint64_t*>& succ_of_class) {
    std::vector<uint8_t> out(32 * kClass6, 0);
    for (int k = 0; k < 32; ++k) {
        const uint64_t* succ64 = succ_of_class[k];
        for (int r = 0; r < 12; ++r)
            for (int v = 0; v < (1 << kBits6[r]); ++v) {
                uint64_t m = 0;
                for (int i = 0; i < kBits6[r]; ++i)
                    if (v & (1 << i)) m |= succ64[kFirst6[r] + i];
                std::memcpy(out.data() + k * kClass6 + kRow6[r] + v * 8, &m, 8);
            }
    }
    return out;
}

}  // namespace

extern "C" {

int trex_nfa_create(const uint64_t* succ, int n_symbols, const int16_t* byte_to_sym, uint64_t start,
                    uint64_t accept, trex_nfa** out) {
    if (!out || !byte_to_sym || (n_symbols > 0 && !succ)) return fail(TREX_ERR_INVALID, "null argument");
    *out = nullptr;
    if (n_symbols < 0 || n_symbols > kMaxSymbols)
        return fail(TREX_ERR_INVALID, "n_symbols=%d outside 0..%d", n_symbols, kMaxSymbols);
    for (int b = 0; b < 256; ++b)
        if (byte_to_sym[b] < -2 || byte_to_sym[b] >= n_symbols)
            return fail(TREX_ERR_INVALID, "byte_to_sym[%d]=%d invalid", b, byte_to_sym[b]);
    TREX_TRY(ensure_module());

    trex_nfa* h = new trex_nfa();
    h->n_symbols = n_symbols;
    h->start = start;
    h->accept = accept;
    const int reject = n_symbols, skip = n_symbols + 1;
    h->n_classes = n_symbols + 2;
    std::vector<uint64_t> ident(64), none(64, 0);
    for (int j = 0; j < 64; ++j) ident[j] = 1ull << j;
    std::vector<uint64_t> tab((size_t)h->n_classes * 256);
    for (int c = 0; c < n_symbols; ++c) nibble_tables(succ + (size_t)c * 64, tab.data() + (size_t)c * 256);
    nibble_tables(none.data(), tab.data() + (size_t)reject * 256);
    nibble_tables(ident.data(), tab.data() + (size_t)skip * 256);
    std::vector<uint8_t> cls(256);
    for (int b = 0; b < 256; ++b)
        cls[b] = (uint8_t)(byte_to_sym[b] == -1 ? reject : byte_to_sym[b] == -2 ? skip : byte_to_sym[b]);

    // window form: every byte outside one aligned 32-byte window rejects
    int window = -1;
    bool ok = true;
    for (int b = 0; b < 256 && ok; ++b) {
        if (byte_to_sym[b] == -1) continue;
        const int w = b & ~31;
        if (window >= 0 && window != w) ok = false;
        window = w;
    }
    if (ok) h->window_base = window < 0 ? 0 : window;

    const std::vector<uint64_t> tab_lds = swizzled(tab);
    int rc = upload((void**)&h->d_tab, tab_lds.data(), tab_lds.size() * 8);
    if (!rc) rc = upload((void**)&h->d_cls, cls.data(), cls.size());
    if (!rc && h->window_base >= 0) {
        h->n_classes_win = 34;
        std::vector<uint64_t> tw((size_t)34 * 256);
        std::vector<uint8_t> cw(256);
        for (int k = 0; k < 32; ++k) {
            const int src = cls[h->window_base + k];
            std::copy(tab.begin() + (size_t)src * 256, tab.begin() + (size_t)(src + 1) * 256, tw.begin() + (size_t)k * 256);
        }
        std::copy(tab.begin() + (size_t)reject * 256, tab.begin() + (size_t)(reject + 1) * 256, tw.begin() + 32 * 256);
        std::copy(tab.begin() + (size_t)skip * 256, tab.begin() + (size_t)(skip + 1) * 256, tw.begin() + 33 * 256);
        for (int b = 0; b < 256; ++b) cw[b] = (uint8_t)(((b & ~31) == h->window_base) ? (b & 31) : 32);
        const std::vector<uint64_t> tw_lds = swizzled(tw);
        rc = upload((void**)&h->d_tab_win, tw_lds.data(), tw_lds.size() * 8);
        if (!rc) rc = upload((void**)&h->d_cls_win, cw.data(), cw.size());
        if (!rc) {
            std::vector<const uint64_t*> src(32);
            for (int k = 0; k < 32; ++k) {
                const int c = cls[h->window_base + k];
                src[k] = c < n_symbols ? succ + (size_t)c * 64 : c == reject ? none.data() : ident.data();
            }
            const std::vector<uint8_t> t5 = slice5_tables(src);
            rc = upload((void**)&h->d_tab5, t5.data(), t5.size());
            if (!rc) {
                const std::vector<uint8_t> t6 = slice6_tables(src);
                rc = upload((void**)&h->d_tab6, t6.data(), t6.size());
            }
        }
    }
    if (rc) { trex_nfa_destroy(h); return rc; }
    h->variant = h->d_tab6 ? TREX_NFA_WINDOW6 : h->window_base >= 0 ? TREX_NFA_WINDOW : TREX_NFA_BYTES;
    *out = h;
    return TREX_OK;
}

int trex_nfa_destroy(trex_nfa* h) {
    if (!h) return TREX_OK;
    if (h->stream) hipStreamSynchronize(h->stream);
    hipFree(h->d_tab);
    hipFree(h->d_cls);
    hipFree(h->d_tab_win);
    hipFree(h->d_cls_win);
    hipFree(h->d_tab5);
    hipFree(h->d_tab6);
    h->in_bytes.release();
    h->in_offsets.release();
    h->out_mask.release();
    h->out_accept.release();
    for (auto& e : h->ev) if (e) hipEventDestroy(e);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
    return TREX_OK;
}

int trex_nfa_info(const trex_nfa* h, int* variant, int* window_base) {
    if (!h) return fail(TREX_ERR_INVALID, "null handle");
    if (variant) *variant = h->variant;
    if (window_base) *window_base = h->window_base;
    return TREX_OK;
}

int trex_nfa_set_variant(trex_nfa* h, int variant) {
    if (!h) return fail(TREX_ERR_INVALID, "null handle");
    if (variant == TREX_NFA_BYTES || (variant == TREX_NFA_WINDOW && h->window_base >= 0) ||
        (variant == TREX_NFA_WINDOW5 && h->d_tab5) || (variant == TREX_NFA_WINDOW6 && h->d_tab6)) {
        h->variant = variant;
        return TREX_OK;
    }
    return fail(TREX_ERR_INVALID, "variant %d not usable for this automaton", variant);
}

int trex_nfa_match_ranges(trex_nfa* h, const uint8_t* d_bytes, const int64_t* d_begin, const int64_t* d_end,
                          int64_t n_strings, uint64_t* d_final_mask, uint8_t* d_accept, void* stream) {
    if (!h) return fail(TREX_ERR_INVALID, "null handle");
    if (n_strings < 0) return fail(TREX_ERR_INVALID, "n_strings < 0");
    if (n_strings == 0) return TREX_OK;
    if (!d_bytes || !d_begin || !d_end || !d_final_mask || !d_accept) return fail(TREX_ERR_INVALID, "null buffer");
    if (h->variant == TREX_NFA_WINDOW5 || h->variant == TREX_NFA_WINDOW6) {
        const bool six = h->variant == TREX_NFA_WINDOW6;
        const int64_t grid5 = (n_strings + kNfa5Block - 1) / kNfa5Block;
        if (grid5 > 0x7fffffff) return fail(TREX_ERR_CAPACITY, "too many strings");
        uint8_t* tab5 = six ? h->d_tab6 : h->d_tab5;
        uint32_t win4 = (uint32_t)h->window_base * 0x01010101u;
        uint64_t start = h->start, accept = h->accept;
        void* args[] = {&d_bytes, &d_begin, &d_end, &n_strings, &tab5, &win4, &start, &accept, &d_final_mask, &d_accept};
        return launch(six ? "nfa64_window6" : "nfa64_window5", (unsigned)grid5, kNfa5Block, 0, (hipStream_t)stream,
                      args);
    }
    const int64_t grid64 = (n_strings + kNfaBlock - 1) / kNfaBlock;
    if (grid64 > 0x7fffffff) return fail(TREX_ERR_CAPACITY, "too many strings");
    const bool win = h->variant == TREX_NFA_WINDOW;
    uint64_t* tab = win ? h->d_tab_win : h->d_tab;
    uint8_t* cls = win ? h->d_cls_win : h->d_cls;
    int n_classes = win ? h->n_classes_win : h->n_classes;
    uint32_t win4 = win ? (uint32_t)h->window_base * 0x01010101u : 0u;
    uint64_t start = h->start, accept = h->accept;
    void* args[] = {&d_bytes, &d_begin, &d_end, &n_strings, &tab, &n_classes, &cls, &win4, &start, &accept,
                    &d_final_mask, &d_accept};
    const unsigned lds = (unsigned)n_classes * 2048u + 256u;
    return launch(win ? "nfa64_window" : "nfa64_bytes", (unsigned)grid64, kNfaBlock, lds, (hipStream_t)stream, args);
}

int trex_nfa_match(trex_nfa* h, const uint8_t* d_bytes, const int64_t* d_offsets, int64_t n_strings,
                   uint64_t* d_final_mask, uint8_t* d_accept, void* stream) {
    if (!d_offsets) return fail(TREX_ERR_INVALID, "null offsets");
    return trex_nfa_match_ranges(h, d_bytes, d_offsets, d_offsets + 1, n_strings, d_final_mask, d_accept, stream);
}

int trex_nfa_match_host(trex_nfa* h, const uint8_t* bytes, const int64_t* offsets, int64_t n_strings,
                        uint64_t* final_mask, uint8_t* accept, float* kernel_ms, float* total_ms) {
    if (!h) return fail(TREX_ERR_INVALID, "null handle");
    if (kernel_ms) *kernel_ms = 0.f;
    if (total_ms) *total_ms = 0.f;
    if (n_strings <= 0) return n_strings < 0 ? fail(TREX_ERR_INVALID, "n_strings < 0") : TREX_OK;
    if (!offsets || !final_mask || !accept) return fail(TREX_ERR_INVALID, "null buffer");
    const int64_t base = offsets[0], total = offsets[n_strings] - offsets[0];
    if (total < 0) return fail(TREX_ERR_INVALID, "offsets decrease");
    if (total > 0 && !bytes) return fail(TREX_ERR_INVALID, "null bytes");
    for (int64_t i = 0; i < n_strings; ++i)
        if (offsets[i + 1] < offsets[i]) return fail(TREX_ERR_INVALID, "offsets decrease at %lld", (long long)i);
    std::lock_guard<std::mutex> lock(h->mu);
    TREX_TRY(h->in_bytes.reserve((size_t)total + 16));
    TREX_TRY(h->in_offsets.reserve(sizeof(int64_t) * (size_t)(n_strings + 1)));
    TREX_TRY(h->out_mask.reserve(8 * (size_t)n_strings));
    TREX_TRY(h->out_accept.reserve((size_t)n_strings));
    std::vector<int64_t> rebased((size_t)n_strings + 1);
    for (int64_t i = 0; i <= n_strings; ++i) rebased[i] = offsets[i] - base;
    TREX_TRY(ensure_host_dispatch(h->stream, h->ev));
    hipStream_t s = h->stream;
    TREX_HIP(hipEventRecord(h->ev[0], s), "hipEventRecord");
    if (total > 0) TREX_HIP(hipMemcpyAsync(h->in_bytes.ptr, bytes + base, (size_t)total, hipMemcpyHostToDevice, s), "H2D bytes");
    TREX_HIP(hipMemcpyAsync(h->in_offsets.ptr, rebased.data(), rebased.size() * 8, hipMemcpyHostToDevice, s), "H2D offsets");
    TREX_HIP(hipEventRecord(h->ev[1], s), "hipEventRecord");
    TREX_TRY(trex_nfa_match(h, h->in_bytes.as<uint8_t>(), h->in_offsets.as<int64_t>(), n_strings,
                            h->out_mask.as<uint64_t>(), h->out_accept.as<uint8_t>(), s));
    TREX_HIP(hipEventRecord(h->ev[2], s), "hipEventRecord");
    TREX_HIP(hipMemcpyAsync(final_mask, h->out_mask.ptr, 8 * (size_t)n_strings, hipMemcpyDeviceToHost, s), "D2H mask");
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

}  // extern "C"

// ---------------------------------------------------------------------------
// Reference-compatible singleton (cuda/nfa_tc_evolution.cu:943-1013)
// ---------------------------------------------------------------------------
namespace {
std::mutex g_nfa_mu;
trex_nfa* g_nfa = nullptr;
int g_nfa_max_b = 0, g_nfa_max_l = 0;
}

extern "C" {

int nfa_tc_engine_init(const float* T_matrices, const float* accept_mask, const float* start_vec, int n, int sigma,
                       int max_B, int max_L) {
    if (!T_matrices || !accept_mask || !start_vec) return fail(TREX_ERR_INVALID, "null argument");
    if (n < 1 || n > 64) return fail(TREX_ERR_INVALID, "n=%d outside 1..64", n);
    if (sigma < 0 || sigma > kMaxSymbols) return fail(TREX_ERR_INVALID, "sigma=%d outside 0..%d", sigma, kMaxSymbols);
    // T[c][dst][src] > 0.5 is an edge src -> dst (the reference's float_to_bitpar, :635-648)
    std::vector<uint64_t> succ((size_t)sigma * 64, 0);
    for (int c = 0; c < sigma; ++c)
        for (int dst = 0; dst < n; ++dst)
            for (int src = 0; src < n; ++src)
                if (T_matrices[(size_t)c * n * n + (size_t)dst * n + src] > 0.5f) succ[(size_t)c * 64 + src] |= 1ull << dst;
    uint64_t start = 0, accept = 0;
    for (int i = 0; i < n; ++i) {
        if (start_vec[i] > 0.0f) start |= 1ull << i;
        if (accept_mask[i] > 0.0f) accept |= 1ull << i;
    }
    // symbol indices >= sigma are skipped, as in simulate_nfa_sequential (:1039)
    std::vector<int16_t> b2s(256);
    for (int b = 0; b < 256; ++b) b2s[b] = (int16_t)(b < sigma ? b : -2);
    std::lock_guard<std::mutex> lock(g_nfa_mu);
    if (g_nfa) { trex_nfa_destroy(g_nfa); g_nfa = nullptr; }
    TREX_TRY(trex_nfa_create(succ.data(), sigma, b2s.data(), start, accept, &g_nfa));
    g_nfa_max_b = max_B;
    g_nfa_max_l = max_L;
    return TREX_OK;
}

int nfa_tc_engine_dispatch(const uint8_t* raw_concat, const int* offsets, int* results, int B, int total_chars,
                           float* kernel_ms, float* total_ms) {
    std::lock_guard<std::mutex> lock(g_nfa_mu);
    if (!g_nfa) return TREX_ERR_NO_DEVICE;   // reference: -1 when not initialised
    if (B < 0 || total_chars < 0) return fail(TREX_ERR_INVALID, "negative size");
    if (B > g_nfa_max_b) return fail(TREX_ERR_CAPACITY, "B=%d exceeds init max_B=%d", B, g_nfa_max_b);
    std::vector<int64_t> offs((size_t)B + 1);
    for (int i = 0; i <= B; ++i) offs[i] = offsets[i];
    for (int i = 0; i < B; ++i)
        if (offs[i + 1] - offs[i] > g_nfa_max_l) return fail(TREX_ERR_CAPACITY, "string %d longer than max_L", i);
    std::vector<uint64_t> mask((size_t)B);
    std::vector<uint8_t> acc((size_t)B);
    if (B > 0) TREX_TRY(trex_nfa_match_host(g_nfa, raw_concat, offs.data(), B, mask.data(), acc.data(), kernel_ms, total_ms));
    for (int i = 0; i < B; ++i) results[i] = acc[i];
    return TREX_OK;
}

void nfa_tc_engine_destroy(void) {
    std::lock_guard<std::mutex> lock(g_nfa_mu);
    if (g_nfa) { trex_nfa_destroy(g_nfa); g_nfa = nullptr; }
}

}  // extern "C"
