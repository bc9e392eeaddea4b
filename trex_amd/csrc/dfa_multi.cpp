// Multi-pattern DFA-16 handle: several resident DFA-16 tables matched in one
// pass over the bytes (kernels in dfa_multi_kernels.hip).
//
// The handle copies each member's pair tables (device to device) into one
// pattern-major array and derives the single-symbol tables from the member's
// byte table, so it does not depend on the member handles after creation.
// A dispatch launches ceil(n_patterns / group) kernels, the patterns split
// as evenly as possible among them; each reads the batch once and advances
// its patterns. The kernels are bound by LDS bank conflicts of the table
// reads. The paired form reads two patterns' pair functions with one 16-byte
// read (default group 4: 33 KB of tables, four workgroups per CU); the
// compact form also packs patterns of at most 4 states eight to a word and
// takes all patterns in one launch (the default when two or more fit).

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdio>
#include <utility>
#include <vector>

#include "runtime_internal.h"
#include "trex_hip.h"

using namespace trex;

namespace {

constexpr int kMultiMax = 8;          // patterns per launch (kernels exist for 1..8)
constexpr int kMultiGroup = 4;        // default patterns per launch
constexpr int kMultiPatterns = 256;   // patterns per handle

// Kernel argument block; the layout of MultiParams in dfa_multi_kernels.hip.
struct MultiParams {
    uint32_t win4[kMultiMax];
    uint32_t start[kMultiMax];
    uint32_t dead[kMultiMax];
    uint32_t accept[kMultiMax];
};

// The compact form's argument block (CompactParams in dfa_compact_kernels.hip):
// the same per launch pattern, plus the pattern's output row.
struct CompactParams {
    uint32_t win4[kMultiMax];
    uint32_t start[kMultiMax];
    uint32_t dead[kMultiMax];
    uint32_t accept[kMultiMax];
    uint32_t slot[kMultiMax];
};

// One launch of the compact form: NF full and NB byte patterns, its tables
// at byte offsets into the handle's compact buffers.
struct CompactLaunch {
    int nf = 0, nb = 0;
    size_t words_off = 0, syms_off = 0;
    CompactParams prm{};
};

}  // namespace

struct trex_dfa_multi {
    int n = 0;
    int form = TREX_DFA_MULTI_K2F_PAIRED;
    int group = kMultiGroup;
    uint8_t* d_f2 = nullptr;   // [n][1024] uint64 pair functions
    uint8_t* d_f1 = nullptr;   // [n][32] uint64 single-symbol functions
    uint8_t* d_t2 = nullptr;   // [n][1024][16] pair walk
    uint8_t* d_t1 = nullptr;   // [n][32][16] single-symbol walk
    std::vector<uint32_t> win4, start, dead, accept;
    std::vector<uint64_t> h_f2;   // [n][1024] pair functions (host copy, for the paired tables)
    uint8_t* d_fp = nullptr;      // paired tables of the current grouping, launch after launch
    std::vector<size_t> fp_off;   // byte offset of each launch's paired tables in d_fp
    int fp_group = -1;            // the group d_fp was built for
    std::vector<uint64_t> h_f1;   // [n][32] single-symbol functions (host copy)
    std::vector<int> n_states;
    uint8_t* d_fc = nullptr;      // compact pair tables of the current grouping
    uint8_t* d_fcs = nullptr;     // and the launches' single-symbol tables in launch order
    std::vector<CompactLaunch> fc;
    int fc_group = -1;            // the group d_fc was built for
};

namespace {

// Launch l of a dispatch covers patterns [first, first + size): the patterns
// split evenly over ceil(n / group) launches (sizes differ by at most one).
std::vector<std::pair<int, int>> launches_of(int n, int group) {
    std::vector<std::pair<int, int>> out;
    const int count = (n + group - 1) / group;
    for (int l = 0, p0 = 0; l < count; ++l) {
        const int P = (n - p0) / (count - l);
        out.push_back({p0, P});
        p0 += P;
    }
    return out;
}

// Paired tables for the current grouping: per launch, (P + 1) / 2 tables of
// 1024 x 16 bytes, entry x = (f2[a][x], f2[b][x]) for the launch's patterns
// a = first + 2j, b = a + 1 (the identity when b is past the launch).
int build_paired(trex_dfa_multi* m) {
    constexpr uint64_t kIdentity = 0xfedcba9876543210ull;   // nibble s = s
    std::vector<uint64_t> host;
    m->fp_off.clear();
    for (const auto& [p0, P] : launches_of(m->n, m->group)) {
        m->fp_off.push_back(host.size() * 8);
        for (int j = 0; j < (P + 1) / 2; ++j) {
            const int a = p0 + 2 * j, b = a + 1;
            for (int x = 0; x < 1024; ++x) {
                host.push_back(m->h_f2[(size_t)a * 1024 + x]);
                host.push_back(b < p0 + P ? m->h_f2[(size_t)b * 1024 + x] : kIdentity);
            }
        }
    }
    hipFree(m->d_fp);
    m->d_fp = nullptr;
    m->fp_group = -1;
    TREX_TRY(upload((void**)&m->d_fp, host.data(), host.size() * 8));
    m->fp_group = m->group;
    return TREX_OK;
}

// A pattern of at most 4 states as one byte per pair entry: 2 bits per state
// (states 0..3; a table of fewer states maps its padding states to themselves).
bool narrow(const trex_dfa_multi* m, int p) { return m->n_states[p] <= 4; }

uint64_t byte_function(uint64_t f) {
    uint64_t b = 0;
    for (int s = 0; s < 4; ++s) b |= ((f >> (4 * s)) & 3u) << (2 * s);
    return b;
}

// Compact tables for the current grouping: per launch, its full patterns
// first (one 64-bit word each), then its patterns of at most 4 states (eight
// bytes to a word); the words of entry x stored pair by pair (NW / 2 tables of
// 1024 x 16 bytes, then 1024 x 8 bytes for an odd last word), and the
// single-symbol functions in the same pattern order.
int build_compact(trex_dfa_multi* m) {
    std::vector<uint64_t> words, syms;
    std::vector<CompactLaunch> fc;
    for (const auto& [p0, P] : launches_of(m->n, m->group)) {
        std::vector<int> order;
        for (int k = p0; k < p0 + P; ++k)
            if (!narrow(m, k)) order.push_back(k);
        const int nf = (int)order.size();
        for (int k = p0; k < p0 + P; ++k)
            if (narrow(m, k)) order.push_back(k);
        CompactLaunch L;
        L.nf = nf;
        L.nb = P - nf;
        L.words_off = words.size() * 8;
        L.syms_off = syms.size() * 8;
        const int nw = nf + (L.nb + 7) / 8;
        std::vector<uint64_t> entry((size_t)1024 * nw, 0);   // [x][word]
        for (int x = 0; x < 1024; ++x) {
            for (int k = 0; k < nf; ++k) entry[(size_t)x * nw + k] = m->h_f2[(size_t)order[k] * 1024 + x];
            for (int i = 0; i < L.nb; ++i)
                entry[(size_t)x * nw + nf + i / 8] |=
                    byte_function(m->h_f2[(size_t)order[nf + i] * 1024 + x]) << (8 * (i % 8));
        }
        for (int j = 0; j < nw / 2; ++j)
            for (int x = 0; x < 1024; ++x) {
                words.push_back(entry[(size_t)x * nw + 2 * j]);
                words.push_back(entry[(size_t)x * nw + 2 * j + 1]);
            }
        if (nw & 1)
            for (int x = 0; x < 1024; ++x) words.push_back(entry[(size_t)x * nw + nw - 1]);
        for (int k = 0; k < P; ++k) {
            const int q = order[k];
            for (int c = 0; c < 32; ++c) syms.push_back(m->h_f1[(size_t)q * 32 + c]);
            L.prm.win4[k] = m->win4[q];
            L.prm.start[k] = m->start[q];
            L.prm.dead[k] = m->dead[q];
            L.prm.accept[k] = m->accept[q];
            L.prm.slot[k] = (uint32_t)q;
        }
        fc.push_back(L);
    }
    hipFree(m->d_fc);
    hipFree(m->d_fcs);
    m->d_fc = m->d_fcs = nullptr;
    m->fc_group = -1;
    TREX_TRY(upload((void**)&m->d_fc, words.data(), words.size() * 8));
    TREX_TRY(upload((void**)&m->d_fcs, syms.data(), syms.size() * 8));
    m->fc = std::move(fc);
    m->fc_group = m->group;
    return TREX_OK;
}

// The default form: compact tables, all patterns of a launch at once, when at
// least two patterns fit one byte (8 seeded 32-symbol patterns, 5 of them of
// 3 states: 0.41 ms per 1 GiB against 0.69 ms paired in launches of four);
// otherwise paired tables in launches of four.
int auto_form(const trex_dfa_multi* m) {
    int small = 0;
    for (int p = 0; p < m->n; ++p) small += narrow(m, p) ? 1 : 0;
    return small >= 2 ? TREX_DFA_MULTI_K2F_COMPACT : TREX_DFA_MULTI_K2F_PAIRED;
}

// (Re)builds the tables the handle's form needs for its grouping.
int ensure_tables(trex_dfa_multi* m) {
    if (m->form == TREX_DFA_MULTI_K2F_PAIRED && m->fp_group != m->group) TREX_TRY(build_paired(m));
    if (m->form == TREX_DFA_MULTI_K2F_COMPACT && m->fc_group != m->group) TREX_TRY(build_compact(m));
    return TREX_OK;
}

}  // namespace

extern "C" {

int trex_dfa_multi_destroy(trex_dfa_multi* m) {
    if (!m) return TREX_OK;
    hipFree(m->d_f2);
    hipFree(m->d_f1);
    hipFree(m->d_t2);
    hipFree(m->d_t1);
    hipFree(m->d_fp);
    hipFree(m->d_fc);
    hipFree(m->d_fcs);
    delete m;
    return TREX_OK;
}

int trex_dfa_multi_create(trex_dfa* const* patterns, int n_patterns, trex_dfa_multi** out) {
    if (!patterns || !out) return fail(TREX_ERR_INVALID, "null argument");
    *out = nullptr;
    if (n_patterns < 1 || n_patterns > kMultiPatterns)
        return fail(TREX_ERR_INVALID, "n_patterns=%d outside 1..%d", n_patterns, kMultiPatterns);
    for (int i = 0; i < n_patterns; ++i) {
        const trex_dfa* h = patterns[i];
        if (!h) return fail(TREX_ERR_INVALID, "pattern %d: null handle", i);
        if (h->n_states > 16 || h->window_base < 0 || h->dead < 0 || !h->d_f2 || !h->d_t2 || h->stride != 16)
            return fail(TREX_ERR_INVALID,
                        "pattern %d: not a DFA-16 whose live bytes lie in one 32-byte window "
                        "(n_states=%d, window=%d); match it with its own handle", i, h->n_states, h->window_base);
    }
    TREX_TRY(ensure_module());
    trex_dfa_multi* m = new trex_dfa_multi();
    m->n = n_patterns;
    const size_t n = (size_t)n_patterns;
    int rc = TREX_OK;
    hipError_t e = hipMalloc((void**)&m->d_f2, n * 1024 * 8);
    if (e == hipSuccess) e = hipMalloc((void**)&m->d_t2, n * 1024 * 16);
    if (e != hipSuccess) rc = hip_fail(e, "hipMalloc(multi-pattern tables)");
    std::vector<uint64_t> f1(n * 32, 0);
    m->h_f2.resize(n * 1024);
    std::vector<uint8_t> t1w(n * 32 * 16, 0), t1(256 * 16), acc(16);
    for (int i = 0; i < n_patterns && !rc; ++i) {
        const trex_dfa* h = patterns[i];
        e = hipMemcpy(m->d_f2 + (size_t)i * 1024 * 8, h->d_f2, 1024 * 8, hipMemcpyDeviceToDevice);
        if (e == hipSuccess) e = hipMemcpy(m->d_t2 + (size_t)i * 1024 * 16, h->d_t2, 1024 * 16, hipMemcpyDeviceToDevice);
        if (e == hipSuccess) e = hipMemcpy(t1.data(), h->d_t1, t1.size(), hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(acc.data(), h->d_acc, acc.size(), hipMemcpyDeviceToHost);
        if (e == hipSuccess)
            e = hipMemcpy(m->h_f2.data() + (size_t)i * 1024, h->d_f2, 1024 * 8, hipMemcpyDeviceToHost);
        if (e != hipSuccess) { rc = hip_fail(e, "hipMemcpy(multi-pattern tables)"); break; }
        const int base = h->window_base;
        for (int c = 0; c < 32; ++c)
            for (int s = 0; s < 16; ++s) {
                const uint8_t nxt = t1[(size_t)(base + c) * 16 + s];   // identity beyond n_states, as in t1
                t1w[((size_t)i * 32 + c) * 16 + s] = nxt;
                f1[(size_t)i * 32 + c] |= (uint64_t)nxt << (4 * s);
            }
        uint32_t mask = 0;
        for (int s = 0; s < 16; ++s) mask |= (acc[s] ? 1u : 0u) << s;
        m->n_states.push_back(h->n_states);
        m->win4.push_back((uint32_t)base * 0x01010101u);
        m->start.push_back((uint32_t)h->start);
        m->dead.push_back((uint32_t)h->dead);
        m->accept.push_back(mask);
    }
    m->h_f1 = f1;
    if (!rc) rc = upload((void**)&m->d_f1, f1.data(), f1.size() * sizeof(uint64_t));
    if (!rc) rc = upload((void**)&m->d_t1, t1w.data(), t1w.size());
    if (!rc) {
        m->form = auto_form(m);
        m->group = m->form == TREX_DFA_MULTI_K2F_COMPACT ? kMultiMax : kMultiGroup;
        rc = ensure_tables(m);
    }
    if (rc) { trex_dfa_multi_destroy(m); return rc; }
    *out = m;
    return TREX_OK;
}

int trex_dfa_multi_set_form(trex_dfa_multi* m, int form) {
    if (!m) return fail(TREX_ERR_INVALID, "null handle");
    if (form == TREX_DFA_AUTO) form = auto_form(m);
    if (form != TREX_DFA_K2_WINDOW_F && form != TREX_DFA_K2_WINDOW && form != TREX_DFA_MULTI_K2F_PAIRED &&
        form != TREX_DFA_MULTI_K2F_COMPACT)
        return fail(TREX_ERR_INVALID, "form %d: use TREX_DFA_K2_WINDOW_F, TREX_DFA_K2_WINDOW, "
                    "TREX_DFA_MULTI_K2F_PAIRED or TREX_DFA_MULTI_K2F_COMPACT", form);
    const int prev = m->form;
    m->form = form;
    const int rc = ensure_tables(m);
    if (rc) m->form = prev;
    return rc;
}

int trex_dfa_multi_set_group(trex_dfa_multi* m, int group) {
    if (!m) return fail(TREX_ERR_INVALID, "null handle");
    if (group < 1 || group > kMultiMax) return fail(TREX_ERR_INVALID, "group=%d outside 1..%d", group, kMultiMax);
    m->group = group;
    return ensure_tables(m);
}

int trex_dfa_multi_info(const trex_dfa_multi* m, int* n_patterns, int* form, int* group) {
    if (!m) return fail(TREX_ERR_INVALID, "null handle");
    if (n_patterns) *n_patterns = m->n;
    if (form) *form = m->form;
    if (group) *group = m->group;
    return TREX_OK;
}

}  // extern "C"

namespace {

int multi_dispatch(trex_dfa_multi* m, const uint8_t* d_bytes, const int64_t* d_begin, const int64_t* d_end,
                   int64_t n_strings, uint8_t* d_final_state, uint8_t* d_accept, void* stream, bool il) {
    if (!m) return fail(TREX_ERR_INVALID, "null handle");
    if (n_strings < 0) return fail(TREX_ERR_INVALID, "n_strings < 0");
    if (n_strings == 0) return TREX_OK;
    if (!d_bytes || !d_begin || !d_end || !d_final_state || !d_accept) return fail(TREX_ERR_INVALID, "null buffer");
    const int64_t grid64 = (n_strings + kBlock - 1) / kBlock;
    if (grid64 > 0x7fffffff) return fail(TREX_ERR_CAPACITY, "too many strings");
    const unsigned grid = (unsigned)grid64;
    if (m->form == TREX_DFA_MULTI_K2F_COMPACT) {
        if (m->fc_group != m->group) return fail(TREX_ERR_INVALID, "compact tables not built");
        for (const CompactLaunch& L : m->fc) {
            const uint8_t* words = m->d_fc + L.words_off;
            const uint8_t* syms = m->d_fcs + L.syms_off;
            CompactParams prm = L.prm;
            void* args[] = {&d_bytes, &d_begin, &d_end, &n_strings, &words, &syms, &prm, &d_final_state, &d_accept};
            char name[48];
            std::snprintf(name, sizeof(name), "dfa16_multi_k2fc_f%d_b%d%s", L.nf, L.nb, il ? "_il" : "");
            TREX_TRY(launch(name, grid, kBlock, 0, (hipStream_t)stream, args));
        }
        return TREX_OK;
    }
    const bool paired = m->form == TREX_DFA_MULTI_K2F_PAIRED;
    const bool f = m->form == TREX_DFA_K2_WINDOW_F || paired;
    const size_t pair_bytes = f ? 1024 * 8 : 1024 * 16, sym_bytes = f ? 32 * 8 : 32 * 16;
    if (paired && m->fp_group != m->group) return fail(TREX_ERR_INVALID, "paired tables not built");
    const auto launches = launches_of(m->n, m->group);
    for (size_t l = 0; l < launches.size(); ++l) {
        const auto [p0, P] = launches[l];
        MultiParams prm{};
        for (int k = 0; k < P; ++k) {
            prm.win4[k] = m->win4[p0 + k];
            prm.start[k] = m->start[p0 + k];
            prm.dead[k] = m->dead[p0 + k];
            prm.accept[k] = m->accept[p0 + k];
        }
        const uint8_t* pairs = paired ? m->d_fp + m->fp_off[l] : (f ? m->d_f2 : m->d_t2) + (size_t)p0 * pair_bytes;
        const uint8_t* syms = (f ? m->d_f1 : m->d_t1) + (size_t)p0 * sym_bytes;
        uint8_t* out_s = d_final_state + (size_t)p0 * (size_t)n_strings;
        uint8_t* out_a = d_accept + (size_t)p0 * (size_t)n_strings;
        void* args[] = {&d_bytes, &d_begin, &d_end, &n_strings, &pairs, &syms, &prm, &out_s, &out_a};
        char name[48];
        std::snprintf(name, sizeof(name), "dfa16_multi_%s_p%d%s", paired ? "k2fp" : f ? "k2f" : "k2w", P,
                      il ? "_il" : "");
        TREX_TRY(launch(name, grid, kBlock, 0, (hipStream_t)stream, args));
    }
    return TREX_OK;
}

}  // namespace

extern "C" {

int trex_dfa_multi_match_ranges(trex_dfa_multi* m, const uint8_t* d_bytes, const int64_t* d_begin,
                                const int64_t* d_end, int64_t n_strings, uint8_t* d_final_state, uint8_t* d_accept,
                                void* stream) {
    return multi_dispatch(m, d_bytes, d_begin, d_end, n_strings, d_final_state, d_accept, stream, false);
}

int trex_dfa_multi_match_interleaved(trex_dfa_multi* m, const uint8_t* d_bytes, const int64_t* d_begin,
                                     const int64_t* d_end, int64_t n_strings, uint8_t* d_final_state,
                                     uint8_t* d_accept, void* stream) {
    return multi_dispatch(m, d_bytes, d_begin, d_end, n_strings, d_final_state, d_accept, stream, true);
}

}  // extern "C"
