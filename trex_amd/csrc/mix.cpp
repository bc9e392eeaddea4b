// Mixed DFA-16 / Boolean NFA-64 batches in one persistent launch: the
// work-queue planner (host) and the launch of mix_persistent.
//
// Plan: every string becomes one WHOLE item, or, above the split length of
// its pattern's kind, a run of chunks (DFA: look-back chunks; NFA: speculative
// chunks plus one fix-up item). Items are grouped per (role, pattern) into
// batches of at most kMixBlock lane items, one workgroup each. Queue order:
//   1. NFA_SPEC batches (no waiting, the heaviest bytes),
//   2. DFA_CHUNK batches, in chunk-position order (a chunk's predecessor is
//      in an earlier or the same batch),
//   3. NFA_FIXUP batches (wait only on 1),
//   4. WHOLE batches, longest first (LPT: a batch lasts as long as its
//      longest string; NFA bytes weigh kNfaWeight DFA bytes).

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "runtime_internal.h"
#include "trex_hip.h"

using namespace trex;

namespace {

constexpr int kMixBlock = 512;      // workgroup of mix_persistent (nibble NFA tables)
constexpr int kMixBlockSlice = 1024;   // workgroup of mix_persistent_slice (6/5-bit slice NFA tables)
constexpr int kNfaClasses = 34;
constexpr unsigned kTabSliceBytes = 32u * (4096u + 8u);   // 6/5-bit slice tables of the 32 window classes
constexpr uint32_t kRoleDfaWhole = 0, kRoleDfaChunk = 1, kRoleNfaWhole = 2, kRoleNfaSpec = 3, kRoleNfaFixup = 4;
constexpr int64_t kNfaWeight = 10;   // NFA-64 bytes cost about ten DFA-16 bytes (LDS lookups per byte)

// Same layout as MixPattern / MixArgs in mix_kernels.hip.
struct MixPattern {
    uint32_t kind;
    uint32_t win4;
    uint32_t start;
    uint32_t dead;
    uint32_t nstates;
    uint32_t reserved;
    uint64_t nstart;
    uint64_t naccept;
    const uint8_t* t1;
    const uint8_t* t2;
    const uint8_t* f2;
    const uint8_t* acc;
    const uint8_t* ntab;
    const uint8_t* ncls;
};

struct MixArgs {
    const uint8_t* bytes;
    const MixPattern* patterns;
    const int32_t* bfirst;
    const int32_t* bcount;
    const int32_t* bpattern;
    const int32_t* brole;
    int32_t n_batches;
    int32_t ck;
    const int64_t* ibeg;
    const int64_t* iend;
    const int32_t* ia;
    const int32_t* ib;
    const int32_t* ic;
    const uint32_t* iflags;
    int32_t n_ck;
    uint32_t* ticket;
    uint32_t* status;
    uint64_t* agg;
    uint32_t* done;
    uint64_t* ckpt;
    uint64_t* spec_exit;
    uint32_t* error;
    uint32_t* converged;
    uint32_t max_spins;
    uint32_t reserved;
    uint64_t max_ticks;
    uint64_t* out_mask;
    uint8_t* out_accept;
};

struct Plan {
    std::vector<int64_t> ibeg, iend;
    std::vector<int32_t> ia, ib, ic;
    std::vector<uint32_t> iflags;
    std::vector<int32_t> bfirst, bcount, bpattern, brole;
    int64_t n_spec = 0, n_slots = 0, n_chunks = 0, n_strings = 0;
    int32_t ck = 64, n_ck = 0;
    int32_t block = kMixBlock;   // lane items per batch = workgroup size of the form planned for

    int32_t add(int64_t b, int64_t e, int32_t a, int32_t out, int32_t c, uint32_t flags) {
        ibeg.push_back(b);
        iend.push_back(e);
        ia.push_back(a);
        ib.push_back(out);
        ic.push_back(c);
        iflags.push_back(flags);
        return (int32_t)(ibeg.size() - 1);
    }
    size_t items() const { return ibeg.size(); }
};

struct BatchRef {
    int32_t first, count, pattern, role;
    int64_t key;   // ordering key within its queue section
};

}  // namespace

struct trex_mix {
    int32_t n_patterns = 0;
    std::vector<int32_t> kinds;
    std::vector<MixPattern> pats;
    bool has_device = false;
    MixPattern* d_pats = nullptr;
    MixPattern* d_pats_slice = nullptr;   // the same with the NFAs' 6/5-bit slice tables
    bool can_slice = true;                // every NFA pattern has slice tables
    int form = TREX_MIX_FORM_NIBBLE;
    Plan plan;
    // The plan's ten arrays, packed at 256-byte aligned offsets: built in
    // pinned host memory by trex_mix_plan, copied to d_plan by the next run on
    // that run's stream, by a copy kernel (no synchronous call while other work
    // is on the device, and no wait behind long copies on the copy engines).
    uint8_t* h_stage = nullptr;
    size_t h_cap = 0, plan_bytes = 0;
    std::vector<uint8_t*> h_retired;   // outgrown staging, freed with the handle (hipHostFree synchronises)
    size_t off[10] = {};
    size_t need[5] = {};               // bytes of d_plan, d_words, d_agg, d_ckpt, d_exit for the staged plan
    bool dirty = false;
    hipEvent_t staged = nullptr;   // recorded behind the last upload from h_stage
    bool staged_pending = false;
    DevBuf d_plan;
    DevBuf d_words;   // ticket, error, converged, done[n_slots], status[n_items]
    DevBuf d_agg, d_ckpt, d_exit;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    int64_t grid = 0;   // persistent workgroups (0: two per compute unit)
};

namespace {

void drop_graph(trex_mix* m) {
    if (m->exec) hipGraphExecDestroy(m->exec);
    if (m->graph) hipGraphDestroy(m->graph);
    m->exec = nullptr;
    m->graph = nullptr;
}

// Grows a plan buffer in stream order (hipFreeAsync / hipMallocAsync on the
// run's stream), by at least half its capacity: unlike hipFree, nothing here
// waits for the work of other streams.
int grow_async(DevBuf& buf, size_t n, hipStream_t s) {
    if (n <= buf.cap) return TREX_OK;
    const size_t want = std::max(n, buf.cap + buf.cap / 2);
    if (buf.ptr) TREX_HIP(hipFreeAsync(buf.ptr, s), "hipFreeAsync");
    buf.ptr = nullptr;
    buf.cap = 0;
    TREX_HIP(hipMallocAsync(&buf.ptr, want, s), "hipMallocAsync");
    buf.cap = want;
    return TREX_OK;
}

void release_async(DevBuf& buf) {
    if (buf.ptr) (void)hipFreeAsync(buf.ptr, nullptr);
    buf.ptr = nullptr;
    buf.cap = 0;
}

// Packs the plan's arrays into h_stage and sizes the device buffers.
int stage_plan(trex_mix* m) {
    const Plan& Q = m->plan;
    auto bytes = [](const auto& v) { return v.size() * sizeof(v[0]); };
    const size_t sizes[10] = {bytes(Q.ibeg), bytes(Q.iend), bytes(Q.ia), bytes(Q.ib), bytes(Q.ic),
                              bytes(Q.iflags), bytes(Q.bfirst), bytes(Q.bcount), bytes(Q.bpattern), bytes(Q.brole)};
    const void* srcs[10] = {Q.ibeg.data(), Q.iend.data(), Q.ia.data(), Q.ib.data(), Q.ic.data(),
                            Q.iflags.data(), Q.bfirst.data(), Q.bcount.data(), Q.bpattern.data(), Q.brole.data()};
    size_t total = 0;
    for (int i = 0; i < 10; ++i) {
        m->off[i] = total;
        total += (sizes[i] + 255) & ~(size_t)255;
    }
    total = std::max<size_t>(total, 256);
    if (m->staged_pending) {   // the last upload may still read h_stage
        TREX_HIP(hipEventSynchronize(m->staged), "hipEventSynchronize");
        m->staged_pending = false;
    }
    if (total > m->h_cap) {
        const size_t want = std::max(total, m->h_cap + m->h_cap / 2);
        uint8_t* p = nullptr;
        TREX_HIP(hipHostMalloc((void**)&p, want, hipHostMallocDefault), "hipHostMalloc plan staging");
        if (m->h_stage) m->h_retired.push_back(m->h_stage);
        m->h_stage = p;
        m->h_cap = want;
    }
    for (int i = 0; i < 10; ++i)
        if (sizes[i]) std::memcpy(m->h_stage + m->off[i], srcs[i], sizes[i]);
    m->plan_bytes = total;
    m->need[0] = total;
    m->need[1] = 4 * (3 + (size_t)Q.n_slots + Q.items());
    m->need[2] = 8 * std::max<size_t>(Q.items(), 1);
    m->need[3] = 8 * std::max<size_t>((size_t)Q.n_spec * Q.n_ck, 1);
    m->need[4] = 8 * std::max<size_t>((size_t)Q.n_spec, 1);
    m->dirty = true;
    return TREX_OK;
}

// Enqueues the copy of a freshly staged plan on `s`.
int upload_plan(trex_mix* m, hipStream_t s) {
    if (!m->dirty) return TREX_OK;
    if (!m->staged) TREX_HIP(hipEventCreateWithFlags(&m->staged, hipEventDisableTiming), "hipEventCreate");
    DevBuf* bufs[5] = {&m->d_plan, &m->d_words, &m->d_agg, &m->d_ckpt, &m->d_exit};
    for (int i = 0; i < 5; ++i) TREX_TRY(grow_async(*bufs[i], m->need[i], s));
    TREX_TRY(trex_memcpy_kernel(m->d_plan.ptr, m->h_stage, (int64_t)m->plan_bytes, s));
    TREX_HIP(hipEventRecord(m->staged, s), "hipEventRecord");
    m->staged_pending = true;
    m->dirty = false;
    return TREX_OK;
}

}  // namespace

extern "C" {

int trex_mix_create(const int32_t* kinds, const void* const* handles, int32_t n_patterns, trex_mix** out) {
    if (!out || !kinds || n_patterns < 1) return fail(TREX_ERR_INVALID, "need kinds and at least one pattern");
    *out = nullptr;
    for (int32_t i = 0; i < n_patterns; ++i)
        if (kinds[i] != TREX_MIX_DFA && kinds[i] != TREX_MIX_NFA)
            return fail(TREX_ERR_INVALID, "pattern %d: kind %d is neither TREX_MIX_DFA nor TREX_MIX_NFA", i, kinds[i]);
    std::vector<MixPattern> pats(n_patterns);
    if (handles) {
        for (int32_t i = 0; i < n_patterns; ++i) {
            MixPattern& P = pats[i];
            std::memset(&P, 0, sizeof(P));
            if (!handles[i]) return fail(TREX_ERR_INVALID, "pattern %d: null handle", i);
            if (kinds[i] == TREX_MIX_DFA) {
                const trex_dfa* d = static_cast<const trex_dfa*>(handles[i]);
                if (d->n_states > 16 || d->window_base < 0 || !d->d_t2 || !d->d_f2)
                    return fail(TREX_ERR_INVALID,
                                "pattern %d: the mixed queue takes DFAs of <= 16 states whose live bytes lie in one "
                                "aligned 32-byte window", i);
                P.kind = 0;
                P.win4 = (uint32_t)d->window_base * 0x01010101u;
                P.start = (uint32_t)d->start;
                P.dead = (uint32_t)d->dead;
                P.nstates = (uint32_t)d->n_states;
                P.t1 = d->d_t1;
                P.t2 = d->d_t2;
                P.f2 = d->d_f2;
                P.acc = d->d_acc;
            } else {
                const trex_nfa* f = static_cast<const trex_nfa*>(handles[i]);
                if (f->window_base < 0 || !f->d_tab_win)
                    return fail(TREX_ERR_INVALID,
                                "pattern %d: the mixed queue takes NFAs whose live bytes lie in one aligned 32-byte "
                                "window", i);
                P.kind = 1;
                P.win4 = (uint32_t)f->window_base * 0x01010101u;
                P.nstart = f->start;
                P.naccept = f->accept;
                P.ntab = reinterpret_cast<const uint8_t*>(f->d_tab_win);
                P.ncls = f->d_cls_win;
            }
        }
        TREX_TRY(ensure_module());
    }
    std::vector<MixPattern> pats_slice = pats;
    bool can_slice_all = true;
    for (int32_t i = 0; i < n_patterns && handles; ++i) {
        if (kinds[i] != TREX_MIX_NFA) continue;
        const trex_nfa* f = static_cast<const trex_nfa*>(handles[i]);
        pats_slice[i].ntab = f->d_tab6;
        pats_slice[i].ncls = nullptr;
        can_slice_all = can_slice_all && f->d_tab6 != nullptr;
    }
    trex_mix* m = new trex_mix();
    m->n_patterns = n_patterns;
    m->kinds.assign(kinds, kinds + n_patterns);
    m->pats = pats;
    m->has_device = handles != nullptr;
    m->can_slice = can_slice_all;
    if (m->has_device) {
        int rc = upload((void**)&m->d_pats, pats.data(), pats.size() * sizeof(MixPattern));
        if (!rc && can_slice_all) rc = upload((void**)&m->d_pats_slice, pats_slice.data(), pats_slice.size() * sizeof(MixPattern));
        if (rc) { hipFree(m->d_pats); delete m; return rc; }
    }
    *out = m;
    return TREX_OK;
}

int trex_mix_destroy(trex_mix* m) {
    if (!m) return TREX_OK;
    drop_graph(m);
    if (m->staged_pending) hipEventSynchronize(m->staged);
    if (m->staged) hipEventDestroy(m->staged);
    if (m->h_stage) hipHostFree(m->h_stage);
    for (uint8_t* p : m->h_retired) hipHostFree(p);
    hipFree(m->d_pats);
    hipFree(m->d_pats_slice);
    for (DevBuf* b : {&m->d_plan, &m->d_words, &m->d_agg, &m->d_ckpt, &m->d_exit})
        release_async(*b);
    delete m;
    return TREX_OK;
}

int trex_mix_plan(trex_mix* m, const int64_t* begin, const int64_t* end, const int32_t* pattern_of, int64_t n_strings,
                  int64_t dfa_split, int64_t dfa_chunk, int64_t nfa_split, int64_t nfa_chunk, int32_t ck) {
    if (!m) return fail(TREX_ERR_INVALID, "null handle");
    if (n_strings < 0 || n_strings > 0x7fffffff) return fail(TREX_ERR_INVALID, "n_strings out of range");
    if (n_strings > 0 && (!begin || !end || !pattern_of)) return fail(TREX_ERR_INVALID, "null array");
    if (dfa_chunk < 16 || dfa_chunk % 16 || dfa_split < 1 || nfa_split < 1)
        return fail(TREX_ERR_INVALID, "dfa_chunk must be a multiple of 16 (>= 16) and the split lengths positive");
    if (ck < 16 || ck % 16 || nfa_chunk < ck || nfa_chunk % ck)
        return fail(TREX_ERR_INVALID, "ck must be a multiple of 16 and nfa_chunk a multiple of ck");
    for (int64_t i = 0; i < n_strings; ++i) {
        if (end[i] < begin[i] || begin[i] < 0) return fail(TREX_ERR_INVALID, "string %lld: bad range", (long long)i);
        if (pattern_of[i] < 0 || pattern_of[i] >= m->n_patterns)
            return fail(TREX_ERR_INVALID, "string %lld: pattern %d out of range", (long long)i, pattern_of[i]);
    }
    drop_graph(m);
    Plan P;
    P.block = m->form == TREX_MIX_FORM_SLICE ? kMixBlockSlice : kMixBlock;
    P.ck = ck;
    P.n_ck = (int32_t)(nfa_chunk / ck);
    P.n_strings = n_strings;
    const int32_t np = m->n_patterns;

    // strings per pattern, split into whole and long
    std::vector<std::vector<int32_t>> whole(np), longs(np);
    for (int64_t i = 0; i < n_strings; ++i) {
        const int32_t p = pattern_of[i];
        const int64_t len = end[i] - begin[i];
        const int64_t split = m->kinds[p] == TREX_MIX_DFA ? dfa_split : nfa_split;
        (len > split ? longs : whole)[p].push_back((int32_t)i);
    }
    std::vector<BatchRef> spec_b, chunk_b, fix_b, whole_b;
    auto cut = [&](int32_t first, int32_t n, int32_t pat, int32_t role, std::vector<BatchRef>& out,
                   const std::vector<int64_t>& keys) {
        for (int32_t o = 0; o < n; o += P.block)
            out.push_back({first + o, std::min(P.block, n - o), pat, role, keys.empty() ? 0 : keys[o]});
    };

    // 1. NFA speculative chunks: each long string's chunks contiguous
    std::vector<std::vector<int32_t>> fix_first(np), fix_count(np), fix_slot(np), fix_sid(np);
    for (int32_t p = 0; p < np; ++p) {
        if (m->kinds[p] != TREX_MIX_NFA || longs[p].empty()) continue;
        std::vector<int32_t> order = longs[p];
        std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
            return end[a] - begin[a] > end[b] - begin[b];
        });
        const int32_t first = (int32_t)P.items();
        for (int32_t sid : order) {
            const int32_t slot = (int32_t)P.n_slots++;
            const int32_t c0 = (int32_t)P.items();
            for (int64_t b = begin[sid]; b < end[sid]; b += nfa_chunk) {
                const int32_t ord = (int32_t)P.n_spec++;
                P.add(b, std::min(end[sid], b + nfa_chunk), slot, -1, ord, b == begin[sid] ? 1u : 0u);
            }
            fix_first[p].push_back(c0);
            fix_count[p].push_back((int32_t)P.items() - c0);
            fix_slot[p].push_back(slot);
            fix_sid[p].push_back(sid);
        }
        P.n_chunks += (int32_t)P.items() - first;
        cut(first, (int32_t)P.items() - first, p, kRoleNfaSpec, spec_b, {});
    }
    // 2. DFA look-back chunks, in chunk-position order
    for (int32_t p = 0; p < np; ++p) {
        if (m->kinds[p] != TREX_MIX_DFA || longs[p].empty()) continue;
        struct C { int32_t sid; int64_t k; };
        std::vector<C> cs;
        for (int32_t sid : longs[p]) {
            const int64_t n = (end[sid] - begin[sid] + dfa_chunk - 1) / dfa_chunk;
            for (int64_t k = 0; k < n; ++k) cs.push_back({sid, k});
        }
        std::stable_sort(cs.begin(), cs.end(), [](const C& a, const C& b) { return a.k < b.k; });
        const int32_t first = (int32_t)P.items();
        std::vector<int64_t> keys;
        // the predecessor of chunk k is the item of chunk k - 1, placed earlier
        std::vector<int32_t> latest(n_strings, -1);
        for (const C& c : cs) {
            const int64_t b = begin[c.sid] + c.k * dfa_chunk;
            const int64_t e = std::min(end[c.sid], b + dfa_chunk);
            const bool last = e == end[c.sid];
            const int32_t it = P.add(b, e, c.k == 0 ? -1 : latest[c.sid], last ? c.sid : -1, 0, c.k == 0 ? 1u : 0u);
            latest[c.sid] = it;
            keys.push_back(c.k);
        }
        P.n_chunks += (int32_t)P.items() - first;
        cut(first, (int32_t)P.items() - first, p, kRoleDfaChunk, chunk_b, keys);
    }
    // 3. NFA fix-ups, one item per long string
    for (int32_t p = 0; p < np; ++p) {
        if (fix_first[p].empty()) continue;
        const int32_t first = (int32_t)P.items();
        for (size_t j = 0; j < fix_first[p].size(); ++j)
            P.add(fix_first[p][j], (int64_t)fix_first[p][j] + fix_count[p][j], fix_slot[p][j], fix_sid[p][j], 0, 0u);
        cut(first, (int32_t)P.items() - first, p, kRoleNfaFixup, fix_b, {});
    }
    // 4. whole strings, longest first
    for (int32_t p = 0; p < np; ++p) {
        if (whole[p].empty()) continue;
        std::vector<int32_t> order = whole[p];
        std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
            return end[a] - begin[a] > end[b] - begin[b];
        });
        const int32_t first = (int32_t)P.items();
        const int64_t w = m->kinds[p] == TREX_MIX_NFA ? kNfaWeight : 1;
        std::vector<int64_t> keys;
        for (int32_t sid : order) {
            P.add(begin[sid], end[sid], -1, sid, 0, 0u);
            keys.push_back(-(end[sid] - begin[sid]) * w);
        }
        cut(first, (int32_t)P.items() - first, p,
            m->kinds[p] == TREX_MIX_DFA ? (int32_t)kRoleDfaWhole : (int32_t)kRoleNfaWhole, whole_b, keys);
    }
    if (P.items() > 0x7fffffff) return fail(TREX_ERR_CAPACITY, "too many work items");

    // queue order
    std::stable_sort(chunk_b.begin(), chunk_b.end(), [](const BatchRef& a, const BatchRef& b) { return a.key < b.key; });
    std::stable_sort(whole_b.begin(), whole_b.end(), [](const BatchRef& a, const BatchRef& b) { return a.key < b.key; });
    for (const auto* sec : {&spec_b, &chunk_b, &fix_b, &whole_b})
        for (const BatchRef& b : *sec) {
            P.bfirst.push_back(b.first);
            P.bcount.push_back(b.count);
            P.bpattern.push_back(b.pattern);
            P.brole.push_back(b.role);
        }
    m->plan = std::move(P);

    if (m->has_device) TREX_TRY(stage_plan(m));
    return TREX_OK;
}

int trex_mix_info(const trex_mix* m, int64_t* n_items, int64_t* n_batches, int64_t* n_chunks, int64_t* n_fixups) {
    if (!m) return fail(TREX_ERR_INVALID, "null handle");
    if (n_items) *n_items = (int64_t)m->plan.items();
    if (n_batches) *n_batches = (int64_t)m->plan.bfirst.size();
    if (n_chunks) *n_chunks = m->plan.n_chunks;
    if (n_fixups) *n_fixups = m->plan.n_slots;
    return TREX_OK;
}

int trex_mix_layout(const trex_mix* m, int64_t* ibeg, int64_t* iend, int32_t* ia, int32_t* ib, uint32_t* iflags,
                    int32_t* bfirst, int32_t* bcount, int32_t* bpattern, int32_t* brole) {
    if (!m) return fail(TREX_ERR_INVALID, "null handle");
    const Plan& P = m->plan;
    auto cp = [](auto* dst, const auto& v) { if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0])); };
    cp(ibeg, P.ibeg);
    cp(iend, P.iend);
    cp(ia, P.ia);
    cp(ib, P.ib);
    cp(iflags, P.iflags);
    cp(bfirst, P.bfirst);
    cp(bcount, P.bcount);
    cp(bpattern, P.bpattern);
    cp(brole, P.brole);
    return TREX_OK;
}

int trex_mix_run(trex_mix* m, const uint8_t* d_bytes, uint64_t* d_out_mask, uint8_t* d_out_accept, uint32_t max_spins,
                 void* stream) {
    if (!m) return fail(TREX_ERR_INVALID, "null handle");
    if (!m->has_device) return fail(TREX_ERR_INVALID, "mix created without pattern handles (plan only)");
    const Plan& Q = m->plan;
    if (Q.bfirst.empty()) return TREX_OK;
    if (!d_bytes || !d_out_mask || !d_out_accept) return fail(TREX_ERR_INVALID, "null buffer");
    if (Q.block != (m->form == TREX_MIX_FORM_SLICE ? kMixBlockSlice : kMixBlock))
        return fail(TREX_ERR_INVALID, "the plan's batches were cut for the other form: plan again after trex_mix_set_form");
    hipStream_t s = (hipStream_t)stream;
    TREX_TRY(upload_plan(m, s));
    uint32_t* words = m->d_words.as<uint32_t>();
    int64_t n_words = 3 + Q.n_slots + (int64_t)Q.items();
    uint32_t zero = 0;
    {
        const unsigned grid = (unsigned)std::min<int64_t>((n_words + kBlock - 1) / kBlock, 4096);
        void* args[] = {&words, &n_words, &zero};
        TREX_TRY(launch("fill_u32", std::max(grid, 1u), kBlock, 0, s, args));
    }
    MixArgs a;
    a.bytes = d_bytes;
    const bool sl = m->form == TREX_MIX_FORM_SLICE;
    a.patterns = sl ? m->d_pats_slice : m->d_pats;
    uint8_t* pb = m->d_plan.as<uint8_t>();
    a.ibeg = reinterpret_cast<const int64_t*>(pb + m->off[0]);
    a.iend = reinterpret_cast<const int64_t*>(pb + m->off[1]);
    a.ia = reinterpret_cast<const int32_t*>(pb + m->off[2]);
    a.ib = reinterpret_cast<const int32_t*>(pb + m->off[3]);
    a.ic = reinterpret_cast<const int32_t*>(pb + m->off[4]);
    a.iflags = reinterpret_cast<const uint32_t*>(pb + m->off[5]);
    a.bfirst = reinterpret_cast<const int32_t*>(pb + m->off[6]);
    a.bcount = reinterpret_cast<const int32_t*>(pb + m->off[7]);
    a.bpattern = reinterpret_cast<const int32_t*>(pb + m->off[8]);
    a.brole = reinterpret_cast<const int32_t*>(pb + m->off[9]);
    a.n_batches = (int32_t)Q.bfirst.size();
    a.ck = Q.ck;
    a.n_ck = Q.n_ck;
    a.ticket = words;
    a.error = words + 1;
    a.converged = words + 2;
    a.done = words + 3;
    a.status = words + 3 + Q.n_slots;
    a.agg = m->d_agg.as<uint64_t>();
    a.ckpt = m->d_ckpt.as<uint64_t>();
    a.spec_exit = m->d_exit.as<uint64_t>();
    // waits end after max_spins polls (0: no count limit) or after
    // the wall-clock bound (trex_set_wait_bound_us, 100 ms by default), whichever comes first
    a.max_spins = max_spins;
    a.reserved = 0;
    a.max_ticks = device_wait_ticks();
    a.out_mask = d_out_mask;
    a.out_accept = d_out_accept;
    // persistent grid: by default two 512-thread workgroups per CU (70 KB of LDS each), or one
    // 1024-thread workgroup (128 KB) with the slice tables
    const int64_t want = m->grid > 0 ? m->grid : (sl ? 1 : 2) * (int64_t)device_cu_count();
    const unsigned grid = (unsigned)std::min<int64_t>(want, (int64_t)Q.bfirst.size());
    void* args[] = {&a};
    if (sl) return launch("mix_persistent_slice", grid, kMixBlockSlice, kTabSliceBytes, s, args);
    return launch("mix_persistent", grid, kMixBlock, kNfaClasses * 2048 + 256, s, args);
}

int trex_mix_set_form(trex_mix* m, int form) {
    if (!m) return fail(TREX_ERR_INVALID, "null handle");
    if (form != TREX_MIX_FORM_NIBBLE && form != TREX_MIX_FORM_SLICE)
        return fail(TREX_ERR_INVALID, "form %d: use TREX_MIX_FORM_NIBBLE or TREX_MIX_FORM_SLICE", form);
    if (form == TREX_MIX_FORM_SLICE && m->has_device && !m->can_slice)
        return fail(TREX_ERR_INVALID, "an NFA pattern has no slice tables");
    m->form = form;
    drop_graph(m);
    return TREX_OK;
}

int trex_mix_set_grid(trex_mix* m, int64_t workgroups) {
    if (!m || workgroups < 0) return fail(TREX_ERR_INVALID, "bad arguments");
    m->grid = workgroups;
    drop_graph(m);
    return TREX_OK;
}

int trex_mix_error(trex_mix* m, uint32_t* error, void* stream) {
    if (!m || !error) return fail(TREX_ERR_INVALID, "null argument");
    *error = 0;
    if (!m->has_device || m->plan.bfirst.empty()) return TREX_OK;
    TREX_HIP(hipMemcpyAsync(error, m->d_words.as<uint32_t>() + 1, 4, hipMemcpyDeviceToHost, (hipStream_t)stream),
             "D2H error");
    TREX_HIP(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
    return TREX_OK;
}

int trex_mix_status(trex_mix* m, void* dst, void* stream) {
    if (!m || !dst) return fail(TREX_ERR_INVALID, "null argument");
    if (!m->has_device || m->plan.bfirst.empty()) return TREX_OK;   // nothing ran: dst untouched
    return trex_memcpy_kernel(dst, m->d_words.ptr, 16, stream);
}

int trex_mix_stats(trex_mix* m, uint32_t* converged, void* stream) {
    if (!m || !converged) return fail(TREX_ERR_INVALID, "null argument");
    *converged = 0;
    if (!m->has_device || m->plan.bfirst.empty()) return TREX_OK;
    TREX_HIP(hipMemcpyAsync(converged, m->d_words.as<uint32_t>() + 2, 4, hipMemcpyDeviceToHost, (hipStream_t)stream),
             "D2H stats");
    TREX_HIP(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
    return TREX_OK;
}

int trex_mix_capture(trex_mix* m, const uint8_t* d_bytes, uint64_t* d_out_mask, uint8_t* d_out_accept,
                     uint32_t max_spins, void* stream) {
    if (!m) return fail(TREX_ERR_INVALID, "null handle");
    if (!stream) return fail(TREX_ERR_INVALID, "graph capture needs a non-default stream");
    drop_graph(m);
    hipStream_t s = (hipStream_t)stream;
    if (m->has_device) TREX_TRY(upload_plan(m, s));   // before the capture: replays launch only the kernels
    TREX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    const int rc = trex_mix_run(m, d_bytes, d_out_mask, d_out_accept, max_spins, stream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) { if (g) hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    m->graph = g;
    TREX_HIP(hipGraphInstantiate(&m->exec, g, nullptr, nullptr, 0), "hipGraphInstantiate");
    return TREX_OK;
}

int trex_mix_replay(trex_mix* m, void* stream) {
    if (!m || !m->exec) return fail(TREX_ERR_INVALID, "no captured graph");
    TREX_HIP(hipGraphLaunch(m->exec, (hipStream_t)stream), "hipGraphLaunch");
    return TREX_OK;
}

}  // extern "C"
