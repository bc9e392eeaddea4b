// Chunk plans for long strings and the chunked DFA-16 launch with decoupled
// look-back (kernels in scan_kernels.hip).

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

#include "runtime_internal.h"
#include "trex_hip.h"

using namespace trex;

namespace {
// Bytes left after each tile of a plan's chunk-interleaved image (default).
constexpr int64_t kImageTileGap = 256;
}  // namespace

struct trex_plan {
    int64_t n_strings = 0;
    int64_t n_chunks = 0;
    int64_t chunk_len = 0;
    DevBuf begin, end, sid, pred, flags;   // per chunk, in processing order
    DevBuf status, agg, jump, control;     // look-back state; control = {ticket, error}
    // chunk-interleaved image (see trex_plan_build_image): tile t = chunks
    // 64t .. 64t+63 in processing order, tile_vecs[t] 16-byte vectors per lane
    std::vector<int64_t> h_begin, h_end;   // chunk ranges in processing order (host copy)
    std::vector<int64_t> h_tile_off;       // byte offset of each tile in the image, n_tiles + 1 entries
    DevBuf tile_off;                       // the same offsets on the device
    int64_t image_gap = kImageTileGap;     // bytes left after each tile (trex_plan_set_image_gap)
    uint32_t hold = 0;                     // test hook: delayed publication (trex_plan_set_hold)
    int chains = kScanChains;              // dependent chains per lane in phase 1 (trex_plan_set_chains)
    std::mutex mu;                         // one launch at a time per plan
};

namespace {

// Image tiles: 64 chunks each, as many 16-byte vectors per lane as the
// tile's longest chunk needs (chunks are sorted longest first, so a tile pads
// little), each followed by image_gap bytes. Every wave streams its own tile
// front to back; with tiles a power of two apart the resident waves would all
// read the same offset modulo that power at once and crowd a few HBM
// channels, which the gap staggers (as the gap between interleaved groups of
// the batch layout does).
int layout_tiles(trex_plan* p) {
    const int64_t n = p->n_chunks, n_tiles = (n + 63) / 64;
    p->h_tile_off.assign((size_t)n_tiles + 1, 0);
    for (int64_t t = 0; t < n_tiles; ++t) {
        int64_t longest = 0;
        for (int64_t k = t * 64; k < std::min(n, t * 64 + 64); ++k)
            longest = std::max(longest, p->h_end[k] - p->h_begin[k]);
        p->h_tile_off[t + 1] = p->h_tile_off[t] + (longest + 15) / 16 * 16 * 64 + p->image_gap;
    }
    hipError_t e = hipMemcpy(p->tile_off.ptr, p->h_tile_off.data(), p->h_tile_off.size() * 8, hipMemcpyHostToDevice);
    return e == hipSuccess ? TREX_OK : hip_fail(e, "plan upload");
}

}  // namespace

extern "C" {

int trex_plan_create(const int64_t* begin, const int64_t* end, int64_t n_strings, int64_t chunk_len,
                     trex_plan** out) {
    if (!out) return fail(TREX_ERR_INVALID, "null out pointer");
    *out = nullptr;
    if (n_strings < 0 || (n_strings > 0 && (!begin || !end))) return fail(TREX_ERR_INVALID, "bad string ranges");
    if (chunk_len < 64 || chunk_len % 16 != 0) return fail(TREX_ERR_INVALID, "chunk_len must be a multiple of 16, >= 64");
    if (n_strings > 0x7fffffff) return fail(TREX_ERR_CAPACITY, "too many strings for one plan");

    // cut strings into chunks of chunk_len (the last one shorter)
    std::vector<int64_t> cb, ce;
    std::vector<int32_t> cs, cp;
    std::vector<uint32_t> cf;
    for (int64_t i = 0; i < n_strings; ++i) {
        const int64_t b = begin[i], e = end[i];
        if (e < b) return fail(TREX_ERR_INVALID, "string %lld has end < begin", (long long)i);
        const int64_t len = e - b;
        const int64_t k = len <= chunk_len ? 1 : (len + chunk_len - 1) / chunk_len;
        for (int64_t j = 0; j < k; ++j) {
            cb.push_back(b + j * chunk_len);
            ce.push_back(std::min(e, b + (j + 1) * chunk_len));
            cs.push_back((int32_t)i);
            cp.push_back(j == 0 ? -1 : (int32_t)(cb.size() - 2));
            cf.push_back((j == 0 ? 1u : 0u) | (j == k - 1 ? 2u : 0u));
        }
    }
    const int64_t n = (int64_t)cb.size();
    if (n > 0x7fffffff) return fail(TREX_ERR_CAPACITY, "too many chunks");
    // processing order: longer chunks first, first chunks ahead of other chunks
    // of equal length; stable, so each chunk keeps its predecessor in front
    std::vector<int64_t> order((size_t)n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
        const int64_t la = ce[a] - cb[a], lb = ce[b] - cb[b];
        if (la != lb) return la > lb;
        return (cf[a] & 1u) > (cf[b] & 1u);
    });
    std::vector<int64_t> pos((size_t)n);
    for (int64_t k = 0; k < n; ++k) pos[order[k]] = k;
    std::vector<int64_t> ob((size_t)n), oe((size_t)n);
    std::vector<int32_t> os((size_t)n), op((size_t)n);
    std::vector<uint32_t> of((size_t)n);
    for (int64_t k = 0; k < n; ++k) {
        const int64_t c = order[k];
        ob[k] = cb[c];
        oe[k] = ce[c];
        os[k] = cs[c];
        op[k] = cp[c] < 0 ? -1 : (int32_t)pos[cp[c]];
        of[k] = cf[c];
        if (op[k] >= k) return fail(TREX_ERR_INVALID, "internal: chunk order puts a predecessor behind");
    }

    trex_plan* p = new trex_plan();
    p->n_strings = n_strings;
    p->n_chunks = n;
    p->chunk_len = chunk_len;
    const size_t m = (size_t)std::max<int64_t>(n, 1);
    int rc = p->begin.reserve(m * 8);
    if (!rc) rc = p->end.reserve(m * 8);
    if (!rc) rc = p->sid.reserve(m * 4);
    if (!rc) rc = p->pred.reserve(m * 4);
    if (!rc) rc = p->flags.reserve(m * 4);
    if (!rc) rc = p->status.reserve(m * 4);
    if (!rc) rc = p->agg.reserve(m * 8);
    if (!rc) rc = p->jump.reserve(m * 4);
    if (!rc) rc = p->control.reserve(16);
    p->h_begin = ob;
    p->h_end = oe;
    if (!rc) rc = p->tile_off.reserve(((n + 63) / 64 + 1) * 8);
    if (!rc) rc = layout_tiles(p);
    if (!rc && n > 0) {
        hipError_t e = hipMemcpy(p->begin.ptr, ob.data(), (size_t)n * 8, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(p->end.ptr, oe.data(), (size_t)n * 8, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(p->sid.ptr, os.data(), (size_t)n * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(p->pred.ptr, op.data(), (size_t)n * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(p->flags.ptr, of.data(), (size_t)n * 4, hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = hip_fail(e, "plan upload");
    }
    if (rc) { trex_plan_destroy(p); return rc; }
    *out = p;
    return TREX_OK;
}

int trex_plan_destroy(trex_plan* p) {
    if (!p) return TREX_OK;
    for (DevBuf* b : {&p->begin, &p->end, &p->sid, &p->pred, &p->flags, &p->status, &p->agg, &p->jump, &p->control,
                      &p->tile_off})
        b->release();
    delete p;
    return TREX_OK;
}

int trex_plan_info(const trex_plan* p, int64_t* n_chunks, int64_t* n_strings) {
    if (!p) return fail(TREX_ERR_INVALID, "null plan");
    if (n_chunks) *n_chunks = p->n_chunks;
    if (n_strings) *n_strings = p->n_strings;
    return TREX_OK;
}

}  // extern "C"

namespace {

// One launch over all chunks of the plan: `d_bytes` is what the plan's ranges
// index, or (image) the plan's chunk-interleaved image.
int match_plan(trex_dfa* h, trex_plan* p, const uint8_t* d_bytes, bool image, uint8_t* d_final_state,
               uint8_t* d_accept, uint32_t max_spins, void* stream) {
    if (!h || !p) return fail(TREX_ERR_INVALID, "null handle or plan");
    if (h->n_states > 16) return fail(TREX_ERR_INVALID, "chunked matching needs a table of at most 16 states");
    if (p->n_chunks == 0) return TREX_OK;
    if (!d_bytes || !d_final_state || !d_accept) return fail(TREX_ERR_INVALID, "null buffer");
    // waits end after max_spins polls (0: no count limit) or after
    // the wall-clock bound (trex_set_wait_bound_us, 100 ms by default), whichever comes first
    uint64_t max_ticks = device_wait_ticks();
    hipStream_t s = (hipStream_t)stream;
    std::lock_guard<std::mutex> lock(p->mu);

    // clear look-back status, ticket and error word with the fill kernel
    {
        uint32_t* st = p->status.as<uint32_t>();
        int64_t n = p->n_chunks;
        uint32_t zero = 0;
        void* a1[] = {&st, &n, &zero};
        TREX_TRY(launch("fill_u32", (unsigned)std::min<int64_t>((n + kBlock - 1) / kBlock, 4096), kBlock, 0, s, a1));
        uint32_t* ctl = p->control.as<uint32_t>();
        int64_t two = 2;
        void* a2[] = {&ctl, &two, &zero};
        TREX_TRY(launch("fill_u32", 1, kBlock, 0, s, a2));
    }
    const bool window = h->window_base >= 0;
    uint32_t win4 = window ? (uint32_t)h->window_base * 0x01010101u : 0u;
    uint32_t start = (uint32_t)h->start;
    uint32_t dead = h->dead >= 0 ? (uint32_t)h->dead : 0u;
    uint32_t n_states = (uint32_t)h->n_states;
    uint8_t* t2 = window ? h->d_t2 : h->d_t1;
    uint8_t* f2 = window ? h->d_f2 : h->d_t1;   // the byte kernel builds its own functions
    int64_t n_chunks = p->n_chunks;
    uint32_t* ticket = p->control.as<uint32_t>();
    uint32_t* error = ticket + 1;
    void* args[] = {&d_bytes, &p->tile_off.ptr, &p->begin.ptr, &p->end.ptr, &p->sid.ptr, &p->pred.ptr,
                    &p->flags.ptr, &n_chunks, &t2, &h->d_t1, &f2, &h->d_acc, &win4, &start, &dead, &n_states,
                    &p->status.ptr, &p->agg.ptr, &p->jump.ptr, &ticket, &error, &max_spins, &max_ticks, &p->hold,
                    &d_final_state, &d_accept};
    // persistent grid: enough waves to fill every CU, never more than the work
    const int64_t waves_needed = (n_chunks + 63) / 64;
    const int64_t blocks = std::min<int64_t>((waves_needed + 3) / 4, (int64_t)device_cu_count() * 8);
    char kernel[48];
    std::snprintf(kernel, sizeof(kernel), "dfa16_chunks_%s%s%s", window ? "window" : "bytes", image ? "_img" : "",
                  p->chains == 2 ? "_c2" : "");
    return launch(kernel, (unsigned)std::max<int64_t>(blocks, 1), kBlock, 0, s, args);
}

}  // namespace

extern "C" {

int trex_dfa_match_plan(trex_dfa* h, trex_plan* p, const uint8_t* d_bytes, uint8_t* d_final_state,
                        uint8_t* d_accept, uint32_t max_spins, void* stream) {
    return match_plan(h, p, d_bytes, false, d_final_state, d_accept, max_spins, stream);
}

int trex_plan_set_chains(trex_plan* p, int chains) {
    if (!p) return fail(TREX_ERR_INVALID, "null plan");
    if (chains != 1 && chains != 2) return fail(TREX_ERR_INVALID, "chains must be 1 or 2");
    std::lock_guard<std::mutex> lock(p->mu);
    p->chains = chains;
    return TREX_OK;
}

int trex_plan_set_image_gap(trex_plan* p, int64_t gap) {
    if (!p) return fail(TREX_ERR_INVALID, "null plan");
    if (gap < 0 || gap > 65536 || gap % 16 != 0)
        return fail(TREX_ERR_INVALID, "image gap %lld: a multiple of 16 in [0, 65536]", (long long)gap);
    std::lock_guard<std::mutex> lock(p->mu);
    // launches already queued read the tile offsets on the device: let them
    // finish before the offsets change under them
    const hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return hip_fail(e, "synchronise before changing the image layout");
    p->image_gap = gap;
    return layout_tiles(p);
}

int trex_plan_set_hold(trex_plan* p, uint32_t hold) {
    if (!p) return fail(TREX_ERR_INVALID, "null plan");
    if (hold > 2000u) return fail(TREX_ERR_INVALID, "hold %u above 2000", hold);   // <= ~3 ms per tile
    std::lock_guard<std::mutex> lock(p->mu);
    p->hold = hold;
    return TREX_OK;
}

int trex_plan_image_size(const trex_plan* p, int64_t* bytes) {
    if (!p || !bytes) return fail(TREX_ERR_INVALID, "null argument");
    *bytes = p->h_tile_off.empty() ? 0 : p->h_tile_off.back();
    return TREX_OK;
}

int trex_plan_build_image(const trex_plan* p, const uint8_t* src, int64_t src_len, uint8_t* dst, int64_t dst_len) {
    if (!p || (p->n_chunks > 0 && (!src || !dst))) return fail(TREX_ERR_INVALID, "null argument");
    const int64_t size = p->h_tile_off.empty() ? 0 : p->h_tile_off.back();
    if (dst_len < size) return fail(TREX_ERR_INVALID, "image buffer holds %lld bytes, the plan needs %lld",
                                    (long long)dst_len, (long long)size);
    const int64_t n = p->n_chunks;
    for (int64_t c = 0; c < n; ++c)
        if (p->h_begin[c] < 0 || p->h_end[c] > src_len)
            return fail(TREX_ERR_INVALID, "chunk %lld lies outside the %lld source bytes", (long long)c,
                        (long long)src_len);
    if (size > 0) std::memset(dst, 0, (size_t)size);
    for (int64_t c = 0; c < n; ++c) {
        uint8_t* out = dst + p->h_tile_off[(size_t)(c >> 6)] + 16 * (c & 63);
        const uint8_t* in = src + p->h_begin[c];
        const int64_t len = p->h_end[c] - p->h_begin[c];
        int64_t i = 0;
        for (; i + 16 <= len; i += 16) std::memcpy(out + i * 64, in + i, 16);
        if (i < len) std::memcpy(out + i * 64, in + i, (size_t)(len - i));
    }
    return TREX_OK;
}

int trex_dfa_match_plan_image(trex_dfa* h, trex_plan* p, const uint8_t* d_image, int64_t image_bytes,
                              uint8_t* d_final_state, uint8_t* d_accept, uint32_t max_spins, void* stream) {
    if (!p) return fail(TREX_ERR_INVALID, "null plan");
    const int64_t need = p->h_tile_off.empty() ? 0 : p->h_tile_off.back();
    if (image_bytes < need)
        return fail(TREX_ERR_INVALID, "image buffer holds %lld bytes, the plan's image is %lld",
                    (long long)image_bytes, (long long)need);
    return match_plan(h, p, d_image, true, d_final_state, d_accept, max_spins, stream);
}

int trex_plan_error(trex_plan* p, uint32_t* error, void* stream) {
    if (!p || !error) return fail(TREX_ERR_INVALID, "null argument");
    uint32_t host = 0;
    hipError_t e = hipMemcpyAsync(&host, p->control.as<uint32_t>() + 1, 4, hipMemcpyDeviceToHost, (hipStream_t)stream);
    if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "read plan error word");
    *error = host;
    return TREX_OK;
}

}  // extern "C"
