// Probabilistic NFA-64: handle creation (rounded transition matrices, column
// mass factors), dispatch of the MFMA forward kernels, host round trip.

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "runtime_internal.h"
#include "trex_hip.h"

using namespace trex;

namespace {

constexpr int kStates = 64;
constexpr int kMaxSymbols = 32;   // four waves x eight symbols of register-resident fragments
constexpr int kSeqPerGroup = 64;

uint32_t f32_bits(float f) { uint32_t x; std::memcpy(&x, &f, 4); return x; }
float bits_f32(uint32_t x) { float f; std::memcpy(&f, &x, 4); return f; }

// IEEE binary16, round to nearest even (finite input).
uint16_t to_f16(float f) {
    const uint32_t x = f32_bits(f);
    const uint16_t sign = (uint16_t)((x >> 16) & 0x8000u);
    const uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x477ff000u) return sign | 0x7c00u;             // rounds to infinity
    if (ax < 0x38800000u) {                                   // below 2^-14: subnormal result
        const float v = bits_f32(ax) * 16777216.0f;           // exact scaling by 2^24
        return sign | (uint16_t)std::nearbyint(v);
    }
    const uint32_t mant = ax & 0x7fffffu;
    uint32_t h = ((uint32_t)((int)(ax >> 23) - 127 + 15) << 10) | (mant >> 13);
    const uint32_t rem = mant & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;
    return sign | (uint16_t)h;
}

float from_f16(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    if (e == 0) {
        const float v = (float)m * (1.0f / 16777216.0f);
        return (h & 0x8000u) ? -v : v;
    }
    if (e == 31) return bits_f32(sign | 0x7f800000u | (m << 13));
    return bits_f32(sign | ((e - 15 + 127) << 23) | (m << 13));
}

// bfloat16, round to nearest even (finite input).
uint16_t to_bf16(float f) {
    const uint32_t x = f32_bits(f);
    return (uint16_t)((x + 0x7fffu + ((x >> 16) & 1u)) >> 16);
}

float from_bf16(uint16_t b) { return bits_f32((uint32_t)b << 16); }

// Round every entry, then give each source column j of T[c] the factor
// colsum(T[c])_j / colsum(round(T[c]))_j (1 for a column with a negative entry
// or a zero rounded sum), so round(T[c]) applied to (factor * s) moves the
// same total mass as T[c] applied to s. The factor is kept in FP64; the
// kernels apply it as an FP32 pair (high part + rest): a factor rounded to
// FP32 alone is off by up to 2^-24 in the same direction on every step with
// that symbol, a mass drift that grows linearly with the sequence length.
void round_tables(const float* T, int n_sym, bool bf16, std::vector<uint16_t>& q, std::vector<double>& scale) {
    const size_t n = (size_t)n_sym * kStates * kStates;
    q.resize(n);
    scale.assign((size_t)n_sym * kStates, 1.0);
    for (size_t i = 0; i < n; ++i) q[i] = bf16 ? to_bf16(T[i]) : to_f16(T[i]);
    for (int c = 0; c < n_sym; ++c)
        for (int j = 0; j < kStates; ++j) {
            double sum = 0, sum_q = 0;
            bool nonneg = true;
            for (int i = 0; i < kStates; ++i) {
                const size_t at = ((size_t)c * kStates + i) * kStates + j;
                sum += T[at];
                sum_q += bf16 ? from_bf16(q[at]) : from_f16(q[at]);
                nonneg = nonneg && T[at] >= 0.f;
            }
            if (nonneg && sum_q > 0) scale[(size_t)c * kStates + j] = sum / sum_q;
        }
}

bool uses_f16(int mode) { return mode == TREX_PNFA_F16X2 || mode == TREX_PNFA_F16; }

// Per-symbol column vectors as the kernels read them, four blocks of
// [n_sym][64] FP32: the column factors f (high part, then the FP32 rest) and
// the column sums of the rounded matrix, colsum(round(T[c])) (high part, then
// rest). The kernels check each product's mass against sum_k colsum_k x_k
// computed on the vector ALUs (round to nearest): the matrix cores' FP32
// accumulation loses a little mass on every step, always in the same
// direction, which over 10^5 steps moves LL by 1e-2. Mass only means
// something for a nonnegative model: otherwise the column sums are left zero,
// which turns the check off.
std::vector<float> operator_tables(const std::vector<double>& scale, const std::vector<uint16_t>& q, bool bf16,
                                   bool nonneg) {
    const size_t m = scale.size();
    std::vector<float> out(4 * m);
    for (size_t i = 0; i < m; ++i) {
        out[i] = (float)scale[i];
        out[m + i] = (float)(scale[i] - (double)out[i]);
        const size_t c = i / kStates, j = i % kStates;
        double cs = 0;
        for (int r = 0; r < kStates; ++r) {
            const uint16_t v = q[(c * kStates + r) * kStates + j];
            cs += bf16 ? from_bf16(v) : from_f16(v);
        }
        if (!nonneg) cs = 0;
        out[2 * m + i] = (float)cs;
        out[3 * m + i] = (float)(cs - (double)out[2 * m + i]);
    }
    return out;
}

// The rounded matrices as the chunk kernel's A operands: for symbol c, row
// block rb, k block kb and lane l = 16 g + i, eight values
// q[c][16 rb + i][16 (2 kb + e / 4) + 4 g + e % 4], e = 0..7 (16 bytes per
// lane, 8 KB per symbol): the k order in which the kernel's product result is
// already its next B operand.
std::vector<uint16_t> chunk_fragments(const std::vector<uint16_t>& q, int n_sym) {
    std::vector<uint16_t> out((size_t)n_sym * 4 * 2 * 64 * 8);
    size_t at = 0;
    for (int c = 0; c < n_sym; ++c)
        for (int rb = 0; rb < 4; ++rb)
            for (int kb = 0; kb < 2; ++kb)
                for (int l = 0; l < 64; ++l)
                    for (int e = 0; e < 8; ++e) {
                        const int row = 16 * rb + (l & 15);
                        const int k = 16 * (2 * kb + e / 4) + 4 * (l >> 4) + e % 4;
                        out[at++] = q[((size_t)c * kStates + row) * kStates + k];
                    }
    return out;
}

// Pinned staging of one chunked call's tables; reused once the copy that
// read it has run (its event has fired). kStageSlots of kStageBytes each are
// made with the handle (a table of 64 Ki chunks fits); a larger table or more
// calls in flight add or grow slots, doubling.
constexpr int kStageSlots = 8;
constexpr size_t kStageBytes = 1u << 20;
struct StageSlot {
    uint8_t* host = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
    bool used = false;
};

const char* mode_name(int mode) {
    switch (mode) {
        case TREX_PNFA_F16X2: return "f16x2";
        case TREX_PNFA_BF16X2: return "bf16x2";
        case TREX_PNFA_F16: return "f16";
        case TREX_PNFA_BF16: return "bf16";
        default: return nullptr;
    }
}

}  // namespace

struct trex_pnfa {
    int n_symbols = 0;
    int mode = TREX_PNFA_F16X2;
    int waves = 8;           // waves per workgroup (4: eight symbols per wave, 8: four)
    int64_t grid = 0;        // most workgroups per kernel launch (0: one per group of 64 sequences)
    bool f16_ok = true;      // every |T| fits binary16
    int e0 = 0;              // s0 = s0n * 2^e0 with max |s0n| in [0.5, 1)
    std::vector<uint16_t> q16, qb16;
    std::vector<double> scale16, scaleb16;
    uint16_t* d_q16 = nullptr;
    uint16_t* d_qb16 = nullptr;
    float* d_scale16 = nullptr;
    float* d_scaleb16 = nullptr;
    float* d_t32 = nullptr;
    float* d_s0 = nullptr;
    float* d_acc = nullptr;
    uint16_t* d_cf16 = nullptr;    // chunk-kernel A operands (FP16, BF16)
    uint16_t* d_cfb16 = nullptr;
    DevBuf in_sym, out_ll, out_lm;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    std::mutex mu;
    std::vector<StageSlot> stages;   // guarded by mu
};

extern "C" {

int trex_pnfa_create(const float* T, int n_symbols, const float* start, const float* accept, trex_pnfa** out) {
    if (!out || !start || !accept || (n_symbols > 0 && !T)) return fail(TREX_ERR_INVALID, "null argument");
    *out = nullptr;
    if (n_symbols < 1 || n_symbols > kMaxSymbols)
        return fail(TREX_ERR_INVALID, "n_symbols=%d outside 1..%d", n_symbols, kMaxSymbols);
    const size_t n = (size_t)n_symbols * kStates * kStates;
    float max_abs = 0.f;
    for (size_t i = 0; i < n; ++i) {
        if (!std::isfinite(T[i])) return fail(TREX_ERR_INVALID, "T[%zu] is not finite", i);
        max_abs = std::fmax(max_abs, std::fabs(T[i]));
    }
    if (max_abs > 3.0e38f) return fail(TREX_ERR_INVALID, "T entries too large");
    float s_max = 0.f;
    for (int i = 0; i < kStates; ++i) {
        if (!std::isfinite(start[i]) || !std::isfinite(accept[i]))
            return fail(TREX_ERR_INVALID, "start/accept[%d] is not finite", i);
        s_max = std::fmax(s_max, std::fabs(start[i]));
    }
    TREX_TRY(ensure_module());

    trex_pnfa* h = new trex_pnfa();
    h->n_symbols = n_symbols;
    h->f16_ok = max_abs <= 65504.f;
    // default: the split state (FP16 when T fits it, else BF16), within 1e-3
    // relative LL of FP64 also over 10^5-step sequences; the single-rounding
    // modes are 20 % faster and reach ~1e-3 only beyond 10^5 steps
    h->mode = h->f16_ok ? TREX_PNFA_F16X2 : TREX_PNFA_BF16X2;
    round_tables(T, n_symbols, false, h->q16, h->scale16);
    round_tables(T, n_symbols, true, h->qb16, h->scaleb16);
    std::vector<float> s0n(start, start + kStates);
    if (s_max > 0.f) {
        std::frexp(s_max, &h->e0);
        for (auto& v : s0n) v = std::ldexp(v, -h->e0);
    }
    int rc = upload((void**)&h->d_q16, h->q16.data(), n * 2);
    if (!rc) rc = upload((void**)&h->d_qb16, h->qb16.data(), n * 2);
    if (!rc) {
        bool nonneg = true;
        for (size_t i = 0; i < n && nonneg; ++i) nonneg = T[i] >= 0.f;
        for (int i = 0; i < kStates && nonneg; ++i) nonneg = start[i] >= 0.f;
        const std::vector<float> p16 = operator_tables(h->scale16, h->q16, false, nonneg);
        const std::vector<float> pb16 = operator_tables(h->scaleb16, h->qb16, true, nonneg);
        rc = upload((void**)&h->d_scale16, p16.data(), p16.size() * 4);
        if (!rc) rc = upload((void**)&h->d_scaleb16, pb16.data(), pb16.size() * 4);
    }
    if (!rc) rc = upload((void**)&h->d_t32, T, n * 4);
    if (!rc) rc = upload((void**)&h->d_s0, s0n.data(), kStates * 4);
    if (!rc) rc = upload((void**)&h->d_acc, accept, kStates * 4);
    if (!rc) {
        const std::vector<uint16_t> cf = chunk_fragments(h->q16, n_symbols), cfb = chunk_fragments(h->qb16, n_symbols);
        rc = upload((void**)&h->d_cf16, cf.data(), cf.size() * 2);
        if (!rc) rc = upload((void**)&h->d_cfb16, cfb.data(), cfb.size() * 2);
    }
    // pinned staging of the chunked path, made here: hipHostMalloc and
    // hipHostFree can wait for the whole device, so a server that keeps the
    // engine resident must not pay them while other work runs
    for (int i = 0; i < kStageSlots && !rc; ++i) {
        h->stages.emplace_back();
        StageSlot& st = h->stages.back();
        hipError_t e = hipEventCreateWithFlags(&st.done, hipEventDisableTiming);
        if (e == hipSuccess) e = hipHostMalloc((void**)&st.host, kStageBytes, hipHostMallocDefault);
        if (e != hipSuccess) rc = hip_fail(e, "pinned staging");
        else st.cap = kStageBytes;
    }
    if (rc) { trex_pnfa_destroy(h); return rc; }
    *out = h;
    return TREX_OK;
}

int trex_pnfa_destroy(trex_pnfa* h) {
    if (!h) return TREX_OK;
    if (h->stream) hipStreamSynchronize(h->stream);
    hipFree(h->d_q16);
    hipFree(h->d_qb16);
    hipFree(h->d_scale16);
    hipFree(h->d_scaleb16);
    hipFree(h->d_t32);
    hipFree(h->d_s0);
    hipFree(h->d_acc);
    hipFree(h->d_cf16);
    hipFree(h->d_cfb16);
    for (StageSlot& st : h->stages) {
        if (st.used) hipEventSynchronize(st.done);
        if (st.done) hipEventDestroy(st.done);
        if (st.host) hipHostFree(st.host);
    }
    h->in_sym.release();
    h->out_ll.release();
    h->out_lm.release();
    for (auto& e : h->ev) if (e) hipEventDestroy(e);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
    return TREX_OK;
}

int trex_pnfa_info(const trex_pnfa* h, int* mode, int* n_symbols, int* f16_ok, int* waves) {
    if (!h) return fail(TREX_ERR_INVALID, "null handle");
    if (mode) *mode = h->mode;
    if (waves) *waves = h->waves;
    if (n_symbols) *n_symbols = h->n_symbols;
    if (f16_ok) *f16_ok = h->f16_ok ? 1 : 0;
    return TREX_OK;
}

int trex_pnfa_set_mode(trex_pnfa* h, int mode) {
    if (!h) return fail(TREX_ERR_INVALID, "null handle");
    if (!mode_name(mode)) return fail(TREX_ERR_INVALID, "unknown mode %d", mode);
    if (uses_f16(mode) && !h->f16_ok) return fail(TREX_ERR_INVALID, "T has entries beyond the FP16 range");
    h->mode = mode;
    return TREX_OK;
}

int trex_pnfa_set_waves(trex_pnfa* h, int waves) {
    if (!h) return fail(TREX_ERR_INVALID, "null handle");
    if (waves != 4 && waves != 8) return fail(TREX_ERR_INVALID, "waves must be 4 or 8, got %d", waves);
    h->waves = waves;
    return TREX_OK;
}

int trex_pnfa_set_grid(trex_pnfa* h, int64_t workgroups) {
    if (!h) return fail(TREX_ERR_INVALID, "null handle");
    if (workgroups < 0 || workgroups > 0x7fffffff) return fail(TREX_ERR_INVALID, "workgroups out of range");
    h->grid = workgroups;
    return TREX_OK;
}

int trex_pnfa_rounded(const trex_pnfa* h, int mode, float* t_rounded, double* col_scale) {
    if (!h) return fail(TREX_ERR_INVALID, "null handle");
    if (!mode_name(mode)) return fail(TREX_ERR_INVALID, "unknown mode %d", mode);
    const bool f16 = uses_f16(mode);
    const auto& q = f16 ? h->q16 : h->qb16;
    const auto& sc = f16 ? h->scale16 : h->scaleb16;
    if (t_rounded)
        for (size_t i = 0; i < q.size(); ++i) t_rounded[i] = f16 ? from_f16(q[i]) : from_bf16(q[i]);
    if (col_scale)   // the factor the kernels apply: high part + rest
        for (size_t i = 0; i < sc.size(); ++i) {
            const float hi = (float)sc[i];
            col_scale[i] = (double)hi + (double)(float)(sc[i] - (double)hi);
        }
    return TREX_OK;
}

namespace {

int pnfa_launch(trex_pnfa* h, const uint8_t* d_symbols, const int64_t* d_begin, const int64_t* d_end, int64_t n_seq,
                int L, double* d_ll, double* d_logmass, void* stream) {
    const int64_t groups = (n_seq + kSeqPerGroup - 1) / kSeqPerGroup;
    if (groups > 0x7fffffff) return fail(TREX_ERR_INVALID, "too many sequences");
    const bool f16 = uses_f16(h->mode);
    const uint16_t* tlp = f16 ? h->d_q16 : h->d_qb16;
    const float* scale = f16 ? h->d_scale16 : h->d_scaleb16;
    int n_sym = h->n_symbols, e0 = h->e0;
    char name[48];
    std::snprintf(name, sizeof(name), "pnfa64_%s_w%d", mode_name(h->mode), h->waves);
    // A grid cap splits the groups over consecutive launches on the stream.
    // Callers put long sequences first, so the first launch's longest group
    // sets the pace as it would for one workgroup taking every cap-th group.
    const int64_t per = h->grid > 0 ? h->grid : groups;
    for (int64_t g0 = 0; g0 < groups; g0 += per) {
        const int64_t first = g0 * kSeqPerGroup;
        int64_t n = std::min<int64_t>(n_seq - first, per * kSeqPerGroup);
        const uint8_t* sym = d_begin ? d_symbols : d_symbols + (d_symbols ? first * (int64_t)L : 0);
        const int64_t* b = d_begin ? d_begin + first : nullptr;
        const int64_t* e = d_end ? d_end + first : nullptr;
        double* ll = d_ll + first;
        double* lm = d_logmass + first;
        void* args[] = {&tlp, &scale, &h->d_t32, &n_sym, &h->d_s0, &e0, &h->d_acc, &sym, &b, &e,
                        &n, &L, &ll, &lm};
        const unsigned grid = (unsigned)((n + kSeqPerGroup - 1) / kSeqPerGroup);
        TREX_TRY(launch(name, grid, 64u * (unsigned)h->waves, 0, (hipStream_t)stream, args));
    }
    return TREX_OK;
}

}  // namespace

int trex_pnfa_loglik(trex_pnfa* h, const uint8_t* d_symbols, int64_t n_seq, int64_t length, double* d_ll,
                     double* d_logmass, void* stream) {
    if (!h) return fail(TREX_ERR_INVALID, "null handle");
    if (n_seq < 0 || length < 0) return fail(TREX_ERR_INVALID, "negative size");
    if (length > 0x7fffffff) return fail(TREX_ERR_INVALID, "length %lld beyond 2^31-1", (long long)length);
    if (n_seq == 0) return TREX_OK;
    if (!d_ll || !d_logmass || (length > 0 && !d_symbols)) return fail(TREX_ERR_INVALID, "null buffer");
    return pnfa_launch(h, d_symbols, nullptr, nullptr, n_seq, (int)length, d_ll, d_logmass, stream);
}

int trex_pnfa_loglik_ranges(trex_pnfa* h, const uint8_t* d_symbols, const int64_t* d_begin, const int64_t* d_end,
                            int64_t n_seq, double* d_ll, double* d_logmass, void* stream) {
    if (!h) return fail(TREX_ERR_INVALID, "null handle");
    if (n_seq < 0) return fail(TREX_ERR_INVALID, "negative size");
    if (n_seq == 0) return TREX_OK;
    if (!d_ll || !d_logmass || !d_begin || !d_end || !d_symbols) return fail(TREX_ERR_INVALID, "null buffer");
    return pnfa_launch(h, d_symbols, d_begin, d_end, n_seq, 0, d_ll, d_logmass, stream);
}

int trex_pnfa_loglik_chunked(trex_pnfa* h, const uint8_t* d_symbols, const int64_t* begin, const int64_t* end,
                             int64_t n_seq, int64_t chunk_len, double* d_ll, double* d_logmass, void* stream) {
    if (!h) return fail(TREX_ERR_INVALID, "null handle");
    if (n_seq < 0 || chunk_len < 0) return fail(TREX_ERR_INVALID, "negative size");
    if (n_seq == 0) return TREX_OK;
    if (!begin || !end || !d_ll || !d_logmass || !d_symbols) return fail(TREX_ERR_INVALID, "null buffer");
    int64_t steps = 0;
    for (int64_t j = 0; j < n_seq; ++j) {
        if (begin[j] < 0 || end[j] < begin[j]) return fail(TREX_ERR_INVALID, "sequence %lld: bad range", (long long)j);
        steps += std::max<int64_t>(end[j] - begin[j] - 1, 0);
    }
    // auto: about four chunks per compute unit, at least 256 symbols each (the
    // chained result's error grows with the number of chunks, ~1e-7 to 5e-7
    // relative LL per chunk measured, so chunks stay long)
    if (chunk_len == 0)
        chunk_len = std::max<int64_t>(256, (steps + 4 * device_cu_count() - 1) / (4 * device_cu_count()));
    if (chunk_len > 0x7fffffff) return fail(TREX_ERR_INVALID, "chunk_len beyond 2^31 - 1");
    int64_t n_chunks = 0;
    for (int64_t j = 0; j < n_seq; ++j)
        n_chunks += (std::max<int64_t>(end[j] - begin[j] - 1, 0) + chunk_len - 1) / chunk_len;
    if (n_chunks > (int64_t)1 << 26) return fail(TREX_ERR_CAPACITY, "%lld chunks: raise chunk_len", (long long)n_chunks);

    // table (staged, then copied by a kernel): cbeg[n_chunks] clen[n_chunks]
    // first[n] count[n] sbegin[n] send[n]; workspace: products, exponents, table
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t o_cbeg = 0, o_clen = al(8 * (size_t)n_chunks), o_first = o_clen + al(4 * (size_t)n_chunks);
    const size_t o_count = o_first + al(8 * (size_t)n_seq), o_sb = o_count + al(4 * (size_t)n_seq);
    const size_t o_se = o_sb + al(8 * (size_t)n_seq), table_bytes = o_se + al(8 * (size_t)n_seq);
    const size_t p_bytes = al((size_t)n_chunks * kStates * kStates * 4), e_bytes = al((size_t)n_chunks * kStates * 4);
    const size_t work_bytes = p_bytes + e_bytes + table_bytes;
    hipStream_t s = (hipStream_t)stream;

    std::lock_guard<std::mutex> lock(h->mu);
    StageSlot* slot = nullptr;
    for (StageSlot& st : h->stages)
        if (!st.used || hipEventQuery(st.done) == hipSuccess) { slot = &st; break; }
    if (!slot) {
        h->stages.emplace_back();
        slot = &h->stages.back();
        TREX_HIP(hipEventCreateWithFlags(&slot->done, hipEventDisableTiming), "hipEventCreate");
    }
    slot->used = false;
    if (slot->cap < table_bytes) {
        const size_t want = std::max(table_bytes, 2 * slot->cap);
        if (slot->host) hipHostFree(slot->host);
        slot->host = nullptr;
        slot->cap = 0;
        TREX_HIP(hipHostMalloc((void**)&slot->host, want, hipHostMallocDefault), "hipHostMalloc");
        slot->cap = want;
    }
    int64_t* cbeg = reinterpret_cast<int64_t*>(slot->host + o_cbeg);
    int32_t* clen = reinterpret_cast<int32_t*>(slot->host + o_clen);
    int64_t* first = reinterpret_cast<int64_t*>(slot->host + o_first);
    int32_t* count = reinterpret_cast<int32_t*>(slot->host + o_count);
    int64_t k = 0;
    for (int64_t j = 0; j < n_seq; ++j) {
        const int64_t st = std::max<int64_t>(end[j] - begin[j] - 1, 0);
        first[j] = k;
        count[j] = (int32_t)((st + chunk_len - 1) / chunk_len);
        for (int64_t b = 0; b < st; b += chunk_len, ++k) {
            cbeg[k] = begin[j] + b;
            clen[k] = (int32_t)std::min<int64_t>(chunk_len, st - b);
        }
    }
    std::memcpy(slot->host + o_sb, begin, 8 * (size_t)n_seq);
    std::memcpy(slot->host + o_se, end, 8 * (size_t)n_seq);

    uint8_t* work = nullptr;
    TREX_HIP(hipMallocAsync((void**)&work, work_bytes, s), "hipMallocAsync");
    uint8_t* tab = work + p_bytes + e_bytes;
    int rc = trex_memcpy_kernel(tab, slot->host, (int64_t)table_bytes, s);
    if (!rc) {
        const hipError_t e = hipEventRecord(slot->done, s);
        if (e != hipSuccess) rc = hip_fail(e, "hipEventRecord");
        slot->used = true;
    }
    const bool f16 = uses_f16(h->mode);
    const int n_sym = h->n_symbols, e0 = h->e0;
    float* d_p = reinterpret_cast<float*>(work);
    int32_t* d_e = reinterpret_cast<int32_t*>(work + p_bytes);
    const int64_t* d_cbeg = reinterpret_cast<const int64_t*>(tab + o_cbeg);
    const int32_t* d_clen = reinterpret_cast<const int32_t*>(tab + o_clen);
    const int64_t* d_first = reinterpret_cast<const int64_t*>(tab + o_first);
    const int32_t* d_count = reinterpret_cast<const int32_t*>(tab + o_count);
    const int64_t* d_sb = reinterpret_cast<const int64_t*>(tab + o_sb);
    const int64_t* d_se = reinterpret_cast<const int64_t*>(tab + o_se);
    if (!rc && n_chunks > 0) {
        const uint16_t* af = f16 ? h->d_cf16 : h->d_cfb16;
        const float* scale = f16 ? h->d_scale16 : h->d_scaleb16;
        void* args[] = {&af, &scale, (void*)&n_sym, &d_symbols, &d_cbeg, &d_clen, &n_chunks, &d_p, &d_e};
        char name[48];
        std::snprintf(name, sizeof(name), "pnfa64_chunk_%s", mode_name(h->mode));
        rc = launch(name, (unsigned)((n_chunks + 3) / 4), 256, 0, s, args);
    }
    if (!rc) {
        void* args[] = {&d_p, &d_e, &d_first, &d_count, &h->d_s0, (void*)&e0, &h->d_t32, &h->d_acc, (void*)&n_sym,
                        &d_symbols, &d_sb, &d_se, &n_seq, &d_ll, &d_logmass};
        rc = launch("pnfa64_combine", (unsigned)((n_seq + 3) / 4), 256, 0, s, args);
    }
    const hipError_t fe = hipFreeAsync(work, s);
    if (!rc && fe != hipSuccess) rc = hip_fail(fe, "hipFreeAsync");
    return rc;
}

int trex_pnfa_loglik_host(trex_pnfa* h, const uint8_t* symbols, int64_t n_seq, int64_t length, double* ll,
                          double* logmass, float* kernel_ms, float* total_ms) {
    if (!h) return fail(TREX_ERR_INVALID, "null handle");
    if (kernel_ms) *kernel_ms = 0.f;
    if (total_ms) *total_ms = 0.f;
    if (n_seq < 0 || length < 0) return fail(TREX_ERR_INVALID, "negative size");
    if (n_seq == 0) return TREX_OK;
    if (!ll || !logmass || (length > 0 && !symbols)) return fail(TREX_ERR_INVALID, "null buffer");
    const size_t n_sym_bytes = (size_t)n_seq * (size_t)length;

    std::lock_guard<std::mutex> lock(h->mu);
    TREX_TRY(h->in_sym.reserve(n_sym_bytes + 16));
    TREX_TRY(h->out_ll.reserve(sizeof(double) * (size_t)n_seq));
    TREX_TRY(h->out_lm.reserve(sizeof(double) * (size_t)n_seq));
    TREX_TRY(ensure_host_dispatch(h->stream, h->ev));
    hipStream_t s = h->stream;
    TREX_HIP(hipEventRecord(h->ev[0], s), "hipEventRecord");
    if (n_sym_bytes) TREX_HIP(hipMemcpyAsync(h->in_sym.ptr, symbols, n_sym_bytes, hipMemcpyHostToDevice, s), "H2D symbols");
    TREX_HIP(hipEventRecord(h->ev[1], s), "hipEventRecord");
    TREX_TRY(trex_pnfa_loglik(h, h->in_sym.as<uint8_t>(), n_seq, length, h->out_ll.as<double>(),
                              h->out_lm.as<double>(), s));
    TREX_HIP(hipEventRecord(h->ev[2], s), "hipEventRecord");
    TREX_HIP(hipMemcpyAsync(ll, h->out_ll.ptr, sizeof(double) * (size_t)n_seq, hipMemcpyDeviceToHost, s), "D2H ll");
    TREX_HIP(hipMemcpyAsync(logmass, h->out_lm.ptr, sizeof(double) * (size_t)n_seq, hipMemcpyDeviceToHost, s),
             "D2H logmass");
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
