// Probabilistic NFA-64: handle creation (rounded transition matrices, column
// mass factors), dispatch of the MFMA forward kernels, host round trip.

#include <hip/hip_runtime_api.h>

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
// same total mass as T[c] applied to s.
void round_tables(const float* T, int n_sym, bool bf16, std::vector<uint16_t>& q, std::vector<float>& scale) {
    const size_t n = (size_t)n_sym * kStates * kStates;
    q.resize(n);
    scale.assign((size_t)n_sym * kStates, 1.f);
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
            if (nonneg && sum_q > 0) scale[(size_t)c * kStates + j] = (float)(sum / sum_q);
        }
}

bool uses_f16(int mode) { return mode == TREX_PNFA_F16X2 || mode == TREX_PNFA_F16; }

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
    std::vector<float> scale16, scaleb16;
    uint16_t* d_q16 = nullptr;
    uint16_t* d_qb16 = nullptr;
    float* d_scale16 = nullptr;
    float* d_scaleb16 = nullptr;
    float* d_t32 = nullptr;
    float* d_s0 = nullptr;
    float* d_acc = nullptr;
    DevBuf in_sym, out_ll, out_lm;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    std::mutex mu;
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
    if (!rc) rc = upload((void**)&h->d_scale16, h->scale16.data(), h->scale16.size() * 4);
    if (!rc) rc = upload((void**)&h->d_scaleb16, h->scaleb16.data(), h->scaleb16.size() * 4);
    if (!rc) rc = upload((void**)&h->d_t32, T, n * 4);
    if (!rc) rc = upload((void**)&h->d_s0, s0n.data(), kStates * 4);
    if (!rc) rc = upload((void**)&h->d_acc, accept, kStates * 4);
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

int trex_pnfa_rounded(const trex_pnfa* h, int mode, float* t_rounded, float* col_scale) {
    if (!h) return fail(TREX_ERR_INVALID, "null handle");
    if (!mode_name(mode)) return fail(TREX_ERR_INVALID, "unknown mode %d", mode);
    const bool f16 = uses_f16(mode);
    const auto& q = f16 ? h->q16 : h->qb16;
    const auto& sc = f16 ? h->scale16 : h->scaleb16;
    if (t_rounded)
        for (size_t i = 0; i < q.size(); ++i) t_rounded[i] = f16 ? from_f16(q[i]) : from_bf16(q[i]);
    if (col_scale) std::memcpy(col_scale, sc.data(), sc.size() * 4);
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
