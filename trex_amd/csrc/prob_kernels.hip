// Probabilistic NFA-64 forward pass on MFMA (gfx950).
//
// For every sequence c_1..c_L: s_t = T[c_t] s_{t-1}, s_0 given, and
//     LL      = log(a . s_L)          (a: accept weights)
//     logmass = log(1 . s_L)
// T[c] is 64 x 64 (row = destination, column = source). The reference's
// probabilistic kernel evolves the same recurrence in FP16 with an FP16
// accumulator and no rescaling and returns a thresholded bit
// (cuda/nfa_tc_evolution.cu:451-561). This kernel:
//
//  * holds T rounded to FP16 (or BF16) as MFMA A-fragments in registers: a
//    workgroup has four waves and wave w owns symbols 8w .. 8w+7 (64 fragments
//    of 16 x 32 per wave);
//  * keeps the 64 sequences' states in LDS in FP32; before the product each
//    state is multiplied by the symbol's column factors
//    colsum(T[c]) / colsum(round(T[c])), so the rounded matrix moves exactly
//    the mass the FP32 one does, then split into a high and a low half (two
//    MFMAs per tile, about 22 bits of state) or rounded once;
//  * each step, every wave gathers the sequences whose current symbol it owns
//    into 16-column tiles (the n-th set bit of a ballot mask -> column n),
//    runs v_mfma_f32_16x16x32_{f16,bf16} with FP32 accumulation, rescales each
//    new state by an exact power of two (largest entry into [0.5, 1)) and adds
//    the exponent to the sequence's integer log2-scale;
//  * recomputes the last step in FP32 from the unrounded T and reduces
//    a . s_L and 1 . s_L in FP64 together with the scale.
//
// MFMA operand layout (16x16x32): lane l holds row (A) / column (B) l & 15 and
// the eight consecutive k = 8 * (l >> 4) + e of the 32-wide k block (any fixed
// assignment of k to (lane, element) gives the same product as long as A and B
// share it); D lane l holds column l & 15, rows 4 * (l >> 4) + v.

#include <hip/hip_runtime.h>
#include <stdint.h>

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

#define PNFA_BLOCK 256

namespace {

constexpr int kStates = 64;
constexpr int kRow = kStates + 4;   // padded LDS row (floats) against bank conflicts
constexpr int kSeqPerGroup = 64;
constexpr int kSymPerWave = 8;
constexpr int kMaxSymbols = 32;
constexpr int kChunk = 64;          // symbols staged per sequence at a time

template <bool BF16>
struct Lp;

template <>
struct Lp<false> {
    typedef f16x8 vec;
    typedef _Float16 scalar;
    __device__ __forceinline__ static f32x4 mfma(vec a, vec b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};

template <>
struct Lp<true> {
    typedef bf16x8 vec;
    typedef __bf16 scalar;
    __device__ __forceinline__ static f32x4 mfma(vec a, vec b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};

// Position of the n-th set bit of m (n < popcount(m)).
__device__ __forceinline__ int nth_set_bit(uint64_t m, int n) {
    int base = 0;
    uint32_t w = (uint32_t)m;
    const int lo_cnt = __popc(w);
    if (n >= lo_cnt) {
        n -= lo_cnt;
        w = (uint32_t)(m >> 32);
        base = 32;
    }
#pragma unroll
    for (int half = 16; half >= 1; half >>= 1) {
        const uint32_t low = w & ((1u << half) - 1u);
        const int c = __popc(low);
        if (n >= c) {
            n -= c;
            w >>= half;
            base += half;
        } else {
            w = low;
        }
    }
    return base;
}

template <bool BF16, bool SPLIT>
__device__ void pnfa_forward(const uint16_t* __restrict__ tlp, const float* __restrict__ colscale,
                             const float* __restrict__ t32, int n_sym, const float* __restrict__ s0, int e0,
                             const float* __restrict__ accw, const uint8_t* __restrict__ symbols, int64_t n_seq,
                             int L, double* __restrict__ out_ll, double* __restrict__ out_logmass)
{
    typedef typename Lp<BF16>::vec vec;
    typedef typename Lp<BF16>::scalar scalar;
    __shared__ __attribute__((aligned(16))) float state[2][kSeqPerGroup][kRow];
    __shared__ __attribute__((aligned(16))) float dscale[kMaxSymbols][kStates];
    __shared__ uint8_t syms[kSeqPerGroup][kChunk];
    __shared__ int32_t expo[kSeqPerGroup];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int64_t seq0 = (int64_t)blockIdx.x * kSeqPerGroup;
    const int nseq = (int)min<int64_t>(kSeqPerGroup, n_seq - seq0);
    const int koff = 8 * (lane >> 4);

    // A fragments of this wave's symbols: A[q][rb][kb] holds rows rb*16 + (lane & 15),
    // columns kb*32 + koff + e
    vec A[kSymPerWave][4][2];
#pragma unroll
    for (int q = 0; q < kSymPerWave; ++q) {
        const int c = wave * kSymPerWave + q;
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
            for (int kb = 0; kb < 2; ++kb) {
                if (c < n_sym) {
                    const uint4 raw = *reinterpret_cast<const uint4*>(
                        tlp + ((size_t)c * kStates + rb * 16 + (lane & 15)) * kStates + kb * 32 + koff);
                    A[q][rb][kb] = __builtin_bit_cast(vec, raw);
                } else {
                    A[q][rb][kb] = vec{};
                }
            }
    }
    for (int i = tid; i < kSeqPerGroup * kStates; i += PNFA_BLOCK) state[0][i / kStates][i % kStates] = s0[i % kStates];
    for (int i = tid; i < kMaxSymbols * kStates; i += PNFA_BLOCK)
        dscale[i / kStates][i % kStates] = i < n_sym * kStates ? colscale[i] : 1.f;
    if (tid < kSeqPerGroup) expo[tid] = e0;
    __syncthreads();

    int cur = 0;
    for (int t = 0; t < L - 1; ++t) {
        if (t % kChunk == 0) {
            for (int i = tid; i < kSeqPerGroup * kChunk; i += PNFA_BLOCK) {
                const int s = i / kChunk, k = i % kChunk;
                syms[s][k] = (s < nseq && t + k < L) ? symbols[(seq0 + s) * (int64_t)L + t + k] : (uint8_t)255;
            }
            __syncthreads();
        }
        const int cj = lane < nseq ? (int)syms[lane][t % kChunk] : 255;
        // a symbol outside the alphabet empties the state
        if (wave == 0 && lane < nseq && cj >= n_sym) {
            float4* row = reinterpret_cast<float4*>(&state[cur ^ 1][lane][0]);
            for (int k = 0; k < kStates / 4; ++k) row[k] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int q = 0; q < kSymPerWave; ++q) {
            const int c = wave * kSymPerWave + q;
            if (c >= n_sym) break;
            const uint64_t m = __ballot(cj == c);
            if (m == 0) continue;
            const int cnt = __popcll(m);
            for (int g = 0; g * 16 < cnt; ++g) {
                const int n = g * 16 + (lane & 15);
                const bool valid = n < cnt;
                const int seq = valid ? nth_set_bit(m, n) : 0;
                vec bhi[2], blo[2];
#pragma unroll
                for (int kb = 0; kb < 2; ++kb) {
                    const float4* src = reinterpret_cast<const float4*>(&state[cur][seq][kb * 32 + koff]);
                    const float4* dsc = reinterpret_cast<const float4*>(&dscale[c][kb * 32 + koff]);
                    const float4 x0 = src[0], x1 = src[1], d0 = dsc[0], d1 = dsc[1];
                    const float x[8] = {x0.x * d0.x, x0.y * d0.y, x0.z * d0.z, x0.w * d0.w,
                                        x1.x * d1.x, x1.y * d1.y, x1.z * d1.z, x1.w * d1.w};
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const float v = valid ? x[e] : 0.f;
                        const scalar h = (scalar)v;
                        bhi[kb][e] = h;
                        if constexpr (SPLIT) blo[kb][e] = (scalar)(v - (float)h);
                    }
                }
                f32x4 d[4];
#pragma unroll
                for (int rb = 0; rb < 4; ++rb) {
                    d[rb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int kb = 0; kb < 2; ++kb) {
                        d[rb] = Lp<BF16>::mfma(A[q][rb][kb], bhi[kb], d[rb]);
                        if constexpr (SPLIT) d[rb] = Lp<BF16>::mfma(A[q][rb][kb], blo[kb], d[rb]);
                    }
                }
                float mx = 0.f;
#pragma unroll
                for (int rb = 0; rb < 4; ++rb)
                    mx = fmaxf(mx, fmaxf(fmaxf(fabsf(d[rb][0]), fabsf(d[rb][1])), fmaxf(fabsf(d[rb][2]), fabsf(d[rb][3]))));
                mx = fmaxf(mx, __shfl_xor(mx, 16, 64));
                mx = fmaxf(mx, __shfl_xor(mx, 32, 64));
                int e2 = 0;
                if (mx > 0.f) (void)frexpf(mx, &e2);     // mx = f * 2^e2 with f in [0.5, 1)
                if (valid) {
#pragma unroll
                    for (int rb = 0; rb < 4; ++rb) {
                        const float4 out = make_float4(ldexpf(d[rb][0], -e2), ldexpf(d[rb][1], -e2),
                                                       ldexpf(d[rb][2], -e2), ldexpf(d[rb][3], -e2));
                        *reinterpret_cast<float4*>(&state[cur ^ 1][seq][rb * 16 + 4 * (lane >> 4)]) = out;
                    }
                    if ((lane >> 4) == 0) expo[seq] += e2;
                }
            }
        }
        __syncthreads();
        cur ^= 1;
    }

    // last step in FP32 from the unrounded T: four threads per sequence, 16 rows each
    const int s = tid >> 2, part = tid & 3;
    double acc_a = 0.0, acc_m = 0.0;
    if (s < nseq) {
        const float* st = state[cur][s];
        if (L >= 1) {
            const int c = symbols[(seq0 + s) * (int64_t)L + L - 1];
            if (c < n_sym) {
                for (int r = part * 16; r < part * 16 + 16; ++r) {
                    const float* trow = t32 + ((size_t)c * kStates + r) * kStates;
                    float v = 0.f;
                    for (int k = 0; k < kStates; ++k) v = fmaf(trow[k], st[k], v);
                    acc_a += (double)v * (double)accw[r];
                    acc_m += (double)v;
                }
            }
        } else {
            for (int r = part * 16; r < part * 16 + 16; ++r) {
                acc_a += (double)st[r] * (double)accw[r];
                acc_m += (double)st[r];
            }
        }
    }
    acc_a += __shfl_xor(acc_a, 1, 64);
    acc_a += __shfl_xor(acc_a, 2, 64);
    acc_m += __shfl_xor(acc_m, 1, 64);
    acc_m += __shfl_xor(acc_m, 2, 64);
    if (s < nseq && part == 0) {
        const double scale = (double)expo[s] * 0.69314718055994530942;   // ln 2
        out_ll[seq0 + s] = log(acc_a) + scale;
        out_logmass[seq0 + s] = log(acc_m) + scale;
    }
}

}  // namespace

#define PNFA_KERNEL(NAME, BF16, SPLIT)                                                                            \
    extern "C" __global__ void __launch_bounds__(PNFA_BLOCK) __attribute__((amdgpu_waves_per_eu(1, 1)))          \
    NAME(const uint16_t* __restrict__ tlp, const float* __restrict__ colscale, const float* __restrict__ t32,     \
         int n_sym, const float* __restrict__ s0, int e0, const float* __restrict__ accw,                         \
         const uint8_t* __restrict__ symbols, int64_t n_seq, int L, double* __restrict__ out_ll,                  \
         double* __restrict__ out_logmass) {                                                                      \
        pnfa_forward<BF16, SPLIT>(tlp, colscale, t32, n_sym, s0, e0, accw, symbols, n_seq, L, out_ll, out_logmass); \
    }

PNFA_KERNEL(pnfa64_f16x2, false, true)
PNFA_KERNEL(pnfa64_bf16x2, true, true)
PNFA_KERNEL(pnfa64_f16, false, false)
PNFA_KERNEL(pnfa64_bf16, true, false)
