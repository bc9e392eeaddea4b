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
//    workgroup of W waves covers 32 symbols, wave w owns the SPW = 32 / W
//    symbols SPW*w .. SPW*w + SPW-1 (8 fragments of 16 x 32 per symbol);
//  * keeps the 64 sequences' states in LDS in FP32; before the product each
//    state is multiplied by the symbol's column factors
//    colsum(T[c]) / colsum(round(T[c])), so the rounded matrix moves exactly
//    the mass the FP32 one does, then split into a high and a low half (two
//    MFMAs per tile, about 22 bits of state) or rounded once;
//  * each step, every wave gathers the sequences whose current symbol it owns
//    into 16-column tiles (the n-th set bit of a ballot mask -> column n, by
//    one ds_permute of the ranks) whose
//    k dimension runs over all of the wave's matrices, each column's state
//    placed in its own symbol's k block and zeros elsewhere; it runs
//    v_mfma_f32_16x16x32_{f16,bf16} with FP32 accumulation, rescales each
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

#include "prob_common.h"

namespace {

using namespace pnfadev;

constexpr int kStates = 64;
constexpr int kRow = kStates + 4;   // padded LDS row (floats) against bank conflicts
constexpr int kSeqPerGroup = 64;
constexpr int kMaxSymbols = 32;
constexpr int kChunk = 64;          // symbols staged per sequence at a time

template <bool BF16, bool SPLIT, int SPW>
__device__ void pnfa_forward(const uint16_t* __restrict__ tlp, const float* __restrict__ colscale,
                             const float* __restrict__ t32, int n_sym, const float* __restrict__ s0, int e0,
                             const float* __restrict__ accw, const uint8_t* __restrict__ symbols,
                             const int64_t* __restrict__ sbegin, const int64_t* __restrict__ send, int64_t n_seq,
                             int L, double* __restrict__ out_ll, double* __restrict__ out_logmass)
{
    typedef typename Lp<BF16>::vec vec;
    constexpr int kBlock = 64 * (kMaxSymbols / SPW);
    constexpr int kTps = kBlock / kSeqPerGroup;   // threads per sequence in the last step
    constexpr int kStage = kChunk + 1;            // staged symbols per sequence: a chunk plus the next one
    constexpr uint8_t kDone = 0xfe;               // no rounded step at this position (the sequence has ended)
    constexpr uint8_t kBad = 0xff;                // symbol outside the alphabet: the state empties
    // One state row per sequence, updated in place: a row is read and then
    // rewritten by the one wave that owns its sequence's symbol at that step.
    __shared__ __attribute__((aligned(16))) float state[kSeqPerGroup][kRow];
    // per-symbol rows padded to kRow floats: lanes of one read that name
    // different symbols start 16 bytes apart in the banks instead of on the
    // same ones (a 64-float row is exactly one turn of the 64 banks)
    // column factors, high part and FP32 rest; row kMaxSymbols is the identity
    // (1, 0), read when the next step takes no factors, so that read needs
    // no branch
    __shared__ __attribute__((aligned(16))) float dscale[kMaxSymbols + 1][kRow];
    __shared__ __attribute__((aligned(16))) float dscale_lo[kMaxSymbols + 1][kRow];
    __shared__ __attribute__((aligned(16))) float dcq[kMaxSymbols][kRow];         // colsum(round T), high
    __shared__ __attribute__((aligned(16))) float dcq_lo[kMaxSymbols][kRow];      // and rest
    __shared__ uint8_t syms[kSeqPerGroup][kChunk + 4];
    __shared__ int32_t estep[2][kSeqPerGroup];   // exponent taken out at a step, by step parity
    __shared__ int32_t expo[kSeqPerGroup];
    __shared__ int64_t sq_beg[kSeqPerGroup];
    __shared__ int32_t sq_len[kSeqPerGroup];
    __shared__ int32_t max_len;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int64_t seq0 = (int64_t)blockIdx.x * kSeqPerGroup;
    const int nseq = (int)min<int64_t>(kSeqPerGroup, n_seq - seq0);
    const int koff = 8 * (lane >> 4);
    const int rsub = 4 * (lane >> 4);   // first of the four result rows of this lane in a row block

    // A fragments of this wave's symbols: A[q][rb][kb] holds rows rb*16 + (lane & 15),
    // columns kb*32 + koff + e
    vec A[SPW][4][2];
#pragma unroll
    for (int q = 0; q < SPW; ++q) {
        const int c = wave * SPW + q;
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
    if (tid == 0) max_len = 0;
    __syncthreads();
    if (tid < kSeqPerGroup) {
        int64_t b = 0, len = 0;
        if (tid < nseq) {
            const int64_t j = seq0 + tid;
            b = sbegin ? sbegin[j] : j * (int64_t)L;
            len = sbegin ? send[j] - b : (int64_t)L;
        }
        sq_beg[tid] = b;
        sq_len[tid] = (int32_t)len;
        atomicMax(&max_len, (int32_t)len);
    }
    auto stage = [&](int t) {
        for (int i = tid; i < kSeqPerGroup * kStage; i += kBlock) {
            const int s = i / kStage, k = i % kStage;
            uint8_t v = kDone;
            if (t + k < sq_len[s] - 1) {
                const uint8_t c = symbols[sq_beg[s] + t + k];
                v = c < n_sym ? c : kBad;
            }
            syms[s][k] = v;
        }
    };
    for (int i = tid; i < (kMaxSymbols + 1) * kStates; i += kBlock) {
        const bool in = i < n_sym * kStates;
        dscale[i / kStates][i % kStates] = in ? colscale[i] : 1.f;
        dscale_lo[i / kStates][i % kStates] = in ? colscale[n_sym * kStates + i] : 0.f;
        if (i < kMaxSymbols * kStates) {
            dcq[i / kStates][i % kStates] = in ? colscale[2 * n_sym * kStates + i] : 0.f;
            dcq_lo[i / kStates][i % kStates] = in ? colscale[3 * n_sym * kStates + i] : 0.f;
        }
    }
    if (tid < 2 * kSeqPerGroup) estep[tid / kSeqPerGroup][tid % kSeqPerGroup] = 0;
    __syncthreads();
    stage(0);
    __syncthreads();
    // The state kept in LDS is already multiplied by the column factors of the
    // symbol it meets next (when that step runs on the rounded matrices).
    for (int i = tid; i < kSeqPerGroup * kStates; i += kBlock) {
        const int s = i / kStates, k = i % kStates;
        const int c0 = syms[s][0];
        state[s][k] = c0 < n_sym ? fmaf(s0[k], dscale[c0][k], s0[k] * dscale_lo[c0][k]) : s0[k];
    }
    int expo_home = e0;   // wave 0, lane j: log2-scale of sequence j
    const int steps = max_len - 1;
    __syncthreads();

    for (int t = 0; t < steps; ++t) {
        if (t % kChunk == 0 && t > 0) {
            stage(t);
            __syncthreads();
        }
        const int par = t & 1;
        const int tc = t % kChunk;
        if (wave == 0 && t > 0) expo_home += estep[par ^ 1][lane];
        const int cj = syms[lane][tc];
        if (wave == 0 && cj >= n_sym) {
            if (cj == kBad) {   // a symbol outside the alphabet empties the state
                float4* row = reinterpret_cast<float4*>(&state[lane][0]);
                for (int k = 0; k < kStates / 4; ++k) row[k] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
            estep[par][lane] = 0;
        }
        // One tile holds up to 16 sequences on any of this wave's symbols: the
        // k dimension runs over the SPW symbols' matrices (SPW x 64), and each
        // column's state sits only in its own symbol's k block.
        const bool sel = cj < n_sym && cj / SPW == wave;
        const uint64_t m = __ballot(sel);
        const int cnt = __popcll(m);
        // column n of the wave's tiles = the n-th selected sequence: each lane
        // (a sequence) pushes its index to lane rank-among-the-selected, the
        // others to lanes past cnt (a permutation, one ds_permute)
        const int below = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        const int dst = sel ? below : cnt + (lane - below);
        const int colseq = __builtin_amdgcn_ds_permute(dst << 2, lane);
        for (int g = 0; g * 16 < cnt; ++g) {
            const int n = g * 16 + (lane & 15);
            const bool valid = n < cnt;
            const int seq = __shfl(colseq, valid ? n : g * 16, 64);
            const int c = syms[seq][tc];
            const int cn = syms[seq][tc + 1];
            const int myq = valid ? c - wave * SPW : -1;
            vec bhi[2], blo[2];
            float mass_in = 0.f;   // sum_k colsum(round T[c])_k x_k over this lane's 16 k
#pragma unroll
            for (int kb = 0; kb < 2; ++kb) {
                const float4* src = reinterpret_cast<const float4*>(&state[seq][kb * 32 + koff]);
                const float4 x0 = src[0], x1 = src[1];
                const float x[8] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
                const float4* qh = reinterpret_cast<const float4*>(&dcq[c][kb * 32 + koff]);
                const float4* ql = reinterpret_cast<const float4*>(&dcq_lo[c][kb * 32 + koff]);
                const float4 qa = qh[0], qb = qh[1], ra = ql[0], rb2 = ql[1];
                const float q[8] = {qa.x, qa.y, qa.z, qa.w, qb.x, qb.y, qb.z, qb.w};
                const float r[8] = {ra.x, ra.y, ra.z, ra.w, rb2.x, rb2.y, rb2.z, rb2.w};
#pragma unroll
                for (int e = 0; e < 8; ++e) mass_in = fmaf(x[e], q[e], fmaf(x[e], r[e], mass_in));
                split8<BF16, SPLIT>(x, bhi[kb], blo[kb]);
            }
            f32x4 d[4];
#pragma unroll
            for (int rb = 0; rb < 4; ++rb) d[rb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int q = 0; q < SPW; ++q) {
                const bool mine = myq == q;
                if (__ballot(mine) == 0) continue;   // no column of this tile on symbol q
                const vec zero{};
                vec bh[2], bl[2];
#pragma unroll
                for (int kb = 0; kb < 2; ++kb) {
                    bh[kb] = mine ? bhi[kb] : zero;
                    if constexpr (SPLIT) bl[kb] = mine ? blo[kb] : zero;
                }
#pragma unroll
                for (int kb = 0; kb < 2; ++kb)
#pragma unroll
                    for (int rb = 0; rb < 4; ++rb) {
                        d[rb] = Lp<BF16>::mfma(A[q][rb][kb], bh[kb], d[rb]);
                        if constexpr (SPLIT) d[rb] = Lp<BF16>::mfma(A[q][rb][kb], bl[kb], d[rb]);
                    }
            }
            // column factors of the next symbol for this lane's 16 result rows
            // (none when the next step is the last one or the sequence ends)
            float4 f[4], fl[4];
            const int fr = cn < n_sym ? cn : kMaxSymbols;   // the identity row otherwise
#pragma unroll
            for (int rb = 0; rb < 4; ++rb) {
                f[rb] = *reinterpret_cast<const float4*>(&dscale[fr][rb * 16 + rsub]);
                fl[rb] = *reinterpret_cast<const float4*>(&dscale_lo[fr][rb * 16 + rsub]);
            }
            float mx = 0.f, mass_out = 0.f;
#pragma unroll
            for (int rb = 0; rb < 4; ++rb) {
                mx = fmaxf(mx, fmaxf(fmaxf(fabsf(d[rb][0]), fabsf(d[rb][1])), fmaxf(fabsf(d[rb][2]), fabsf(d[rb][3]))));
                mass_out += (d[rb][0] + d[rb][1]) + (d[rb][2] + d[rb][3]);
            }
            // the product's mass back to what its operands carry: the matrix cores'
            // FP32 accumulation loses a little mass on every step, in one direction
            mass_in = column_sum(mass_in);
            mass_out = column_sum(mass_out);
            const float fix = mass_out > 0.f && mass_in > 0.f ? mass_in / mass_out : 1.f;
            mx = column_max(mx) * fix;
            const int e2 = mx > 0.f ? __builtin_amdgcn_frexp_expf(mx) : 0;   // mx = f * 2^e2, f in [0.5, 1)
            if (valid) {
                // fix and the power of two in one factor: scaling by 2^-e2
                // commutes with each rounding, so this equals applying it last
                const float sc = ldexpf(fix, -e2);
#pragma unroll
                for (int rb = 0; rb < 4; ++rb) {
                    const f32x4 dv = d[rb] * sc;
                    const float4 out = make_float4(fmaf(dv[0], f[rb].x, dv[0] * fl[rb].x),
                                                   fmaf(dv[1], f[rb].y, dv[1] * fl[rb].y),
                                                   fmaf(dv[2], f[rb].z, dv[2] * fl[rb].z),
                                                   fmaf(dv[3], f[rb].w, dv[3] * fl[rb].w));
                    *reinterpret_cast<float4*>(&state[seq][rb * 16 + rsub]) = out;
                }
                if ((lane >> 4) == 0) estep[par][seq] = e2;
            }
        }
        __syncthreads();
    }
    if (wave == 0) {
        if (steps >= 1) expo_home += estep[(steps - 1) & 1][lane];
        expo[lane] = expo_home;
    }
    __syncthreads();

    // last step in FP32 from the unrounded T: kTps threads per sequence
    constexpr int kRows = kStates / kTps;
    const int s = tid / kTps, part = tid % kTps;
    double acc_a = 0.0, acc_m = 0.0;
    if (s < nseq) {
        const float* st = state[s];
        if (sq_len[s] >= 1) {
            const int c = symbols[sq_beg[s] + sq_len[s] - 1];
            if (c < n_sym) {
                for (int r = part * kRows; r < part * kRows + kRows; ++r) {
                    const float* trow = t32 + ((size_t)c * kStates + r) * kStates;
                    float v = 0.f;
                    for (int k = 0; k < kStates; ++k) v = fmaf(trow[k], st[k], v);
                    acc_a += (double)v * (double)accw[r];
                    acc_m += (double)v;
                }
            }
        } else {
            for (int r = part * kRows; r < part * kRows + kRows; ++r) {
                acc_a += (double)st[r] * (double)accw[r];
                acc_m += (double)st[r];
            }
        }
    }
#pragma unroll
    for (int o = 1; o < kTps; o <<= 1) {
        acc_a += __shfl_xor(acc_a, o, 64);
        acc_m += __shfl_xor(acc_m, o, 64);
    }
    if (s < nseq && part == 0) {
        const double scale = (double)expo[s] * 0.69314718055994530942;   // ln 2
        out_ll[seq0 + s] = log(acc_a) + scale;
        out_logmass[seq0 + s] = log(acc_m) + scale;
    }
}

}  // namespace

// Kernels: pnfa64_<mode>_w<W>, W = waves per workgroup (32 / symbols per wave);
// one workgroup of 64 * W threads per 64 sequences. Sequence j is
// symbols[sbegin[j] .. send[j]), or symbols[j * L .. (j + 1) * L) when sbegin
// is null.
#define PNFA_KERNEL(NAME, BF16, SPLIT, SPW)                                                                       \
    extern "C" __global__ void __launch_bounds__(64 * (32 / SPW))                                                  \
    __attribute__((amdgpu_waves_per_eu(32 / SPW / 4, 32 / SPW / 4)))                                              \
    NAME(const uint16_t* __restrict__ tlp, const float* __restrict__ colscale, const float* __restrict__ t32,     \
         int n_sym, const float* __restrict__ s0, int e0, const float* __restrict__ accw,                         \
         const uint8_t* __restrict__ symbols, const int64_t* __restrict__ sbegin,                                 \
         const int64_t* __restrict__ send, int64_t n_seq, int L, double* __restrict__ out_ll,                     \
         double* __restrict__ out_logmass) {                                                                      \
        pnfa_forward<BF16, SPLIT, SPW>(tlp, colscale, t32, n_sym, s0, e0, accw, symbols, sbegin, send, n_seq, L,  \
                                       out_ll, out_logmass);                                                      \
    }

#define PNFA_KERNELS(W, SPW)                                  \
    PNFA_KERNEL(pnfa64_f16x2_w##W, false, true, SPW)          \
    PNFA_KERNEL(pnfa64_bf16x2_w##W, true, true, SPW)          \
    PNFA_KERNEL(pnfa64_f16_w##W, false, false, SPW)           \
    PNFA_KERNEL(pnfa64_bf16_w##W, true, false, SPW)

PNFA_KERNELS(4, 8)
PNFA_KERNELS(8, 4)
