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
//    one ds_permute of the ranks, worked out during the step before, as it
//    reads only the staged symbols) whose
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
    constexpr int kRing = 2 * kChunk;             // staged symbols per sequence: this chunk and the next
    constexpr uint8_t kDone = 0xfe;               // no rounded step at this position (the sequence has ended)
    // One state row per sequence, updated in place: a row is read and then
    // rewritten by the one wave that owns its sequence's symbol at that step.
    __shared__ __attribute__((aligned(16))) float state[kSeqPerGroup][kRow];
    // per-symbol rows padded to kRow floats: lanes of one read that name
    // different symbols start 16 bytes apart in the banks instead of on the
    // same ones (a 64-float row is exactly one turn of the 64 banks)
    __shared__ __attribute__((aligned(16))) float dscale[kMaxSymbols][kRow];      // column factors, high
    __shared__ __attribute__((aligned(16))) float dscale_lo[kMaxSymbols][kRow];   // and FP32 rest
    __shared__ __attribute__((aligned(16))) float dcq[kMaxSymbols][kRow];         // colsum(round T), high
    __shared__ __attribute__((aligned(16))) float dcq_lo[kMaxSymbols][kRow];      // and rest
    // symbol at position p of sequence s: syms[s][p % kRing] (rows of an odd
    // number of words: lanes reading one position of different sequences hit
    // different banks)
    __shared__ uint8_t syms[kSeqPerGroup][kRing + 4];
    __shared__ int32_t expo[kSeqPerGroup];    // log2-scale of each sequence's state
    __shared__ int32_t empty[kSeqPerGroup];   // a symbol outside the alphabet: the state is zero from there on
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
    // positions p0 .. p0 + kChunk - 1 of every sequence into the ring. A
    // symbol outside the alphabet empties the state, which then stays zero:
    // the sequence is marked and its results are -inf whatever it steps
    // through afterwards (the position itself takes no step).
    auto stage = [&](int p0) {
        for (int i = tid; i < kSeqPerGroup * kChunk; i += kBlock) {
            const int s = i / kChunk, k = i % kChunk;
            uint8_t v = kDone;
            if (p0 + k < sq_len[s] - 1) {
                const uint8_t c = symbols[sq_beg[s] + p0 + k];
                if (c < n_sym) v = c;
                else empty[s] = 1;
            }
            syms[s][(p0 + k) % kRing] = v;
        }
    };
    for (int i = tid; i < kMaxSymbols * kStates; i += kBlock) {
        const bool in = i < n_sym * kStates;
        dscale[i / kStates][i % kStates] = in ? colscale[i] : 1.f;
        dscale_lo[i / kStates][i % kStates] = in ? colscale[n_sym * kStates + i] : 0.f;
        dcq[i / kStates][i % kStates] = in ? colscale[2 * n_sym * kStates + i] : 0.f;
        dcq_lo[i / kStates][i % kStates] = in ? colscale[3 * n_sym * kStates + i] : 0.f;
    }
    if (tid < kSeqPerGroup) {
        expo[tid] = e0;
        empty[tid] = 0;
    }
    __syncthreads();
    stage(0);
    stage(kChunk);
    for (int i = tid; i < kSeqPerGroup * kStates; i += kBlock) state[i / kStates][i % kStates] = s0[i % kStates];
    const int steps = max_len - 1;
    __syncthreads();

    // Which sequence each of this wave's tile columns holds at step t: the
    // sequences whose symbol the wave owns, in lane order. Each lane pushes
    // its index and symbol to lane rank-among-the-selected, the others to
    // lanes past cnt (a permutation, one ds_permute); two lane swaps then
    // copy columns 0..15 to every row of 16 lanes. It reads only the symbol
    // ring, so the next step's routing is worked out during this step's
    // tile, in three parts (route_part).
    struct Route {
        int cj;       // lane j: symbol of sequence j at this step
        int cnt;      // sequences on this wave's symbols
        int colseq;   // lane n: sequence of column n (bits 0-7) and its symbol (bits 8-15)
        int seq, c;   // the first tile's column (lane & 15): its sequence and symbol
    };
    auto route_part = [&](Route& r, int t, int part) {
        if (part == 0) r.cj = syms[lane][t % kRing];
        if (part == 1) {
            const bool sel = r.cj < n_sym && r.cj / SPW == wave;
            const uint64_t m = __ballot(sel);
            r.cnt = __popcll(m);
            const int below =
                (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            const int dst = sel ? below : r.cnt + (lane - below);
            r.colseq = __builtin_amdgcn_ds_permute(dst << 2, lane | (r.cj << 8));
        }
        if (part == 2) {
            const unsigned v0 = (unsigned)r.colseq;
            const auto h = __builtin_amdgcn_permlane32_swap(v0, v0, false, false);   // lanes 32..63 <- 0..31
            const auto q = __builtin_amdgcn_permlane16_swap(h[0], h[0], false, false);   // odd rows <- even rows
            const int v = (int)q[0];
            r.seq = v & 0xff;
            // a lane past the tile's columns reads any row of the symbol tables
            r.c = (lane & 15) < r.cnt ? (v >> 8) & 0xff : wave * SPW;
        }
    };

    // One tile: up to 16 sequences (column n = lane & 15) on any of this
    // wave's symbols; the k dimension runs over the SPW symbols' matrices
    // (SPW x 64), and each column's state sits only in its own symbol's k
    // block. hook(-2) runs at the start, hook(-1) after the operand split,
    // hook(q) after symbol q's products (issued or skipped): the next step's
    // routing goes out between them.
    auto tile = [&](int seq, int c, bool valid, auto&& hook) {
        const int myq = valid ? c - wave * SPW : -1;
        hook(-2);
        vec bhi[2], blo[2];
        float mass_in = 0.f;   // sum_k colsum(round T[c])_k x_k over this lane's 16 k
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) {
            const int k0 = kb * 32 + koff;
            const float4* src = reinterpret_cast<const float4*>(&state[seq][k0]);
            const float4 x0 = src[0], x1 = src[1];
            float x[8] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
            const float4* fh = reinterpret_cast<const float4*>(&dscale[c][k0]);
            const float4* fo = reinterpret_cast<const float4*>(&dscale_lo[c][k0]);
            const float4 fa = fh[0], fb = fh[1], oa = fo[0], ob = fo[1];
            const float f[8] = {fa.x, fa.y, fa.z, fa.w, fb.x, fb.y, fb.z, fb.w};
            const float fl[8] = {oa.x, oa.y, oa.z, oa.w, ob.x, ob.y, ob.z, ob.w};
            const float4* qh = reinterpret_cast<const float4*>(&dcq[c][k0]);
            const float4* ql = reinterpret_cast<const float4*>(&dcq_lo[c][k0]);
            const float4 qa = qh[0], qb = qh[1], ra = ql[0], rb2 = ql[1];
            const float q[8] = {qa.x, qa.y, qa.z, qa.w, qb.x, qb.y, qb.z, qb.w};
            const float rr[8] = {ra.x, ra.y, ra.z, ra.w, rb2.x, rb2.y, rb2.z, rb2.w};
            // the symbol's column factors: the rounded matrix then moves the
            // mass the FP32 one does
#pragma unroll
            for (int e = 0; e < 8; ++e) x[e] = fmaf(x[e], f[e], x[e] * fl[e]);
#pragma unroll
            for (int e = 0; e < 8; ++e) mass_in = fmaf(x[e], q[e], fmaf(x[e], rr[e], mass_in));
            split8<BF16, SPLIT>(x, bhi[kb], blo[kb]);
        }
        hook(-1);
        f32x4 d[4];
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) d[rb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int q = 0; q < SPW; ++q) {
            const bool mine = myq == q;
            if (__ballot(mine) == 0) {   // no column of this tile on symbol q
                hook(q);
                continue;
            }
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
            hook(q);
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
        // the quotient by one Newton step on the reciprocal: within half an
        // ulp and unbiased, as the fix is applied on every step (the
        // reciprocal alone errs in one direction, a drift over 10^5 steps)
        float fix = 1.f;
        if (mass_out > 0.f && mass_in > 0.f) {
            const float rcp = __builtin_amdgcn_rcpf(mass_out);
            const float q0 = mass_in * rcp;
            fix = fmaf(fmaf(-mass_out, q0, mass_in), rcp, q0);
        }
        mx = column_max(mx) * fix;
        const int e2 = mx > 0.f ? __builtin_amdgcn_frexp_expf(mx) : 0;   // mx = f * 2^e2, f in [0.5, 1)
        if (valid) {
            // fix and the power of two in one factor: scaling by 2^-e2
            // commutes with each rounding, so this equals applying it last
            const float sc = ldexpf(fix, -e2);
#pragma unroll
            for (int rb = 0; rb < 4; ++rb) {
                const f32x4 dv = d[rb] * sc;
                *reinterpret_cast<float4*>(&state[seq][rb * 16 + rsub]) = make_float4(dv[0], dv[1], dv[2], dv[3]);
            }
            if ((lane >> 4) == 0) atomicAdd(&expo[seq], e2);   // the step's owner alone
        }
    };

    Route r;
    for (int part = 0; part < 3; ++part) route_part(r, 0, part);
    for (int t = 0; t < steps; ++t) {
        // the chunk after this one, into the half of the ring the previous
        // chunk held (first read at step t + kChunk - 1, barriers apart)
        if (t % kChunk == 0 && t > 0) stage(t + kChunk);
        Route next;
        // the first tile always runs (with no column named it does no
        // products); the next step's routing goes out between its products
        tile(r.seq, r.c, (lane & 15) < r.cnt, [&](int point) {
            if (point < 1) route_part(next, t + 1, point + 2);
        });
        if (r.cnt > 16) {   // more tiles: the column map again (not kept through the first tile)
            Route again;
            route_part(again, t, 0);
            route_part(again, t, 1);
            for (int g = 1; g * 16 < r.cnt; ++g) {
                const int n = g * 16 + (lane & 15);
                const bool valid = n < r.cnt;
                const int v = __shfl(again.colseq, valid ? n : g * 16, 64);
                tile(v & 0xff, (v >> 8) & 0xff, valid, [](int) {});
            }
        }
        __syncthreads();
        r = next;
    }

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
        out_ll[seq0 + s] = empty[s] ? -INFINITY : log(acc_a) + scale;
        out_logmass[seq0 + s] = empty[s] ? -INFINITY : log(acc_m) + scale;
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
