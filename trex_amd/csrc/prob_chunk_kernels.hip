// Parallel-in-time probabilistic NFA-64 (gfx950): long sequences split into
// chunks whose 64 x 64 products run concurrently, then chained per sequence.
//
// The forward pass of one sequence is s_{L-1} = M(c_{L-2}) ... M(c_0) s_0
// followed by one FP32 step with the unrounded T[c_{L-1}], where
// M(c) = round(T[c]) diag(f[c]) and f[c] are the column mass factors (the
// operator the batch kernels in prob_kernels.hip apply on every step but the
// last). A chunk of positions [b, b + n) has the product
//     P = M(c_{b+n-1}) ... M(c_b),
// so the chunks of a sequence are independent, and
//     s_{L-1} = P_J ... P_1 s_0.
//
// pnfa64_chunk_<mode>: one wave per chunk. Every column of P evolves like a
// state vector under the same symbol, so each step is one dense
// 64 x 64 x 64 product: 4 (row blocks) x 4 (column blocks) x 2 (k blocks) x
// 2 (high / low half of the state) v_mfma_f32_16x16x32 with FP32
// accumulation. The k order of the operands is chosen so that the result
// layout is the next step's B operand: lane l = 16 g + n of column block cb
// holds rows 16 rb + 4 g + v (rb, v < 4) of column 16 cb + n, and k position
// (g, e) of k block kb stands for source state 16 (2 kb + e / 4) + 4 g + e % 4,
// which the same lane holds. The product never leaves registers and the
// waves of a workgroup never wait for each other. The A operand (the rounded
// T[c], 8 KB per symbol, stored on the host in this k order) is read from
// global memory (L2-resident) each step, the next step's loads issued before
// the current product. Each column is rescaled by an exact power of two per
// step (largest entry into [0.5, 1)); the exponents are summed per column.
//
// pnfa64_combine: one wave per sequence, lane r holding s_r in FP64:
// s <- P_j (2^{E_j} * s) chunk by chunk (exponents aligned to the largest
// before the product, the state renormalised after it, the log2-scale kept as
// an integer), then the last step with the unrounded T in FP64 and
// LL = log(a . s_L), log-mass = log(1 . s_L), plus the scale times ln 2.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "prob_common.h"

namespace {

using namespace pnfadev;

constexpr int kStates = 64;
constexpr int kWavesPerBlock = 4;

template <bool BF16>
__device__ __forceinline__ void load_afrag(const uint4* __restrict__ afrag, uint32_t c, int lane,
                                           typename Lp<BF16>::vec (&A)[4][2]) {
    typedef typename Lp<BF16>::vec vec;
    const uint4* base = afrag + (size_t)c * (4 * 2 * 64) + lane;
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) A[rb][kb] = __builtin_bit_cast(vec, base[(rb * 2 + kb) * 64]);
}

template <bool BF16, bool SPLIT>
__device__ void chunk_product(const uint4* __restrict__ afrag, const float* __restrict__ colscale, int n_sym,
                              const uint8_t* __restrict__ symbols, const int64_t* __restrict__ cbeg,
                              const int32_t* __restrict__ clen, int64_t n_chunks, float* __restrict__ out_p,
                              int32_t* __restrict__ out_e)
{
    typedef typename Lp<BF16>::vec vec;
    const int lane = threadIdx.x & 63;
    const int64_t chunk = (int64_t)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (chunk >= n_chunks) return;   // the whole wave leaves
    const int g = lane >> 4, n = lane & 15;

    // D[cb][rb][v]: P[16 rb + 4 g + v][16 cb + n]; start from the identity
    f32x4 D[4][4];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
            for (int v = 0; v < 4; ++v) D[cb][rb][v] = (16 * rb + 4 * g + v == 16 * cb + n) ? 1.f : 0.f;
    int E[4] = {0, 0, 0, 0};

    const uint8_t* sp = symbols + cbeg[chunk];
    const int len = clen[chunk];
    // per-symbol column vectors, four blocks of n_sym * 64 floats: factor
    // (high, rest), column sum of the rounded matrix (high, rest)
    const int blk = n_sym * kStates;
    vec A[4][2], An[4][2];
    float4 F[4], Fn[4], Fl[4], Fln[4], Q[4], Qn[4], Ql[4], Qln[4];
    auto load_cols = [&](uint32_t s, float4 (&f)[4], float4 (&fl)[4], float4 (&q)[4], float4 (&ql)[4]) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float* at = colscale + s * kStates + 16 * r + 4 * g;
            f[r] = *reinterpret_cast<const float4*>(at);
            fl[r] = *reinterpret_cast<const float4*>(at + blk);
            q[r] = *reinterpret_cast<const float4*>(at + 2 * blk);
            ql[r] = *reinterpret_cast<const float4*>(at + 3 * blk);
        }
    };
    uint32_t c = len > 0 ? sp[0] : 0u;
    if (len > 0 && c < (uint32_t)n_sym) {
        load_afrag<BF16>(afrag, c, lane, A);
        load_cols(c, F, Fl, Q, Ql);
    }
    for (int t = 0; t < len; ++t) {
        if (c >= (uint32_t)n_sym) {   // a symbol outside the alphabet empties every column
#pragma unroll
            for (int cb = 0; cb < 4; ++cb)
#pragma unroll
                for (int rb = 0; rb < 4; ++rb) D[cb][rb] = f32x4{0.f, 0.f, 0.f, 0.f};
            break;
        }
        // next step's operands, in flight during this step's products
        const uint32_t cn = t + 1 < len ? sp[t + 1] : (uint32_t)n_sym;
        if (cn < (uint32_t)n_sym) {
            load_afrag<BF16>(afrag, cn, lane, An);
            load_cols(cn, Fn, Fln, Qn, Qln);
        }
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
            vec bh[2], bl[2];
            float mass_in = 0.f;   // sum_k colsum(round T)_k x_k over this lane's 16 k
#pragma unroll
            for (int kb = 0; kb < 2; ++kb) {
                const f32x4 d0 = D[cb][2 * kb], d1 = D[cb][2 * kb + 1];
                const float4 f0 = F[2 * kb], f1 = F[2 * kb + 1], g0 = Fl[2 * kb], g1 = Fl[2 * kb + 1];
                const float x[8] = {fmaf(d0[0], f0.x, d0[0] * g0.x), fmaf(d0[1], f0.y, d0[1] * g0.y),
                                    fmaf(d0[2], f0.z, d0[2] * g0.z), fmaf(d0[3], f0.w, d0[3] * g0.w),
                                    fmaf(d1[0], f1.x, d1[0] * g1.x), fmaf(d1[1], f1.y, d1[1] * g1.y),
                                    fmaf(d1[2], f1.z, d1[2] * g1.z), fmaf(d1[3], f1.w, d1[3] * g1.w)};
                const float4 q0 = Q[2 * kb], q1 = Q[2 * kb + 1], r0 = Ql[2 * kb], r1 = Ql[2 * kb + 1];
                const float qh[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
                const float ql[8] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w};
#pragma unroll
                for (int e = 0; e < 8; ++e) mass_in = fmaf(x[e], qh[e], fmaf(x[e], ql[e], mass_in));
                split8<BF16, SPLIT>(x, bh[kb], bl[kb]);
            }
            f32x4 acc[4];
#pragma unroll
            for (int rb = 0; rb < 4; ++rb) {
                acc[rb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kb = 0; kb < 2; ++kb) {
                    acc[rb] = Lp<BF16>::mfma(A[rb][kb], bh[kb], acc[rb]);
                    if constexpr (SPLIT) acc[rb] = Lp<BF16>::mfma(A[rb][kb], bl[kb], acc[rb]);
                }
            }
            float mx = 0.f, mass_out = 0.f;
#pragma unroll
            for (int rb = 0; rb < 4; ++rb) {
                mx = fmaxf(mx, fmaxf(fmaxf(fabsf(acc[rb][0]), fabsf(acc[rb][1])),
                                     fmaxf(fabsf(acc[rb][2]), fabsf(acc[rb][3]))));
                mass_out += (acc[rb][0] + acc[rb][1]) + (acc[rb][2] + acc[rb][3]);
            }
            // the product's mass back to what the operands carry (ratio 1 - O(1e-7))
            mass_in = column_sum(mass_in);
            mass_out = column_sum(mass_out);
            const float fix = mass_out > 0.f && mass_in > 0.f ? mass_in / mass_out : 1.f;
            mx = column_max(mx) * fix;
            const int e2 = mx > 0.f ? __builtin_amdgcn_frexp_expf(mx) : 0;
#pragma unroll
            for (int rb = 0; rb < 4; ++rb)
#pragma unroll
                for (int v = 0; v < 4; ++v) D[cb][rb][v] = ldexpf(acc[rb][v] * fix, -e2);
            E[cb] += e2;
        }
        c = cn;
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
            for (int kb = 0; kb < 2; ++kb) A[rb][kb] = An[rb][kb];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            F[r] = Fn[r];
            Fl[r] = Fln[r];
            Q[r] = Qn[r];
            Ql[r] = Qln[r];
        }
    }

    float* P = out_p + chunk * (kStates * kStates);
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
            for (int v = 0; v < 4; ++v) P[(16 * rb + 4 * g + v) * kStates + 16 * cb + n] = D[cb][rb][v];
    if (g == 0) {
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) out_e[chunk * kStates + 16 * cb + n] = E[cb];
    }
}

__device__ __forceinline__ double wave_max(double x) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) x = fmax(x, __shfl_xor(x, o, 64));
    return x;
}

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

__device__ __forceinline__ int wave_max_int(int x) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) x = max(x, __shfl_xor(x, o, 64));
    return x;
}

}  // namespace

// Chunk products: chunk i covers symbols[cbeg[i] .. cbeg[i] + clen[i]);
// out_p [n_chunks][64][64] (row = destination), out_e [n_chunks][64]:
// the product is out_p[i] * diag(2^out_e[i]).
#define PNFA_CHUNK_KERNEL(NAME, BF16, SPLIT)                                                                      \
    extern "C" __global__ void __launch_bounds__(64 * kWavesPerBlock)                                             \
    NAME(const uint4* __restrict__ afrag, const float* __restrict__ colscale, int n_sym,                         \
         const uint8_t* __restrict__ symbols, const int64_t* __restrict__ cbeg, const int32_t* __restrict__ clen, \
         int64_t n_chunks, float* __restrict__ out_p, int32_t* __restrict__ out_e) {                             \
        chunk_product<BF16, SPLIT>(afrag, colscale, n_sym, symbols, cbeg, clen, n_chunks, out_p, out_e);        \
    }

PNFA_CHUNK_KERNEL(pnfa64_chunk_f16x2, false, true)
PNFA_CHUNK_KERNEL(pnfa64_chunk_bf16x2, true, true)
PNFA_CHUNK_KERNEL(pnfa64_chunk_f16, false, false)
PNFA_CHUNK_KERNEL(pnfa64_chunk_bf16, true, false)

// Sequence j = symbols[sbegin[j] .. send[j]); its chunks are
// first[j] .. first[j] + count[j] - 1 (the rounded steps, all but the last
// symbol). s0 = s0n * 2^e0; t32 [n_sym][64][64] unrounded; accw [64].
extern "C" __global__ void __launch_bounds__(64 * kWavesPerBlock)
pnfa64_combine(const float* __restrict__ chunk_p, const int32_t* __restrict__ chunk_e,
               const int64_t* __restrict__ first, const int32_t* __restrict__ count,
               const float* __restrict__ s0n, int e0, const float* __restrict__ t32, const float* __restrict__ accw,
               int n_sym, const uint8_t* __restrict__ symbols, const int64_t* __restrict__ sbegin,
               const int64_t* __restrict__ send, int64_t n_seq, double* __restrict__ out_ll,
               double* __restrict__ out_logmass)
{
    __shared__ double vs[kWavesPerBlock][kStates];
    const int r = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int64_t j = (int64_t)blockIdx.x * kWavesPerBlock + w;
    if (j >= n_seq) return;   // the whole wave leaves; no workgroup barrier below
    double s = (double)s0n[r];
    int64_t scale = e0;
    const int64_t c0 = first[j];
    const int nc = count[j];
    // software pipeline: chunk k + 1's row of P and exponent load while chunk k is applied
    float4 row[kStates / 4], row_n[kStates / 4];
    int ek = 0, ek_n = 0;
    auto fetch = [&](int k, float4 (&dst)[kStates / 4], int& e) {
        const float4* src = reinterpret_cast<const float4*>(chunk_p + (c0 + k) * (kStates * kStates) + r * kStates);
#pragma unroll
        for (int q = 0; q < kStates / 4; ++q) dst[q] = src[q];
        e = chunk_e[(c0 + k) * kStates + r];
    };
    if (nc > 0) fetch(0, row, ek);
    for (int k = 0; k < nc; ++k) {
        if (k + 1 < nc) fetch(k + 1, row_n, ek_n);
        // v = s * 2^E, aligned to the largest exponent so no entry overflows
        int xs = 0;
        const double m = frexp(s, &xs);
        const int tot = s != 0.0 ? xs + ek : INT32_MIN;
        const int top = wave_max_int(tot);
        if (top == INT32_MIN) break;   // the state is empty: it stays empty
        vs[w][r] = s != 0.0 ? ldexp(m, tot - top) : 0.0;
        __builtin_amdgcn_wave_barrier();
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < kStates / 4; ++q) {
            const float4 p = row[q];
            acc = fma((double)p.x, vs[w][4 * q], acc);
            acc = fma((double)p.y, vs[w][4 * q + 1], acc);
            acc = fma((double)p.z, vs[w][4 * q + 2], acc);
            acc = fma((double)p.w, vs[w][4 * q + 3], acc);
        }
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int q = 0; q < kStates / 4; ++q) row[q] = row_n[q];
        ek = ek_n;
        scale += top;
        const double mx = wave_max(fabs(acc));
        int e = 0;
        if (mx > 0.0) frexp(mx, &e);
        s = ldexp(acc, -e);
        scale += e;
    }
    // last step: the unrounded T of the last symbol, in FP64
    const int64_t len = send[j] - sbegin[j];
    double sl = s;
    if (len >= 1) {
        const uint32_t c = symbols[send[j] - 1];
        vs[w][r] = s;
        __builtin_amdgcn_wave_barrier();
        sl = 0.0;
        if (c < (uint32_t)n_sym) {
            const float* trow = t32 + ((size_t)c * kStates + r) * kStates;
            for (int k = 0; k < kStates; ++k) sl = fma((double)trow[k], vs[w][k], sl);
        }
        __builtin_amdgcn_wave_barrier();
    }
    const double a = wave_sum(sl * (double)accw[r]);
    const double mass = wave_sum(sl);
    if (r == 0) {
        const double ln2 = 0.69314718055994530942;
        out_ll[j] = log(a) + (double)scale * ln2;
        out_logmass[j] = log(mass) + (double)scale * ln2;
    }
}
