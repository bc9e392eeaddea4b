// Fused multi-pattern DFA-16 walk for gfx950: one read of each string feeds
// P resident pattern tables (P = 1..8 per launch).
//
// The single-pattern window kernels (dfa_kernels.hip) stream each string from
// HBM once per pattern; matching P patterns over the same batch reads it P
// times. Here a lane loads each 16-byte vector once and advances P states with
// it, so HBM traffic is that of one pattern and the work per byte grows with
// P in LDS reads and VALU (the patterns' steps are independent, so P chains
// also hide each other's LDS latency). This is the capability of the
// reference's PackedEngine (src/packed_engine.py:25-126: several patterns per
// pass over the input), on the DFA-16 window form.
//
// Every pattern of a launch is a DFA of at most 16 states whose live bytes
// lie in one aligned 32-byte window [base_p, base_p + 32) (bytes outside it
// lead every state to the absorbing dead state). The pair index
// sym0 | sym1 << 5 (sym = byte & 31) does not depend on the window, and the
// window check (WinCheck) only accumulates the bytes, so one check serves
// all patterns: at the end each pattern whose window saw an outside byte is
// forced to its dead state. Head and tail bytes (outside whole 16-byte
// vectors) take single-symbol steps and enter the same check.
//
// Two table forms (the single-pattern k2f / k2w):
//   F  pair functions f2[p][pair] (uint64, 16 nibbles, 8 KB per pattern):
//      the LDS address does not depend on the state; s = (f >> 4s) & 15
//   W  state walk t2[p][pair * 16 + s] (uint8, 16 KB per pattern): one byte
//      read, the address depends on the state
// and single-symbol tables f1[p][sym] (uint64) / t1[p][sym * 16 + s] (uint8).
//
// Layouts: begin/end ranges over contiguous strings (CSR or staggered), or
// the interleaved layout (64-string groups stored as 1 KiB rows; see
// dfa_kernels.hip). Outputs are pattern-major: out_state[p * n + i],
// out_accept[p * n + i].

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.h"

#define TREX_MULTI_BLOCK 256

namespace {

using namespace trexdev;

constexpr int kMultiMax = 8;     // patterns per launch
constexpr int kIlGroup = 64;   // strings per group of the interleaved layout

}  // namespace

// Per-pattern scalars of one launch (kernel argument, read through the
// scalar unit).
struct MultiParams {
    uint32_t win4[kMultiMax];     // window base replicated in four bytes
    uint32_t start[kMultiMax];
    uint32_t dead[kMultiMax];
    uint32_t accept[kMultiMax];   // bit s set = state s accepts
};

namespace {

// Each pattern's pair table sits in LDS at a stride of kPairBytes + 16: with
// a stride that is a multiple of 128 bytes the compiler pairs the reads of
// two patterns into one ds_read2st64, and since both entries of a lane then
// lie in the same bank, every such read conflicts with itself (measured: 24
// cycles per paired read, against 7 for a single read).
constexpr int kPad = 16;

struct MultiF {
    static constexpr int kPairBytes = 1024 * 8;
    static constexpr int kSymBytes = 32 * 8;
    __device__ __forceinline__ static uint32_t pair(const uint8_t* __restrict__ t, uint32_t s, uint32_t x) {
        const uint64_t f = reinterpret_cast<const uint64_t*>(t)[x];
        return (uint32_t)(f >> (s << 2)) & 15u;
    }
    __device__ __forceinline__ static uint32_t sym(const uint8_t* __restrict__ t, uint32_t s, uint32_t c) {
        const uint64_t f = reinterpret_cast<const uint64_t*>(t)[c];
        return (uint32_t)(f >> (s << 2)) & 15u;
    }
};

struct MultiW {
    static constexpr int kPairBytes = 1024 * 16;
    static constexpr int kSymBytes = 32 * 16;
    __device__ __forceinline__ static uint32_t pair(const uint8_t* __restrict__ t, uint32_t s, uint32_t x) {
        return t[(x << 4) | s];
    }
    __device__ __forceinline__ static uint32_t sym(const uint8_t* __restrict__ t, uint32_t s, uint32_t c) {
        return t[(c << 4) | s];
    }
};

template <class Form, int P, int PF, bool IL>
__device__ __forceinline__ void multi_kernel(const uint8_t* __restrict__ bytes, const int64_t* __restrict__ begin,
                                           const int64_t* __restrict__ endp, int64_t n_strings,
                                           const uint8_t* __restrict__ pairg, const uint8_t* __restrict__ symg,
                                           const MultiParams& prm, uint8_t* __restrict__ out_state,
                                           uint8_t* __restrict__ out_accept)
{
    constexpr int kStride = Form::kPairBytes + kPad;
    __shared__ __attribute__((aligned(16))) uint8_t tp[P * kStride];
    __shared__ __attribute__((aligned(16))) uint8_t ts[P * Form::kSymBytes];
#pragma unroll
    for (int k = 0; k < P; ++k) stage_table(tp + k * kStride, pairg + (size_t)k * Form::kPairBytes, Form::kPairBytes);
    stage_table(ts, symg, P * Form::kSymBytes);
    __syncthreads();

    const int64_t sid = (int64_t)blockIdx.x * TREX_MULTI_BLOCK + threadIdx.x;
    if (sid >= n_strings) return;
    int64_t p = begin[sid];
    const int64_t end = endp[sid];
    uint32_t s[P];
#pragma unroll
    for (int k = 0; k < P; ++k) s[k] = prm.start[k];
    WinCheck w; w.init();

    auto word = [&](uint32_t x) {
        w.add(x);
        const uint32_t lo = pair_lo(x), hi = pair_hi(x);
#pragma unroll
        for (int k = 0; k < P; ++k) s[k] = Form::pair(tp + k * kStride, s[k], lo);
#pragma unroll
        for (int k = 0; k < P; ++k) s[k] = Form::pair(tp + k * kStride, s[k], hi);
    };
    auto vec = [&](uint4 v) { word(v.x); word(v.y); word(v.z); word(v.w); };
    auto byte = [&](uint32_t b) {
        w.add(b * 0x01010101u);
#pragma unroll
        for (int k = 0; k < P; ++k) s[k] = Form::sym(ts + k * Form::kSymBytes, s[k], b & 31u);
    };

    if constexpr (IL) {
        const int64_t len = end - p;
        const int64_t nvec = len >> 4;
        stream_body_nt<PF, kIlGroup>(reinterpret_cast<const uint4*>(bytes + p), nvec, vec);
        const uint8_t* tail = bytes + p + nvec * (16 * kIlGroup);
        for (int t = 0; t < (int)(len & 15); ++t) byte(tail[t]);
    } else {
        const int64_t head_end = min(end, p + head_len(bytes, p));
        for (; p < head_end; ++p) byte(bytes[p]);
        const int64_t nvec = (end - p) >> 4;
        stream_body<PF>(reinterpret_cast<const uint4*>(bytes + p), nvec, vec);
        p += nvec << 4;
        for (; p < end; ++p) byte(bytes[p]);
    }

#pragma unroll
    for (int k = 0; k < P; ++k) {
        const uint32_t q = w.outside(prm.win4[k]) ? prm.dead[k] : s[k];
        out_state[(int64_t)k * n_strings + sid] = (uint8_t)q;
        out_accept[(int64_t)k * n_strings + sid] = (uint8_t)((prm.accept[k] >> q) & 1u);
    }
}

// Paired form: patterns 2j and 2j + 1 of a launch share one table, entry x =
// (f2_{2j}[x], f2_{2j+1}[x]) in 16 bytes, so one ds_read_b128 serves both
// (an odd launch's last pattern is paired with the identity). A 16-byte read
// occupies a whole quad of banks; random entries spread over the 16 quads.
template <int P, int PF, bool IL>
__device__ __forceinline__ void multi_kernel_paired(const uint8_t* __restrict__ bytes,
                                                    const int64_t* __restrict__ begin,
                                                    const int64_t* __restrict__ endp, int64_t n_strings,
                                                    const uint8_t* __restrict__ pairg,
                                                    const uint8_t* __restrict__ symg, const MultiParams& prm,
                                                    uint8_t* __restrict__ out_state, uint8_t* __restrict__ out_accept)
{
    constexpr int PP = (P + 1) / 2;
    __shared__ __attribute__((aligned(16))) uint8_t tp[PP * 16384];
    __shared__ __attribute__((aligned(16))) uint8_t ts[P * MultiF::kSymBytes];
    stage_table(tp, pairg, PP * 16384);
    stage_table(ts, symg, P * MultiF::kSymBytes);
    __syncthreads();

    const int64_t sid = (int64_t)blockIdx.x * TREX_MULTI_BLOCK + threadIdx.x;
    if (sid >= n_strings) return;
    int64_t p = begin[sid];
    const int64_t end = endp[sid];
    uint32_t s[2 * PP];
#pragma unroll
    for (int k = 0; k < 2 * PP; ++k) s[k] = k < P ? prm.start[k] : 0u;
    WinCheck w; w.init();

    auto pstep = [&](uint32_t x) {
#pragma unroll
        for (int j = 0; j < PP; ++j) {
            const uint4 v = reinterpret_cast<const uint4*>(tp + j * 16384)[x];
            const uint64_t f0 = (uint64_t)v.x | ((uint64_t)v.y << 32);
            const uint64_t f1 = (uint64_t)v.z | ((uint64_t)v.w << 32);
            s[2 * j] = (uint32_t)(f0 >> (s[2 * j] << 2)) & 15u;
            s[2 * j + 1] = (uint32_t)(f1 >> (s[2 * j + 1] << 2)) & 15u;
        }
    };
    auto word = [&](uint32_t x) {
        w.add(x);
        pstep(pair_lo(x));
        pstep(pair_hi(x));
    };
    auto vec = [&](uint4 v) { word(v.x); word(v.y); word(v.z); word(v.w); };
    auto byte = [&](uint32_t b) {
        w.add(b * 0x01010101u);
#pragma unroll
        for (int k = 0; k < P; ++k) s[k] = MultiF::sym(ts + k * MultiF::kSymBytes, s[k], b & 31u);
    };

    if constexpr (IL) {
        const int64_t len = end - p;
        const int64_t nvec = len >> 4;
        stream_body_nt<PF, kIlGroup>(reinterpret_cast<const uint4*>(bytes + p), nvec, vec);
        const uint8_t* tail = bytes + p + nvec * (16 * kIlGroup);
        for (int t = 0; t < (int)(len & 15); ++t) byte(tail[t]);
    } else {
        const int64_t head_end = min(end, p + head_len(bytes, p));
        for (; p < head_end; ++p) byte(bytes[p]);
        const int64_t nvec = (end - p) >> 4;
        stream_body<PF>(reinterpret_cast<const uint4*>(bytes + p), nvec, vec);
        p += nvec << 4;
        for (; p < end; ++p) byte(bytes[p]);
    }

#pragma unroll
    for (int k = 0; k < P; ++k) {
        const uint32_t q = w.outside(prm.win4[k]) ? prm.dead[k] : s[k];
        out_state[(int64_t)k * n_strings + sid] = (uint8_t)q;
        out_accept[(int64_t)k * n_strings + sid] = (uint8_t)((prm.accept[k] >> q) & 1u);
    }
}

template <int P>
constexpr int paired_waves() {
    const int blocks = (160 * 1024) / (((P + 1) / 2) * 16384 + P * MultiF::kSymBytes);
    return blocks < 1 ? 1 : (blocks > 4 ? 4 : blocks);
}

// Waves per SIMD the LDS footprint allows (160 KB per CU, one wave per SIMD
// per workgroup), at most 4: the register budget the compiler is held to.
template <class Form, int P>
constexpr int multi_waves() {
    const int blocks = (160 * 1024) / (P * (Form::kPairBytes + kPad + Form::kSymBytes));
    return blocks < 1 ? 1 : (blocks > 4 ? 4 : blocks);
}

}  // namespace

// pairg: [P][pair table] of the form, symg: [P][single-symbol table].
#define TREX_MULTI_KERNEL(NAME, FORM, P, IL)                                                                     \
    extern "C" __global__ void __launch_bounds__(TREX_MULTI_BLOCK)                                               \
    __attribute__((amdgpu_waves_per_eu(multi_waves<FORM, P>())))                                                 \
    NAME(const uint8_t* __restrict__ bytes, const int64_t* __restrict__ begin, const int64_t* __restrict__ end, \
         int64_t n_strings, const uint8_t* __restrict__ pairg, const uint8_t* __restrict__ symg, MultiParams prm, \
         uint8_t* __restrict__ out_state, uint8_t* __restrict__ out_accept) {                                 \
        multi_kernel<FORM, P, 8, IL>(bytes, begin, end, n_strings, pairg, symg, prm, out_state, out_accept);    \
    }

#define TREX_MULTI_KERNELS_P(FORM, TAG, P)                            \
    TREX_MULTI_KERNEL(dfa16_multi_##TAG##_p##P, FORM, P, false)         \
    TREX_MULTI_KERNEL(dfa16_multi_##TAG##_p##P##_il, FORM, P, true)

#define TREX_MULTI_KERNELS(FORM, TAG)                                                         \
    TREX_MULTI_KERNELS_P(FORM, TAG, 1) TREX_MULTI_KERNELS_P(FORM, TAG, 2)                       \
    TREX_MULTI_KERNELS_P(FORM, TAG, 3) TREX_MULTI_KERNELS_P(FORM, TAG, 4)                       \
    TREX_MULTI_KERNELS_P(FORM, TAG, 5) TREX_MULTI_KERNELS_P(FORM, TAG, 6)                       \
    TREX_MULTI_KERNELS_P(FORM, TAG, 7) TREX_MULTI_KERNELS_P(FORM, TAG, 8)

TREX_MULTI_KERNELS(MultiF, k2f)
TREX_MULTI_KERNELS(MultiW, k2w)

// Paired form. pairg: [(P + 1) / 2][1024] x 16 bytes, symg: [P][32] uint64.
#define TREX_PAIRED_KERNEL(NAME, P, IL)                                                                        \
    extern "C" __global__ void __launch_bounds__(TREX_MULTI_BLOCK)                                             \
    __attribute__((amdgpu_waves_per_eu(paired_waves<P>())))                                                   \
    NAME(const uint8_t* __restrict__ bytes, const int64_t* __restrict__ begin, const int64_t* __restrict__ end, \
         int64_t n_strings, const uint8_t* __restrict__ pairg, const uint8_t* __restrict__ symg, MultiParams prm, \
         uint8_t* __restrict__ out_state, uint8_t* __restrict__ out_accept) {                                 \
        multi_kernel_paired<P, 8, IL>(bytes, begin, end, n_strings, pairg, symg, prm, out_state, out_accept);  \
    }

#define TREX_PAIRED_KERNELS_P(P)                                 \
    TREX_PAIRED_KERNEL(dfa16_multi_k2fp_p##P, P, false)          \
    TREX_PAIRED_KERNEL(dfa16_multi_k2fp_p##P##_il, P, true)

TREX_PAIRED_KERNELS_P(1) TREX_PAIRED_KERNELS_P(2) TREX_PAIRED_KERNELS_P(3) TREX_PAIRED_KERNELS_P(4)
TREX_PAIRED_KERNELS_P(5) TREX_PAIRED_KERNELS_P(6) TREX_PAIRED_KERNELS_P(7) TREX_PAIRED_KERNELS_P(8)
