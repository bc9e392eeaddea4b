// Multi-pattern DFA-16 walk with compact pair tables (gfx950).
//
// The multi-pattern kernels (dfa_multi_kernels.hip) read one 8-byte pair
// function per pattern and byte pair, and are bound by those LDS reads. A
// pattern of at most 4 states needs 2 bits per state: its pair function fits
// one byte, so eight such patterns share one 8-byte word. Here a launch holds
// NF patterns with full functions (16 nibbles, one word each) and NB
// patterns of at most 4 states (one byte each, eight per word); entry x of
// the launch's table is NW = NF + ceil(NB / 8) words, stored word pair by word
// pair: table j (j < NW / 2) holds words 2j, 2j+1 of every entry as 16 bytes,
// so each step reads NW / 2 16-byte entries at stride 16 (random entries
// spread over all 16 bank quads), plus one 8-byte entry when NW is odd.
//
// Full pattern k:  s = (w_k >> 4s) & 15
// Byte pattern i:  s = bits [8 (i % 8) + 2s, +2) of w_{NF + i/8} (one bit-field
//                  extract from the word's half that holds the byte)
//
// Head and tail bytes take the 16-nibble single-symbol functions of every
// pattern (a state below 4 reads a nibble below 4). The window check and the
// outputs are those of the other multi-pattern kernels; launch pattern k
// writes row slot[k] of the pattern-major outputs (the host orders a launch's
// patterns full ones first).

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.h"

#define TREX_COMPACT_BLOCK 256

namespace {

using namespace trexdev;

constexpr int kCompactMax = 8;   // patterns per launch
constexpr int kIlGroup = 64;     // strings per group of the interleaved layout

}  // namespace

// Per-pattern scalars of one launch, in launch order.
struct CompactParams {
    uint32_t win4[kCompactMax];
    uint32_t start[kCompactMax];
    uint32_t dead[kCompactMax];
    uint32_t accept[kCompactMax];   // bit s set = state s accepts
    uint32_t slot[kCompactMax];     // output row of the pattern
};

namespace {

template <int NF, int NB, int PF, bool IL>
__device__ __forceinline__ void compact_kernel(const uint8_t* __restrict__ bytes, const int64_t* __restrict__ begin,
                                               const int64_t* __restrict__ endp, int64_t n_strings,
                                               const uint8_t* __restrict__ wordsg, const uint8_t* __restrict__ symg,
                                               const CompactParams& prm, uint8_t* __restrict__ out_state,
                                               uint8_t* __restrict__ out_accept)
{
    constexpr int P = NF + NB;
    constexpr int NW = NF + (NB + 7) / 8;
    constexpr int kPairTables = NW / 2;
    __shared__ __attribute__((aligned(16))) uint8_t tw[1024 * NW * 8];
    __shared__ __attribute__((aligned(16))) uint8_t ts[P * 32 * 8];
    stage_table(tw, wordsg, 1024 * NW * 8);
    stage_table(ts, symg, P * 32 * 8);
    __syncthreads();

    const int64_t sid = (int64_t)blockIdx.x * TREX_COMPACT_BLOCK + threadIdx.x;
    if (sid >= n_strings) return;
    int64_t p = begin[sid];
    const int64_t end = endp[sid];
    uint32_t s[P];
#pragma unroll
    for (int k = 0; k < P; ++k) s[k] = prm.start[k];
    WinCheck w; w.init();

    auto pstep = [&](uint32_t x) {
        uint64_t W[NW];
#pragma unroll
        for (int j = 0; j < kPairTables; ++j) {
            const uint4 v = reinterpret_cast<const uint4*>(tw + j * 16384)[x];
            W[2 * j] = (uint64_t)v.x | ((uint64_t)v.y << 32);
            W[2 * j + 1] = (uint64_t)v.z | ((uint64_t)v.w << 32);
        }
        if constexpr (NW & 1) W[NW - 1] = reinterpret_cast<const uint64_t*>(tw + kPairTables * 16384)[x];
#pragma unroll
        for (int k = 0; k < NF; ++k) s[k] = (uint32_t)(W[k] >> (s[k] << 2)) & 15u;
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            // the byte's dword, then one bit-field extract (offset 8 (i % 4) + 2s)
            const uint64_t wv = W[NF + i / 8];
            const uint32_t dw = (i % 8) < 4 ? (uint32_t)wv : (uint32_t)(wv >> 32);
            s[NF + i] = __builtin_amdgcn_ubfe(dw, 8u * (i % 4) + (s[NF + i] << 1), 2u);
        }
    };
    auto word = [&](uint32_t x) {
        w.add(x);
        // the word's table reads go out once the previous word's steps are
        // done (an empty asm ties the word to a state): one word ahead at
        // most, or the compiler hoists a whole block's reads and runs out of
        // registers
        asm volatile("" : "+v"(x) : "v"(s[0]));
        pstep(pair_lo(x));
        pstep(pair_hi(x));
    };
    auto vec = [&](uint4 v) { word(v.x); word(v.y); word(v.z); word(v.w); };
    auto byte = [&](uint32_t b) {
        w.add(b * 0x01010101u);
#pragma unroll
        for (int k = 0; k < P; ++k) {
            const uint64_t f = reinterpret_cast<const uint64_t*>(ts)[k * 32 + (b & 31u)];
            s[k] = (uint32_t)(f >> (s[k] << 2)) & 15u;
        }
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
        const int64_t row = (int64_t)prm.slot[k] * n_strings + sid;
        out_state[row] = (uint8_t)q;
        out_accept[row] = (uint8_t)((prm.accept[k] >> q) & 1u);
    }
}

// Waves per SIMD the LDS footprint allows (160 KB per CU), at most 4 (3 for
// the one split whose code spills to scratch at 128 registers).
template <int NF, int NB>
constexpr int compact_waves() {
    constexpr int NW = NF + (NB + 7) / 8;
    constexpr int cap = (NF == 2 && NB == 3) ? 3 : 4;
    const int blocks = (160 * 1024) / (1024 * NW * 8 + (NF + NB) * 256);
    return blocks < 1 ? 1 : (blocks > cap ? cap : blocks);
}

}  // namespace

// wordsg: the launch's compact pair tables (NW / 2 tables of 1024 x 16 bytes,
// then 1024 x 8 bytes when NW is odd); symg: [P][32] uint64 single-symbol
// functions in launch order.
#define TREX_COMPACT_KERNEL(NF, NB, IL, SUFFIX)                                                                 \
    extern "C" __global__ void __launch_bounds__(TREX_COMPACT_BLOCK)                                            \
    __attribute__((amdgpu_waves_per_eu(compact_waves<NF, NB>())))                                                \
    dfa16_multi_k2fc_f##NF##_b##NB##SUFFIX(const uint8_t* __restrict__ bytes, const int64_t* __restrict__ begin, \
                                          const int64_t* __restrict__ end, int64_t n_strings,                   \
                                          const uint8_t* __restrict__ wordsg, const uint8_t* __restrict__ symg, \
                                          CompactParams prm, uint8_t* __restrict__ out_state,                  \
                                          uint8_t* __restrict__ out_accept) {                                  \
        compact_kernel<NF, NB, 8, IL>(bytes, begin, end, n_strings, wordsg, symg, prm, out_state, out_accept);  \
    }

#define TREX_COMPACT_BOTH(NF, NB) TREX_COMPACT_KERNEL(NF, NB, false, ) TREX_COMPACT_KERNEL(NF, NB, true, _il)

// every split of 1..8 patterns into full and byte patterns
TREX_COMPACT_BOTH(0, 1) TREX_COMPACT_BOTH(0, 2) TREX_COMPACT_BOTH(0, 3) TREX_COMPACT_BOTH(0, 4)
TREX_COMPACT_BOTH(0, 5) TREX_COMPACT_BOTH(0, 6) TREX_COMPACT_BOTH(0, 7) TREX_COMPACT_BOTH(0, 8)
TREX_COMPACT_BOTH(1, 0) TREX_COMPACT_BOTH(1, 1) TREX_COMPACT_BOTH(1, 2) TREX_COMPACT_BOTH(1, 3)
TREX_COMPACT_BOTH(1, 4) TREX_COMPACT_BOTH(1, 5) TREX_COMPACT_BOTH(1, 6) TREX_COMPACT_BOTH(1, 7)
TREX_COMPACT_BOTH(2, 0) TREX_COMPACT_BOTH(2, 1) TREX_COMPACT_BOTH(2, 2) TREX_COMPACT_BOTH(2, 3)
TREX_COMPACT_BOTH(2, 4) TREX_COMPACT_BOTH(2, 5) TREX_COMPACT_BOTH(2, 6)
TREX_COMPACT_BOTH(3, 0) TREX_COMPACT_BOTH(3, 1) TREX_COMPACT_BOTH(3, 2) TREX_COMPACT_BOTH(3, 3)
TREX_COMPACT_BOTH(3, 4) TREX_COMPACT_BOTH(3, 5)
TREX_COMPACT_BOTH(4, 0) TREX_COMPACT_BOTH(4, 1) TREX_COMPACT_BOTH(4, 2) TREX_COMPACT_BOTH(4, 3)
TREX_COMPACT_BOTH(4, 4)
TREX_COMPACT_BOTH(5, 0) TREX_COMPACT_BOTH(5, 1) TREX_COMPACT_BOTH(5, 2) TREX_COMPACT_BOTH(5, 3)
TREX_COMPACT_BOTH(6, 0) TREX_COMPACT_BOTH(6, 1) TREX_COMPACT_BOTH(6, 2)
TREX_COMPACT_BOTH(7, 0) TREX_COMPACT_BOTH(7, 1)
TREX_COMPACT_BOTH(8, 0)
