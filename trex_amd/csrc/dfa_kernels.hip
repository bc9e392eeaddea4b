// DFA byte-walk kernels for gfx950 (CDNA4, wave64).
//
// One lane walks one string through a byte transition table held in LDS and
// writes the final state and its accept bit. This is the work of the
// reference's monoid_batch_kernel / prefix_fused_kernel
// (cuda/monoid_batch.cu:58-95, cuda/prefix_compose.cu:217-258), reorganised
// for HBM streaming:
//
//  * string i is bytes[begin[i] .. end[i]); a CSR batch passes
//    begin = offsets, end = offsets + 1, and the staggered HBM layout (strings
//    grouped per wave, groups separated by a gap so that concurrent waves read
//    different HBM channels) passes its own begin/end arrays;
//  * the body of each string is read with 16-byte global loads, PF vectors
//    in flight per lane (two blocks of PF, double-buffered). The _il kernels
//    issue them non-temporal: a wave load there covers whole cache lines that
//    are never read again (measured: +12-14 % HBM read rate). Contiguous
//    strings keep ordinary loads: a lane's next loads hit the cache line its
//    last one brought in (non-temporal loads cost 2.5x there);
//  * the "window" kernels fold two bytes per table step ("K=4": one 32-bit
//    word = two pair steps) when every byte that does not lead straight to
//    the dead state lies in one aligned 32-byte window [base, base+32); the
//    pair index is sym0 | sym1 << 5 with sym = byte & 31. A byte outside the
//    window sends every state to the absorbing dead state, so the kernels
//    only record whether one was seen (two bit operations per word) and force
//    the dead state at the end. Two pair-table forms:
//      k2w  state walk, t2[pair * 16 + s] (16 KB, uint8)
//      k2f  function table, f2[pair] = 16 nibbles (8 KB, uint64): the LDS
//           read does not depend on the state, so it leaves the dependent
//           chain, which becomes s = (f >> 4s) & 15
//      k2fr the k2f table in four interleaved copies, lane l reading copy
//           l & 3: lanes of different copies never meet in an LDS bank
//  * dfa_k1 / dfa16_k1 are the general walk, one lookup per byte:
//    t1[byte * stride + s], for any table of up to 256 states.
//  * every kernel has an _il form for the interleaved layout: the strings of
//    a wave (64 consecutive strings) are stored as rows of 64 x 16 bytes,
//    vector k of lane j at group base + (64 k + j) * 16, so each load of the
//    wave reads one contiguous 1 KiB row. begin[i] is the address of string
//    i's vector 0 (16-byte aligned), end[i] - begin[i] its length; its last
//    len % 16 bytes open vector len / 16.
//
// Bytes before the first 16-byte boundary and after the last one go through
// the exact per-byte table t1 in every kernel.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.h"

#define TREX_BLOCK 256
// The string kernels run one wave per 64 strings; four waves per SIMD must stay
// resident, which caps them at 128 VGPRs.
#define TREX_STRING_KERNEL __global__ void __launch_bounds__(TREX_BLOCK) __attribute__((amdgpu_waves_per_eu(4)))

namespace {

using namespace trexdev;

constexpr int kIlGroup = 64;   // strings per group of the interleaved layout (one wave)

// Pair-step policies over the LDS pair table: stage() fills it from the
// global table, lane_table() is where this lane reads it, word() takes two
// pair steps.
struct StepK2W {
    static constexpr int kTableBytes = 1024 * 16;
    __device__ __forceinline__ static void stage(uint8_t* dst, const uint8_t* src) { stage_table(dst, src, kTableBytes); }
    __device__ __forceinline__ static const uint8_t* lane_table(const uint8_t* t) { return t; }
    __device__ __forceinline__ static uint32_t word(const uint8_t* __restrict__ t, uint32_t s, uint32_t x) {
        s = t[(pair_lo(x) << 4) | s];
        return t[(pair_hi(x) << 4) | s];
    }
};

struct StepK2F {
    static constexpr int kTableBytes = 1024 * 8;
    __device__ __forceinline__ static void stage(uint8_t* dst, const uint8_t* src) { stage_table(dst, src, kTableBytes); }
    __device__ __forceinline__ static const uint8_t* lane_table(const uint8_t* t) { return t; }
    __device__ __forceinline__ static uint32_t word(const uint8_t* __restrict__ t, uint32_t s, uint32_t x) {
        const uint64_t* f = reinterpret_cast<const uint64_t*>(t);
        const uint64_t f0 = f[pair_lo(x)];
        const uint64_t f1 = f[pair_hi(x)];
        s = (uint32_t)(f0 >> (s << 2)) & 15u;
        return (uint32_t)(f1 >> (s << 2)) & 15u;
    }
};

// k2f with the function table in kCopies interleaved copies: entry p of copy
// c sits at (p * kCopies + c) * 8 bytes and lane l reads copy l % kCopies, so
// the lanes of one LDS access that read different copies use disjoint banks.
struct StepK2FR {
    static constexpr int kCopies = 4;
    static constexpr int kTableBytes = 1024 * 8 * kCopies;
    __device__ __forceinline__ static void stage(uint8_t* dst, const uint8_t* src) {
        const uint64_t* s = reinterpret_cast<const uint64_t*>(src);
        uint64_t* d = reinterpret_cast<uint64_t*>(dst);
        for (int i = threadIdx.x; i < 1024 * kCopies; i += blockDim.x) d[i] = s[i / kCopies];
    }
    __device__ __forceinline__ static const uint8_t* lane_table(const uint8_t* t) {
        return t + (threadIdx.x % kCopies) * 8;
    }
    __device__ __forceinline__ static uint32_t word(const uint8_t* __restrict__ t, uint32_t s, uint32_t x) {
        const uint64_t* f = reinterpret_cast<const uint64_t*>(t);
        const uint64_t f0 = f[pair_lo(x) * kCopies];
        const uint64_t f1 = f[pair_hi(x) * kCopies];
        s = (uint32_t)(f0 >> (s << 2)) & 15u;
        return (uint32_t)(f1 >> (s << 2)) & 15u;
    }
};

template <class Step>
__device__ __forceinline__ uint32_t window_vec(const uint8_t* __restrict__ t, uint32_t s, uint4 v, WinCheck& w) {
    w.add(v.x); w.add(v.y); w.add(v.z); w.add(v.w);
    s = Step::word(t, s, v.x);
    s = Step::word(t, s, v.y);
    s = Step::word(t, s, v.z);
    s = Step::word(t, s, v.w);
    return s;
}

template <int STRIDE_LOG2>
__device__ __forceinline__ uint32_t k1_byte(const uint8_t* __restrict__ t1, uint32_t s, uint32_t b, uint32_t stride) {
    if constexpr (STRIDE_LOG2 >= 0) return t1[(b << STRIDE_LOG2) | s];
    else return t1[b * stride + s];
}

template <int STRIDE_LOG2>
__device__ __forceinline__ uint32_t k1_word(const uint8_t* __restrict__ t1, uint32_t s, uint32_t x, uint32_t stride) {
    s = k1_byte<STRIDE_LOG2>(t1, s, x & 255u, stride);
    s = k1_byte<STRIDE_LOG2>(t1, s, (x >> 8) & 255u, stride);
    s = k1_byte<STRIDE_LOG2>(t1, s, (x >> 16) & 255u, stride);
    s = k1_byte<STRIDE_LOG2>(t1, s, x >> 24, stride);
    return s;
}

template <int STRIDE_LOG2>
__device__ __forceinline__ uint32_t k1_vec(const uint8_t* __restrict__ t1, uint32_t s, uint4 v, uint32_t stride) {
    s = k1_word<STRIDE_LOG2>(t1, s, v.x, stride);
    s = k1_word<STRIDE_LOG2>(t1, s, v.y, stride);
    s = k1_word<STRIDE_LOG2>(t1, s, v.z, stride);
    s = k1_word<STRIDE_LOG2>(t1, s, v.w, stride);
    return s;
}

template <class Step, int PF, bool IL>
__device__ __forceinline__ void window_kernel(const uint8_t* __restrict__ bytes, const int64_t* __restrict__ begin,
                                              const int64_t* __restrict__ endp, int64_t n_strings,
                                              const uint8_t* __restrict__ tpg, const uint8_t* __restrict__ t1g,
                                              const uint8_t* __restrict__ accg, uint32_t win4, uint32_t start,
                                              uint32_t dead, uint8_t* __restrict__ out_state,
                                              uint8_t* __restrict__ out_accept)
{
    __shared__ __attribute__((aligned(16))) uint8_t tp[Step::kTableBytes];
    __shared__ __attribute__((aligned(16))) uint8_t t1[256 * 16];
    __shared__ __attribute__((aligned(16))) uint8_t acc[16];
    Step::stage(tp, tpg);
    stage_table(t1, t1g, 256 * 16);
    if (threadIdx.x < 16) acc[threadIdx.x] = accg[threadIdx.x];
    __syncthreads();

    const int64_t sid = (int64_t)blockIdx.x * TREX_BLOCK + threadIdx.x;
    if (sid >= n_strings) return;
    int64_t p = begin[sid];
    const int64_t end = endp[sid];
    uint32_t s = start;
    const uint8_t* tl = Step::lane_table(tp);
    WinCheck w; w.init();

    if constexpr (IL) {
        const int64_t len = end - p;
        const int64_t nvec = len >> 4;
        stream_body_nt<PF, kIlGroup>(reinterpret_cast<const uint4*>(bytes + p), nvec,
                                        [&](uint4 x) { s = window_vec<Step>(tl, s, x, w); });
        const uint8_t* tail = bytes + p + nvec * (16 * kIlGroup);
        for (int t = 0; t < (int)(len & 15); ++t) s = t1[((uint32_t)tail[t] << 4) | s];
    } else {
        const int64_t head_end = min(end, p + head_len(bytes, p));
        for (; p < head_end; ++p) s = t1[((uint32_t)bytes[p] << 4) | s];
        const int64_t nvec = (end - p) >> 4;
        stream_body<PF>(reinterpret_cast<const uint4*>(bytes + p), nvec,
                        [&](uint4 x) { s = window_vec<Step>(tl, s, x, w); });
        p += nvec << 4;
        for (; p < end; ++p) s = t1[((uint32_t)bytes[p] << 4) | s];
    }
    if (w.outside(win4)) s = dead;

    out_state[sid] = (uint8_t)s;
    out_accept[sid] = acc[s];
}

template <int STRIDE_LOG2, bool IL>
__device__ __forceinline__ void k1_kernel(const uint8_t* __restrict__ bytes, const int64_t* __restrict__ begin,
                                          const int64_t* __restrict__ endp, int64_t n_strings,
                                          const uint8_t* __restrict__ t1g, const uint8_t* __restrict__ accg,
                                          uint32_t stride, uint32_t start, uint8_t* __restrict__ out_state,
                                          uint8_t* __restrict__ out_accept)
{
    extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
    uint8_t* t1 = lds;
    uint8_t* acc = lds + 256 * stride;
    stage_table(t1, t1g, (int)(256 * stride));
    stage_table(acc, accg, (int)stride);
    __syncthreads();

    const int64_t sid = (int64_t)blockIdx.x * TREX_BLOCK + threadIdx.x;
    if (sid >= n_strings) return;
    int64_t p = begin[sid];
    const int64_t end = endp[sid];
    uint32_t s = start;

    if constexpr (IL) {
        const int64_t len = end - p;
        const int64_t nvec = len >> 4;
        stream_body_nt<8, kIlGroup>(reinterpret_cast<const uint4*>(bytes + p), nvec,
                                       [&](uint4 x) { s = k1_vec<STRIDE_LOG2>(t1, s, x, stride); });
        const uint8_t* tail = bytes + p + nvec * (16 * kIlGroup);
        for (int t = 0; t < (int)(len & 15); ++t) s = k1_byte<STRIDE_LOG2>(t1, s, tail[t], stride);
    } else {
        const int64_t head_end = min(end, p + head_len(bytes, p));
        for (; p < head_end; ++p) s = k1_byte<STRIDE_LOG2>(t1, s, bytes[p], stride);
        const int64_t nvec = (end - p) >> 4;
        stream_body<8>(reinterpret_cast<const uint4*>(bytes + p), nvec,
                       [&](uint4 x) { s = k1_vec<STRIDE_LOG2>(t1, s, x, stride); });
        p += nvec << 4;
        for (; p < end; ++p) s = k1_byte<STRIDE_LOG2>(t1, s, bytes[p], stride);
    }

    out_state[sid] = (uint8_t)s;
    out_accept[sid] = acc[s];
}

// k2f over contiguous strings with the read order rotated per wave.
//
// Strings laid out back to back at equal lengths (a packed batch of 4 KiB
// strings) are read by every resident wave at the same offset modulo 4 KiB at
// once, which crowds a few HBM channels (measured: 4.1 TB/s for the bare read
// pattern, 6.0 TB/s once each wave starts 256 B x (wave % 16) further into
// its strings). Lane i therefore reads its body vectors r .. nvec-1 first,
// r = (16 x (wave % 16)) mod nvec, then 0 .. r-1:
//  * phase A composes the function of vectors r .. nvec-1 from `init` (the
//    identity on the states that can occur, other slots started at a state
//    that can) with the 16-byte rows of t2, until the function is constant
//    in every lane of the wave; from then on the wave walks the constant's
//    state with the nibble functions of f2, as the unrotated kernel does;
//  * phase B walks vectors 0 .. r-1 from the state after the head bytes;
//  * the body's result is phase A's walked state, or its function applied to
//    phase B's state (the dead state stays dead); then the tail bytes.
// Same results as the unrotated kernel. The host picks it for automata that
// synchronise (the function turns constant within a few bytes); one that
// never does would compose a whole phase A.
template <int PF>
__device__ __forceinline__ void rot_kernel(const uint8_t* __restrict__ bytes, const int64_t* __restrict__ begin,
                                           const int64_t* __restrict__ endp, int64_t n_strings,
                                           const uint8_t* __restrict__ f2g, const uint8_t* __restrict__ t2g,
                                           const uint8_t* __restrict__ t1g, const uint8_t* __restrict__ accg,
                                           uint32_t win4, uint32_t start, uint32_t dead, Fn16 init,
                                           uint8_t* __restrict__ out_state, uint8_t* __restrict__ out_accept)
{
    __shared__ __attribute__((aligned(16))) uint8_t f2[1024 * 8];
    __shared__ __attribute__((aligned(16))) uint8_t t2[1024 * 16];
    __shared__ __attribute__((aligned(16))) uint8_t t1[256 * 16];
    __shared__ __attribute__((aligned(16))) uint8_t acc[16];
    stage_table(f2, f2g, 1024 * 8);
    stage_table(t2, t2g, 1024 * 16);
    stage_table(t1, t1g, 256 * 16);
    if (threadIdx.x < 16) acc[threadIdx.x] = accg[threadIdx.x];
    __syncthreads();

    const int64_t sid = (int64_t)blockIdx.x * TREX_BLOCK + threadIdx.x;
    if (sid >= n_strings) return;
    int64_t p = begin[sid];
    const int64_t end = endp[sid];
    uint32_t s = start;
    WinCheck w; w.init();

    const int64_t head_end = min(end, p + head_len(bytes, p));
    for (; p < head_end; ++p) s = t1[((uint32_t)bytes[p] << 4) | s];
    const int64_t nvec = (end - p) >> 4;
    const uint4* v = reinterpret_cast<const uint4*>(bytes + p);
    const int64_t r = nvec > 0 ? (int64_t)(16u * ((uint32_t)(sid >> 6) & 15u)) % nvec : 0;

    // phase A: vectors r .. nvec-1
    bool walking = false;
    uint32_t sa = 0;
    Fn16 f = init;
    auto mixed = [&](const uint4& x) {   // some lane of the wave still composes
        const uint32_t words[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            w.add(words[k]);
            if (walking) {
                sa = StepK2F::word(f2, sa, words[k]);
            } else {
                f = fn_then(f, load_row(t2, pair_lo(words[k])));
                f = fn_then(f, load_row(t2, pair_hi(words[k])));
            }
        }
        if (!walking && fn_is_const(f)) {
            walking = true;
            sa = f.d[0] & 0xffu;
        }
    };
    // the wave-wide mode test once per block of PF vectors: a walking wave
    // runs each block straight-line, as the unrotated kernel does (with the
    // compiler's schedule of the table reads: 1-2 % faster here than the
    // one-vector-ahead order the look-back scan uses)
    stream_blocks<PF>(v + r, nvec - r, [&](const uint4 (&a)[PF]) {
        if (__ballot(!walking) == 0ull) {
#pragma unroll
            for (int k = 0; k < PF; ++k) sa = window_vec<StepK2F>(f2, sa, a[k], w);
        } else {
#pragma unroll
            for (int k = 0; k < PF; ++k) mixed(a[k]);
        }
    }, [&](const uint4& x) {
        if (__ballot(!walking) == 0ull) sa = window_vec<StepK2F>(f2, sa, x, w);
        else mixed(x);
    });
    // phase B: vectors 0 .. r-1 from the head's state
    stream_body<PF>(v, r, [&](uint4 x) { s = window_vec<StepK2F>(f2, s, x, w); });
    // the dead state is absorbing; `init` may leave its slot out (when no
    // window byte leads to it, it is reached only through a byte the window
    // check or the head's exact walk already sends to it)
    s = s == dead ? dead : (walking ? sa : fn_apply(f, s));
    p += nvec << 4;
    for (; p < end; ++p) s = t1[((uint32_t)bytes[p] << 4) | s];
    if (w.outside(win4)) s = dead;

    out_state[sid] = (uint8_t)s;
    out_accept[sid] = acc[s];
}

}  // namespace

// ---------------------------------------------------------------------------
// Window kernels: 16-state table, two bytes per step.
//   tpg: pair table (k2w: 16 KB uint8, k2f / k2fr: 8 KB uint64)
//   t1g [256 * 16] byte table, accg [16]
//   win4 = window base byte replicated in all four bytes (base % 32 == 0)
// ---------------------------------------------------------------------------
#define TREX_WINDOW_KERNEL(NAME, STEP, PF, IL)                                                                \
    extern "C" __global__ void __launch_bounds__(TREX_BLOCK)                                                  \
    NAME(const uint8_t* __restrict__ bytes, const int64_t* __restrict__ begin, const int64_t* __restrict__ end, \
         int64_t n_strings, const uint8_t* __restrict__ tpg, const uint8_t* __restrict__ t1g,                \
         const uint8_t* __restrict__ accg, uint32_t win4, uint32_t start, uint32_t dead,                     \
         uint8_t* __restrict__ out_state, uint8_t* __restrict__ out_accept) {                                \
        window_kernel<STEP, PF, IL>(bytes, begin, end, n_strings, tpg, t1g, accg, win4, start, dead,         \
                                    out_state, out_accept);                                                   \
    }

#define TREX_WINDOW_KERNELS(BASE, STEP)                        \
    TREX_WINDOW_KERNEL(BASE##_pf8, STEP, 8, false)             \
    TREX_WINDOW_KERNEL(BASE##_pf16, STEP, 16, false)           \
    TREX_WINDOW_KERNEL(BASE##_il_pf8, STEP, 8, true)           \
    TREX_WINDOW_KERNEL(BASE##_il_pf16, STEP, 16, true)

TREX_WINDOW_KERNELS(dfa16_k2w, StepK2W)
TREX_WINDOW_KERNELS(dfa16_k2f, StepK2F)
TREX_WINDOW_KERNELS(dfa16_k2fr, StepK2FR)

// k2f with the per-wave rotated read order (contiguous strings only).
//   f2g: nibble pair functions (8 KB), t2g: pair rows [1024][16] (16 KB);
//   init0..3: the 16 slots of the composition's starting function.
extern "C" __global__ void __launch_bounds__(TREX_BLOCK)
dfa16_k2f_rot_pf8(const uint8_t* __restrict__ bytes, const int64_t* __restrict__ begin,
                  const int64_t* __restrict__ end, int64_t n_strings, const uint8_t* __restrict__ f2g,
                  const uint8_t* __restrict__ t2g, const uint8_t* __restrict__ t1g,
                  const uint8_t* __restrict__ accg, uint32_t win4, uint32_t start, uint32_t dead, uint32_t init0,
                  uint32_t init1, uint32_t init2, uint32_t init3, uint8_t* __restrict__ out_state,
                  uint8_t* __restrict__ out_accept) {
    rot_kernel<8>(bytes, begin, end, n_strings, f2g, t2g, t1g, accg, win4, start, dead,
                  Fn16{{init0, init1, init2, init3}}, out_state, out_accept);
}

// ---------------------------------------------------------------------------
// General byte walk.
//   t1g [256 * stride] with t1[byte * stride + s]; accg [stride]
//   Dynamic LDS: 256 * stride + stride bytes (stride a multiple of 16).
// ---------------------------------------------------------------------------
#define TREX_K1_KERNEL(NAME, STRIDE_LOG2, IL)                                                                 \
    extern "C" TREX_STRING_KERNEL                                                                             \
    NAME(const uint8_t* __restrict__ bytes, const int64_t* __restrict__ begin, const int64_t* __restrict__ end, \
         int64_t n_strings, const uint8_t* __restrict__ t1g, const uint8_t* __restrict__ accg, uint32_t stride, \
         uint32_t start, uint8_t* __restrict__ out_state, uint8_t* __restrict__ out_accept) {                 \
        k1_kernel<STRIDE_LOG2, IL>(bytes, begin, end, n_strings, t1g, accg, STRIDE_LOG2 >= 0 ? 16u : stride,  \
                                   start, out_state, out_accept);                                             \
    }

TREX_K1_KERNEL(dfa16_k1, 4, false)
TREX_K1_KERNEL(dfa16_k1_il, 4, true)
TREX_K1_KERNEL(dfa_k1, -1, false)
TREX_K1_KERNEL(dfa_k1_il, -1, true)

// ---------------------------------------------------------------------------
// Plain 16-byte copy, the reference point for HBM traffic counters.
// n16 = number of 16-byte vectors.
// ---------------------------------------------------------------------------
extern "C" __global__ void __launch_bounds__(TREX_BLOCK)
copy_u4(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n16)
{
    const int64_t stride = (int64_t)gridDim.x * TREX_BLOCK;
    for (int64_t i = (int64_t)blockIdx.x * TREX_BLOCK + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}
