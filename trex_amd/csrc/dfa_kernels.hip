// DFA byte-walk kernels for gfx950 (CDNA4, wave64).
//
// One lane walks one string of a CSR batch (bytes + int64 offsets) through a
// byte transition table held in LDS and writes the final state and its
// accept bit. This is the work of the reference's monoid_batch_kernel /
// prefix_fused_kernel (cuda/monoid_batch.cu:58-95, cuda/prefix_compose.cu:217-258),
// reorganised for HBM streaming:
//
//  * the body of each string is read with 16-byte global loads, four vectors
//    in flight per lane while the previous four are consumed;
//  * dfa16_k2w folds two bytes per LDS lookup ("K=4": one 32-bit word = two
//    pair lookups) when every byte that does not lead straight to the dead
//    state lies in one aligned 32-byte window [base, base+32). Its table is
//    t2[(sym0 | sym1 << 5) * 16 + s] = delta(delta(s, base+sym0), base+sym1).
//    A byte outside the window sends every state to the absorbing dead
//    state, so the kernel only records whether one was seen (two bit ops per
//    word) and forces the dead state at the end;
//  * dfa_k1 is the general walk over any table of up to 256 states, one
//    lookup per byte: t1[byte * stride + s].
//
// Bytes before the first 16-byte boundary and after the last one go through
// the exact per-byte table t1 in both kernels.

#include <hip/hip_runtime.h>
#include <stdint.h>

#define TREX_BLOCK 256

namespace {

__device__ __forceinline__ uint32_t k2w_word(const uint8_t* __restrict__ t2, uint32_t s, uint32_t x) {
    uint32_t i0 = (x & 31u) | (((x >> 8) & 31u) << 5);
    uint32_t i1 = ((x >> 16) & 31u) | (((x >> 24) & 31u) << 5);
    s = t2[(i0 << 4) | s];
    s = t2[(i1 << 4) | s];
    return s;
}

struct WinCheck {
    uint32_t any_hi;   // OR of the top three bits of every body byte
    uint32_t all;      // AND of every body word
    __device__ __forceinline__ void init() { any_hi = 0u; all = 0xffffffffu; }
    __device__ __forceinline__ void add(uint32_t x) { any_hi |= x & 0xe0e0e0e0u; all &= x; }
    // true iff some body byte had top bits != the window's
    __device__ __forceinline__ bool outside(uint32_t win4) const {
        return ((any_hi & ~win4) | (win4 & ~all)) != 0u;
    }
};

__device__ __forceinline__ uint32_t k2w_vec(const uint8_t* __restrict__ t2, uint32_t s, uint4 v, WinCheck& w) {
    w.add(v.x); w.add(v.y); w.add(v.z); w.add(v.w);
    s = k2w_word(t2, s, v.x);
    s = k2w_word(t2, s, v.y);
    s = k2w_word(t2, s, v.z);
    s = k2w_word(t2, s, v.w);
    return s;
}

template <int STRIDE_LOG2>
__device__ __forceinline__ uint32_t k1_byte(const uint8_t* __restrict__ t1, uint32_t s, uint32_t b, uint32_t stride) {
    if (STRIDE_LOG2 >= 0) return t1[(b << STRIDE_LOG2) | s];
    return t1[b * stride + s];
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

// Copy `n` bytes (multiple of 16) from global to LDS with the whole block.
__device__ __forceinline__ void stage_table(uint8_t* __restrict__ dst, const uint8_t* __restrict__ src, int n) {
    const uint4* s4 = reinterpret_cast<const uint4*>(src);
    uint4* d4 = reinterpret_cast<uint4*>(dst);
    for (int i = threadIdx.x; i < (n >> 4); i += blockDim.x) d4[i] = s4[i];
}

}  // namespace

// ---------------------------------------------------------------------------
// dfa16_k2w: 16-state table, window fast path, two bytes per lookup.
//   t2g  [1024 * 16] pair table, t1g [256 * 16] byte table, accg [16]
//   win4 = window base byte replicated in all four bytes (base % 32 == 0)
// ---------------------------------------------------------------------------
extern "C" __global__ void __launch_bounds__(TREX_BLOCK)
dfa16_k2w(const uint8_t* __restrict__ bytes, const int64_t* __restrict__ offsets, int64_t n_strings,
          const uint8_t* __restrict__ t2g, const uint8_t* __restrict__ t1g, const uint8_t* __restrict__ accg,
          uint32_t win4, uint32_t start, uint32_t dead,
          uint8_t* __restrict__ out_state, uint8_t* __restrict__ out_accept)
{
    __shared__ __attribute__((aligned(16))) uint8_t t2[1024 * 16];
    __shared__ __attribute__((aligned(16))) uint8_t t1[256 * 16];
    __shared__ __attribute__((aligned(16))) uint8_t acc[16];
    stage_table(t2, t2g, 1024 * 16);
    stage_table(t1, t1g, 256 * 16);
    if (threadIdx.x < 16) acc[threadIdx.x] = accg[threadIdx.x];
    __syncthreads();

    const int64_t sid = (int64_t)blockIdx.x * TREX_BLOCK + threadIdx.x;
    if (sid >= n_strings) return;
    int64_t p = offsets[sid];
    const int64_t end = offsets[sid + 1];
    uint32_t s = start;

    const int64_t head_end = min(end, p + (int64_t)((16u - ((uint32_t)(uintptr_t)(bytes + p) & 15u)) & 15u));
    for (; p < head_end; ++p) s = t1[((uint32_t)bytes[p] << 4) | s];

    WinCheck w; w.init();
    const uint4* v = reinterpret_cast<const uint4*>(bytes + p);
    const int64_t nvec = (end - p) >> 4;
    int64_t i = 0;
    if (nvec >= 4) {
        uint4 a0 = v[0], a1 = v[1], a2 = v[2], a3 = v[3];
        for (i = 4; i + 4 <= nvec; i += 4) {
            const uint4 b0 = v[i], b1 = v[i + 1], b2 = v[i + 2], b3 = v[i + 3];
            s = k2w_vec(t2, s, a0, w); s = k2w_vec(t2, s, a1, w);
            s = k2w_vec(t2, s, a2, w); s = k2w_vec(t2, s, a3, w);
            a0 = b0; a1 = b1; a2 = b2; a3 = b3;
        }
        s = k2w_vec(t2, s, a0, w); s = k2w_vec(t2, s, a1, w);
        s = k2w_vec(t2, s, a2, w); s = k2w_vec(t2, s, a3, w);
    }
    for (; i < nvec; ++i) s = k2w_vec(t2, s, v[i], w);
    p += nvec << 4;
    for (; p < end; ++p) s = t1[((uint32_t)bytes[p] << 4) | s];
    if (w.outside(win4)) s = dead;

    out_state[sid] = (uint8_t)s;
    out_accept[sid] = acc[s];
}

// ---------------------------------------------------------------------------
// dfa_k1: general byte walk, up to 256 states.
//   t1g [256 * stride] with t1[byte * stride + s]; accg [stride]
//   Dynamic LDS: 256 * stride + stride bytes (stride a multiple of 16).
// ---------------------------------------------------------------------------
template <int STRIDE_LOG2>
__device__ __forceinline__ void dfa_k1_body(const uint8_t* __restrict__ bytes, const int64_t* __restrict__ offsets,
                                            int64_t n_strings, const uint8_t* __restrict__ t1g,
                                            const uint8_t* __restrict__ accg, uint32_t stride, uint32_t start,
                                            uint8_t* __restrict__ out_state, uint8_t* __restrict__ out_accept)
{
    extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
    uint8_t* t1 = lds;
    uint8_t* acc = lds + 256 * stride;
    stage_table(t1, t1g, (int)(256 * stride));
    stage_table(acc, accg, (int)stride);
    __syncthreads();

    const int64_t sid = (int64_t)blockIdx.x * TREX_BLOCK + threadIdx.x;
    if (sid >= n_strings) return;
    int64_t p = offsets[sid];
    const int64_t end = offsets[sid + 1];
    uint32_t s = start;

    const int64_t head_end = min(end, p + (int64_t)((16u - ((uint32_t)(uintptr_t)(bytes + p) & 15u)) & 15u));
    for (; p < head_end; ++p) s = k1_byte<STRIDE_LOG2>(t1, s, bytes[p], stride);

    const uint4* v = reinterpret_cast<const uint4*>(bytes + p);
    const int64_t nvec = (end - p) >> 4;
    int64_t i = 0;
    if (nvec >= 4) {
        uint4 a0 = v[0], a1 = v[1], a2 = v[2], a3 = v[3];
        for (i = 4; i + 4 <= nvec; i += 4) {
            const uint4 b0 = v[i], b1 = v[i + 1], b2 = v[i + 2], b3 = v[i + 3];
            s = k1_vec<STRIDE_LOG2>(t1, s, a0, stride); s = k1_vec<STRIDE_LOG2>(t1, s, a1, stride);
            s = k1_vec<STRIDE_LOG2>(t1, s, a2, stride); s = k1_vec<STRIDE_LOG2>(t1, s, a3, stride);
            a0 = b0; a1 = b1; a2 = b2; a3 = b3;
        }
        s = k1_vec<STRIDE_LOG2>(t1, s, a0, stride); s = k1_vec<STRIDE_LOG2>(t1, s, a1, stride);
        s = k1_vec<STRIDE_LOG2>(t1, s, a2, stride); s = k1_vec<STRIDE_LOG2>(t1, s, a3, stride);
    }
    for (; i < nvec; ++i) s = k1_vec<STRIDE_LOG2>(t1, s, v[i], stride);
    p += nvec << 4;
    for (; p < end; ++p) s = k1_byte<STRIDE_LOG2>(t1, s, bytes[p], stride);

    out_state[sid] = (uint8_t)s;
    out_accept[sid] = acc[s];
}

extern "C" __global__ void __launch_bounds__(TREX_BLOCK)
dfa16_k1(const uint8_t* __restrict__ bytes, const int64_t* __restrict__ offsets, int64_t n_strings,
         const uint8_t* __restrict__ t1g, const uint8_t* __restrict__ accg, uint32_t stride, uint32_t start,
         uint8_t* __restrict__ out_state, uint8_t* __restrict__ out_accept)
{
    dfa_k1_body<4>(bytes, offsets, n_strings, t1g, accg, 16u, start, out_state, out_accept);
}

extern "C" __global__ void __launch_bounds__(TREX_BLOCK)
dfa_k1(const uint8_t* __restrict__ bytes, const int64_t* __restrict__ offsets, int64_t n_strings,
       const uint8_t* __restrict__ t1g, const uint8_t* __restrict__ accg, uint32_t stride, uint32_t start,
       uint8_t* __restrict__ out_state, uint8_t* __restrict__ out_accept)
{
    dfa_k1_body<-1>(bytes, offsets, n_strings, t1g, accg, stride, start, out_state, out_accept);
}
