// Bit-parallel Boolean NFA-64 kernels for gfx950 (wave64).
//
// One lane walks one string. The state is a 64-bit mask (bit j = NFA state j
// active); one step is the union of the successor masks of the active states,
// the bit-parallel step of the reference's V4 kernel
// (cuda/nfa_tc_evolution.cu:683-745), which tests 64 rows per step and only
// handles two symbols. Here any symbol alphabet is handled with
// nibble-sliced successor tables in LDS:
//
//     N[c][p][v] = union of succ[c][4p + i] over the bits i of v   (uint64)
//
// so a step is 16 independent LDS lookups, one per nibble p of the state:
//     s' = OR_p N[c][p][(s >> 4p) & 15]
// The lookups depend on the state only through their index; all 16 are
// issued together.
//
// Byte -> class: `cls[byte]` names the table of that byte. Tables are built
// on the host for every symbol plus two extra classes, "reject" (all zero:
// the empty set, absorbing) and "skip" (the identity). In the window kernel,
// every byte outside one aligned 32-byte window [base, base+32) rejects and
// every byte inside it is its own class (byte & 31); a byte outside the
// window is only recorded (two bit operations per word) and forces the
// empty set at the end, exactly as if it had been stepped.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.h"

#define TREX_NFA_BLOCK 512

namespace {

using namespace trexdev;

template <bool WINDOW>
__device__ __forceinline__ void nfa_kernel(const uint8_t* __restrict__ bytes, const int64_t* __restrict__ begin,
                                           const int64_t* __restrict__ endp, int64_t n_strings,
                                           const uint64_t* __restrict__ tabg, int n_classes,
                                           const uint8_t* __restrict__ clsg, uint32_t win4, uint64_t start,
                                           uint64_t accept, uint64_t* __restrict__ out_mask,
                                           uint8_t* __restrict__ out_accept)
{
    extern __shared__ __attribute__((aligned(16))) uint64_t lds64[];
    uint64_t* tab = lds64;                                        // [n_classes][256]
    uint8_t* cls = reinterpret_cast<uint8_t*>(lds64 + (size_t)n_classes * 256);   // [256]
    {
        const uint4* s4 = reinterpret_cast<const uint4*>(tabg);
        uint4* d4 = reinterpret_cast<uint4*>(tab);
        for (int i = threadIdx.x; i < n_classes * 128; i += blockDim.x) d4[i] = s4[i];
        if (threadIdx.x < 256) cls[threadIdx.x] = clsg[threadIdx.x];
    }
    __syncthreads();

    const int64_t sid = (int64_t)blockIdx.x * TREX_NFA_BLOCK + threadIdx.x;
    if (sid >= n_strings) return;
    int64_t p = begin[sid];
    const int64_t end = endp[sid];
    uint64_t s = start;

    const int64_t head_end = min(end, p + (int64_t)((16u - ((uint32_t)(uintptr_t)(bytes + p) & 15u)) & 15u));
    for (; p < head_end; ++p) s = nfa_step(tab, cls[bytes[p]], s);

    WinCheck w;
    w.init();
    const int64_t nvec = (end - p) >> 4;
    const uint4* v = reinterpret_cast<const uint4*>(bytes + p);
    auto word = [&](uint32_t x) {
        if constexpr (WINDOW) {
            w.add(x);
            s = nfa_step(tab, x & 31u, s);
            s = nfa_step(tab, (x >> 8) & 31u, s);
            s = nfa_step(tab, (x >> 16) & 31u, s);
            s = nfa_step(tab, (x >> 24) & 31u, s);
        } else {
            s = nfa_step(tab, cls[x & 255u], s);
            s = nfa_step(tab, cls[(x >> 8) & 255u], s);
            s = nfa_step(tab, cls[(x >> 16) & 255u], s);
            s = nfa_step(tab, cls[x >> 24], s);
        }
    };
    int64_t i = 0;
    if (nvec >= 2) {
        uint4 a = v[0];
        for (i = 1; i < nvec; ++i) {
            const uint4 b = v[i];
            word(a.x); word(a.y); word(a.z); word(a.w);
            a = b;
        }
        word(a.x); word(a.y); word(a.z); word(a.w);
    } else if (nvec == 1) {
        const uint4 a = v[0];
        word(a.x); word(a.y); word(a.z); word(a.w);
        i = 1;
    }
    p += nvec << 4;
    for (; p < end; ++p) s = nfa_step(tab, cls[bytes[p]], s);
    if constexpr (WINDOW) {
        if (w.outside(win4)) s = 0;
    }
    out_mask[sid] = s;
    out_accept[sid] = (s & accept) != 0 ? 1 : 0;
}

#define TREX_NFA5_BLOCK 1024

template <bool SIX>
__device__ __forceinline__ void nfa_slice_kernel(const uint8_t* __restrict__ bytes, const int64_t* __restrict__ begin,
                                            const int64_t* __restrict__ endp, int64_t n_strings,
                                            const uint8_t* __restrict__ tabg, uint32_t win4, uint64_t start,
                                            uint64_t accept, uint64_t* __restrict__ out_mask,
                                            uint8_t* __restrict__ out_accept)
{
    constexpr uint32_t kBytes = SIX ? kTab6Bytes : kTab5Bytes;
    __shared__ __attribute__((aligned(16))) char tb[kBytes];
    stage_table(reinterpret_cast<uint8_t*>(tb), tabg, (int)kBytes);
    auto step = [&](uint32_t k, uint64_t x) { return SIX ? nfa6_step(tb, k, x) : nfa5_step(tb, k, x); };
    __syncthreads();

    const int64_t sid = (int64_t)blockIdx.x * TREX_NFA5_BLOCK + threadIdx.x;
    if (sid >= n_strings) return;
    int64_t p = begin[sid];
    const int64_t end = endp[sid];
    uint64_t s = start;
    WinCheck w;
    w.init();
    auto byte = [&](uint32_t b) {
        w.add(b * 0x01010101u);
        s = step(b & 31u, s);
    };
    const int64_t head_end = min(end, p + head_len(bytes, p));
    for (; p < head_end; ++p) byte(bytes[p]);
    const int64_t nvec = (end - p) >> 4;
    auto word = [&](uint32_t x) {
        w.add(x);
        s = step(x & 31u, s);
        s = step((x >> 8) & 31u, s);
        s = step((x >> 16) & 31u, s);
        s = step((x >> 24) & 31u, s);
    };
    stream_body<2>(reinterpret_cast<const uint4*>(bytes + p), nvec,
                   [&](uint4 v) { word(v.x); word(v.y); word(v.z); word(v.w); });
    p += nvec << 4;
    for (; p < end; ++p) byte(bytes[p]);
    if (w.outside(win4)) s = 0;
    out_mask[sid] = s;
    out_accept[sid] = (s & accept) != 0 ? 1 : 0;
}

}  // namespace

// tabg: 32 window classes of 5-bit slice tables, class k at byte k * 3336
// (13 rows of 32 uint64 entries, 8 bytes of padding); kTab5Bytes in total.
extern "C" __global__ void __launch_bounds__(TREX_NFA5_BLOCK)
nfa64_window5(const uint8_t* __restrict__ bytes, const int64_t* __restrict__ begin, const int64_t* __restrict__ end,
              int64_t n_strings, const uint8_t* __restrict__ tabg, uint32_t win4, uint64_t start, uint64_t accept,
              uint64_t* __restrict__ out_mask, uint8_t* __restrict__ out_accept)
{
    nfa_slice_kernel<false>(bytes, begin, end, n_strings, tabg, win4, start, accept, out_mask, out_accept);
}

// tabg: 32 window classes of 6/5-bit slice tables, class k at byte k * 4104
// (12 rows: 64, 64, 32, 32, 32, 32 uint64 entries per half; 8 bytes of
// padding); kTab6Bytes in total.
extern "C" __global__ void __launch_bounds__(TREX_NFA5_BLOCK)
nfa64_window6(const uint8_t* __restrict__ bytes, const int64_t* __restrict__ begin, const int64_t* __restrict__ end,
              int64_t n_strings, const uint8_t* __restrict__ tabg, uint32_t win4, uint64_t start, uint64_t accept,
              uint64_t* __restrict__ out_mask, uint8_t* __restrict__ out_accept)
{
    nfa_slice_kernel<true>(bytes, begin, end, n_strings, tabg, win4, start, accept, out_mask, out_accept);
}

// tabg: [n_classes][16][16] uint64 nibble tables; clsg: [256] class of every
// byte. Window kernel: classes 0..31 are the window bytes base + k (n_classes
// >= 32); the byte map is used for the unaligned head and tail bytes only.
// Dynamic LDS: n_classes * 2048 + 256 bytes.
extern "C" __global__ void __launch_bounds__(TREX_NFA_BLOCK)
nfa64_window(const uint8_t* __restrict__ bytes, const int64_t* __restrict__ begin, const int64_t* __restrict__ end,
             int64_t n_strings, const uint64_t* __restrict__ tabg, int n_classes, const uint8_t* __restrict__ clsg,
             uint32_t win4, uint64_t start, uint64_t accept, uint64_t* __restrict__ out_mask,
             uint8_t* __restrict__ out_accept)
{
    nfa_kernel<true>(bytes, begin, end, n_strings, tabg, n_classes, clsg, win4, start, accept, out_mask, out_accept);
}

extern "C" __global__ void __launch_bounds__(TREX_NFA_BLOCK)
nfa64_bytes(const uint8_t* __restrict__ bytes, const int64_t* __restrict__ begin, const int64_t* __restrict__ end,
            int64_t n_strings, const uint64_t* __restrict__ tabg, int n_classes, const uint8_t* __restrict__ clsg,
            uint32_t win4, uint64_t start, uint64_t accept, uint64_t* __restrict__ out_mask,
            uint8_t* __restrict__ out_accept)
{
    nfa_kernel<false>(bytes, begin, end, n_strings, tabg, n_classes, clsg, win4, start, accept, out_mask, out_accept);
}
