// Chunked DFA-16 matching with a decoupled look-back across the chunks of a
// string (gfx950, wave64).
//
// Long strings are cut into chunks so that one string's bytes are walked by
// many lanes at once. Chunk c covers bytes[begin[c] .. end[c]) of string
// sid[c]; pred[c] is the chunk just before it in the same string (-1 for a
// string's first chunk) and flags[c] bit 1 marks a string's last chunk.
//
//  Phase 1 (no waiting): a first chunk walks its bytes from the start state.
//    Any other chunk composes the transition function of its bytes,
//    F = T[b_n] o ... o T[b_1], held as 16 state bytes in four registers
//    (byte s = image of state s): each step gathers G[F[s]] for all s with
//    three byte permutes per register (v_perm_b32), G being the 16-byte row
//    of the pair table (two bytes per step) or of the byte table. Once F maps
//    every state to one state it stays constant-valued, so the lane switches
//    to walking that single state.
//    The chunk then publishes AGGREGATE (its function), or PREFIX (its final
//    state) for a first chunk.
//  Phase 2 (look-back, cuda/monoid_scan.cu:128-222 restated over functions):
//    walk back over predecessors, composing their aggregates, until one with
//    PREFIX is found; the chunk's prefix state is F(acc(p)). Waiting is only
//    for predecessors' phase-1 results, which never wait on anything.
//
// Forward progress: chunks are claimed from an atomic ticket in index order,
// and a chunk's predecessor always has a smaller index, so every chunk waited
// on has been claimed by a running wave (the reference's blockIdx.x order,
// cuda/monoid_scan.cu:141, gives no such guarantee). Spins are bounded by
// `max_spins`; on expiry the kernel sets `error` and moves on, so every wave
// always exits.
//
// status[c]: 0 = not ready, 1 = aggregate in agg[c], 2 | state << 8 = prefix.
// status, the ticket and the error word are cleared by fill_u32 before each
// launch (a kernel, not a memset, so the sequence can be graph-captured).

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.h"

#define TREX_BLOCK 256

namespace {

using namespace trexdev;

constexpr uint32_t kStatusAggregate = 1u;
constexpr uint32_t kStatusPrefix = 2u;

// The per-lane scanner: either walking one state or composing a function.
template <bool WINDOW>
struct Scanner {
    const uint8_t* t2;   // pair table [1024][16] (WINDOW)
    const uint8_t* t1;   // byte table [256][16]
    bool walking;
    uint32_t s;
    Fn16 f;
    WinCheck w;

    __device__ __forceinline__ void byte(uint32_t b) {
        if (walking) s = t1[(b << 4) | s];
        else f = fn_then(f, load_row(t1, b));
    }
    __device__ __forceinline__ void word(uint32_t x) {
        if constexpr (WINDOW) {
            w.add(x);
            if (walking) {
                s = t2[(pair_lo(x) << 4) | s];
                s = t2[(pair_hi(x) << 4) | s];
            } else {
                f = fn_then(f, load_row(t2, pair_lo(x)));
                f = fn_then(f, load_row(t2, pair_hi(x)));
            }
        } else {
            byte(x & 255u); byte((x >> 8) & 255u); byte((x >> 16) & 255u); byte(x >> 24);
        }
    }
    // After a block: a constant function becomes a single walked state.
    __device__ __forceinline__ void settle() {
        if (!walking && fn_is_const(f)) {
            walking = true;
            s = f.d[0] & 0xffu;
        }
    }
};

template <bool WINDOW>
__device__ void chunk_scan(const uint8_t* __restrict__ bytes, const int64_t* __restrict__ cbegin,
                           const int64_t* __restrict__ cend, const int32_t* __restrict__ csid,
                           const int32_t* __restrict__ cpred, const uint32_t* __restrict__ cflags, int64_t n_chunks,
                           const uint8_t* __restrict__ t2g, const uint8_t* __restrict__ t1g,
                           const uint8_t* __restrict__ accg, uint32_t win4, uint32_t start, uint32_t dead,
                           uint32_t* __restrict__ status, uint64_t* __restrict__ agg, uint32_t* __restrict__ ticket,
                           uint32_t* __restrict__ error, uint32_t max_spins, uint8_t* __restrict__ out_state,
                           uint8_t* __restrict__ out_accept)
{
    __shared__ __attribute__((aligned(16))) uint8_t t2[WINDOW ? 1024 * 16 : 16];
    __shared__ __attribute__((aligned(16))) uint8_t t1[256 * 16];
    __shared__ __attribute__((aligned(16))) uint8_t acc[16];
    {
        const uint4* s4 = reinterpret_cast<const uint4*>(t1g);
        for (int i = threadIdx.x; i < 256; i += blockDim.x) reinterpret_cast<uint4*>(t1)[i] = s4[i];
        if constexpr (WINDOW) {
            const uint4* p4 = reinterpret_cast<const uint4*>(t2g);
            for (int i = threadIdx.x; i < 1024; i += blockDim.x) reinterpret_cast<uint4*>(t2)[i] = p4[i];
        }
        if (threadIdx.x < 16) acc[threadIdx.x] = accg[threadIdx.x];
    }
    __syncthreads();

    const uint32_t lane = threadIdx.x & 63u;
    while (true) {
        // one ticket per wave: 64 consecutive chunks
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(ticket, 64u);
        base = __shfl(base, 0, 64);
        if ((int64_t)base >= n_chunks) break;
        const int64_t c = (int64_t)base + lane;
        const bool active = c < n_chunks;

        Scanner<WINDOW> sc;
        sc.t2 = t2;
        sc.t1 = t1;
        sc.w.init();
        int32_t pred = -1;
        uint32_t flags = 0;
        int64_t p = 0, end = 0;
        if (active) {
            pred = cpred[c];
            flags = cflags[c];
            p = cbegin[c];
            end = cend[c];
        }
        sc.walking = pred < 0;
        sc.s = start;
        sc.f = fn_identity();

        // ---- phase 1: walk or compose this chunk's bytes ----
        const int64_t head_end = min(end, p + (int64_t)((16u - ((uint32_t)(uintptr_t)(bytes + p) & 15u)) & 15u));
        for (; p < head_end; ++p) sc.byte(bytes[p]);
        const int64_t nvec = (end - p) >> 4;
        const uint4* v = reinterpret_cast<const uint4*>(bytes + p);
        int64_t i = 0;
        for (; i + 4 <= nvec; i += 4) {
            const uint4 a0 = v[i], a1 = v[i + 1], a2 = v[i + 2], a3 = v[i + 3];
            sc.word(a0.x); sc.word(a0.y); sc.word(a0.z); sc.word(a0.w);
            sc.word(a1.x); sc.word(a1.y); sc.word(a1.z); sc.word(a1.w);
            sc.word(a2.x); sc.word(a2.y); sc.word(a2.z); sc.word(a2.w);
            sc.word(a3.x); sc.word(a3.y); sc.word(a3.z); sc.word(a3.w);
            sc.settle();
        }
        for (; i < nvec; ++i) {
            const uint4 a = v[i];
            sc.word(a.x); sc.word(a.y); sc.word(a.z); sc.word(a.w);
        }
        p += nvec << 4;
        for (; p < end; ++p) sc.byte(bytes[p]);
        if constexpr (WINDOW) {
            if (sc.w.outside(win4)) { sc.walking = true; sc.s = dead; }
        }

        // ---- publish ----
        uint64_t my_fn = 0;
        if (active) {
            if (pred < 0) {
                store_release(&status[c], kStatusPrefix | (sc.s << 8));
            } else {
                my_fn = sc.walking ? fn_pack(fn_const(sc.s)) : fn_pack(sc.f);
                agg[c] = my_fn;
                store_release(&status[c], kStatusAggregate);
            }
        }

        // ---- phase 2: look back to the nearest published prefix ----
        uint32_t state = sc.s;
        if (active && pred >= 0) {
            uint64_t acc_fn = 0xfedcba9876543210ull;   // identity
            int32_t j = pred;
            uint32_t spins = 0;
            bool done = false;
            while (!done) {
                const uint32_t st = load_relaxed(&status[j]);
                if (st == 0u) {
                    if (++spins > max_spins) {
                        atomicOr(error, 1u);
                        state = dead;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                    continue;
                }
                fence_acquire();   // agg[j] was written before status[j] was released
                if ((st & 0xffu) == kStatusPrefix) {
                    state = nib(my_fn, nib(acc_fn, (st >> 8) & 15u));
                    done = true;
                } else {
                    acc_fn = nib_then(agg[j], acc_fn);   // earlier chunk applies first
                    j = cpred[j];
                    if (j < 0) {   // cannot happen: a string's first chunk publishes a prefix
                        atomicOr(error, 2u);
                        state = dead;
                        break;
                    }
                }
            }
            store_release(&status[c], kStatusPrefix | (state << 8));
        }
        if (active && (flags & 2u)) {
            const int32_t sid = csid[c];
            out_state[sid] = (uint8_t)state;
            out_accept[sid] = acc[state];
        }
    }
}

}  // namespace

extern "C" __global__ void __launch_bounds__(TREX_BLOCK)
dfa16_chunks_window(const uint8_t* __restrict__ bytes, const int64_t* __restrict__ cbegin,
                    const int64_t* __restrict__ cend, const int32_t* __restrict__ csid,
                    const int32_t* __restrict__ cpred, const uint32_t* __restrict__ cflags, int64_t n_chunks,
                    const uint8_t* __restrict__ t2g, const uint8_t* __restrict__ t1g,
                    const uint8_t* __restrict__ accg, uint32_t win4, uint32_t start, uint32_t dead,
                    uint32_t* __restrict__ status, uint64_t* __restrict__ agg, uint32_t* __restrict__ ticket,
                    uint32_t* __restrict__ error, uint32_t max_spins, uint8_t* __restrict__ out_state,
                    uint8_t* __restrict__ out_accept)
{
    chunk_scan<true>(bytes, cbegin, cend, csid, cpred, cflags, n_chunks, t2g, t1g, accg, win4, start, dead, status,
                     agg, ticket, error, max_spins, out_state, out_accept);
}

extern "C" __global__ void __launch_bounds__(TREX_BLOCK)
dfa16_chunks_bytes(const uint8_t* __restrict__ bytes, const int64_t* __restrict__ cbegin,
                   const int64_t* __restrict__ cend, const int32_t* __restrict__ csid,
                   const int32_t* __restrict__ cpred, const uint32_t* __restrict__ cflags, int64_t n_chunks,
                   const uint8_t* __restrict__ t2g, const uint8_t* __restrict__ t1g,
                   const uint8_t* __restrict__ accg, uint32_t win4, uint32_t start, uint32_t dead,
                   uint32_t* __restrict__ status, uint64_t* __restrict__ agg, uint32_t* __restrict__ ticket,
                   uint32_t* __restrict__ error, uint32_t max_spins, uint8_t* __restrict__ out_state,
                   uint8_t* __restrict__ out_accept)
{
    chunk_scan<false>(bytes, cbegin, cend, csid, cpred, cflags, n_chunks, t2g, t1g, accg, win4, start, dead, status,
                      agg, ticket, error, max_spins, out_state, out_accept);
}

// Sets dst[0 .. n) = value (status words, tickets, error flags).
extern "C" __global__ void __launch_bounds__(TREX_BLOCK)
fill_u32(uint32_t* __restrict__ dst, int64_t n, uint32_t value)
{
    const int64_t stride = (int64_t)gridDim.x * TREX_BLOCK;
    for (int64_t i = (int64_t)blockIdx.x * TREX_BLOCK + threadIdx.x; i < n; i += stride) dst[i] = value;
}
