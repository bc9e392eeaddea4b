// Device helpers shared by the gfx950 kernels (DFA walks, look-back scan,
// Boolean NFA, mixed work queue). Header-only, internal linkage.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace trexdev {

// ---------------------------------------------------------------------------
// Byte streaming
// ---------------------------------------------------------------------------

// Number of bytes from `p` to the next 16-byte boundary of the address.
__device__ __forceinline__ int64_t head_len(const uint8_t* bytes, int64_t p) {
    return (int64_t)((16u - ((uint32_t)(uintptr_t)(bytes + p) & 15u)) & 15u);
}

// Streams the 16-byte vectors v[0], v[VSTRIDE], ..., v[(nvec - 1) * VSTRIDE]
// through `consume` in order with two blocks of PF vectors: the loads of block
// i + 1 are issued before block i is consumed, so PF loads are in flight while
// the table walk runs. VSTRIDE = 1: a string stored contiguously; VSTRIDE = 64:
// the interleaved layout, where one wave's 64 strings share each 1 KiB row.
template <int PF, int VSTRIDE = 1, class F>
__device__ __forceinline__ void stream_body(const uint4* __restrict__ v, int64_t nvec, F&& consume) {
    int64_t i = 0;
    if (nvec >= PF) {
        uint4 a[PF];
#pragma unroll
        for (int k = 0; k < PF; ++k) a[k] = v[(int64_t)k * VSTRIDE];
        for (i = PF; i + PF <= nvec; i += PF) {
            uint4 b[PF];
#pragma unroll
            for (int k = 0; k < PF; ++k) b[k] = v[(i + k) * VSTRIDE];
#pragma unroll
            for (int k = 0; k < PF; ++k) consume(a[k]);
#pragma unroll
            for (int k = 0; k < PF; ++k) a[k] = b[k];
        }
#pragma unroll
        for (int k = 0; k < PF; ++k) consume(a[k]);
    }
    for (; i < nvec; ++i) consume(v[i * VSTRIDE]);
}

// As stream_body, but whole blocks of PF vectors go to `block` (given the
// array) and only the last nvec % PF vectors go to `one`: a consumer that
// tests a wave-wide condition once per block (instead of once per vector)
// keeps each block's code straight-line, so the compiler can issue all of a
// block's table reads ahead of the dependent chain that uses them.
template <int PF, int VSTRIDE = 1, class FB, class F1>
__device__ __forceinline__ void stream_blocks(const uint4* __restrict__ v, int64_t nvec, FB&& block, F1&& one) {
    int64_t i = 0;
    if (nvec >= PF) {
        uint4 a[PF];
#pragma unroll
        for (int k = 0; k < PF; ++k) a[k] = v[(int64_t)k * VSTRIDE];
        for (i = PF; i + PF <= nvec; i += PF) {
            uint4 b[PF];
#pragma unroll
            for (int k = 0; k < PF; ++k) b[k] = v[(i + k) * VSTRIDE];
            block(a);
#pragma unroll
            for (int k = 0; k < PF; ++k) a[k] = b[k];
        }
        block(a);
    }
    for (; i < nvec; ++i) one(v[i * VSTRIDE]);
}

// Two streams consumed side by side: pairs (a[i], b[i]) for i < nb, then
// a[nb] alone when `extra_a` (a holds nb or nb + 1 vectors). PF vectors of
// each stream are in flight while the previous block is consumed, so the two
// consumers' dependent chains interleave.
template <int PF, int VSTRIDE = 1, class F2, class F1>
__device__ __forceinline__ void stream_body2(const uint4* __restrict__ va, const uint4* __restrict__ vb, int64_t nb,
                                             bool extra_a, F2&& both, F1&& only_a) {
    int64_t i = 0;
    if (nb >= PF) {
        uint4 a[PF], b[PF];
#pragma unroll
        for (int k = 0; k < PF; ++k) {
            a[k] = va[(int64_t)k * VSTRIDE];
            b[k] = vb[(int64_t)k * VSTRIDE];
        }
        for (i = PF; i + PF <= nb; i += PF) {
            uint4 a2[PF], b2[PF];
#pragma unroll
            for (int k = 0; k < PF; ++k) {
                a2[k] = va[(i + k) * VSTRIDE];
                b2[k] = vb[(i + k) * VSTRIDE];
            }
#pragma unroll
            for (int k = 0; k < PF; ++k) both(a[k], b[k]);
#pragma unroll
            for (int k = 0; k < PF; ++k) {
                a[k] = a2[k];
                b[k] = b2[k];
            }
        }
#pragma unroll
        for (int k = 0; k < PF; ++k) both(a[k], b[k]);
    }
    for (; i < nb; ++i) both(va[i * VSTRIDE], vb[i * VSTRIDE]);
    if (extra_a) only_a(va[nb * VSTRIDE]);
}

// v[i] by a non-temporal load (data read once: no reuse in the caches).
__device__ __forceinline__ uint4 load16_nt(const uint4* __restrict__ v, int64_t i) {
    typedef unsigned int v4u __attribute__((ext_vector_type(4)));
    const v4u x = __builtin_nontemporal_load(reinterpret_cast<const v4u*>(v) + i);
    return make_uint4(x.x, x.y, x.z, x.w);
}

// stream_body with non-temporal loads, for access patterns where one wave
// load covers whole cache lines that are not read again.
template <int PF, int VSTRIDE = 1, class F>
__device__ __forceinline__ void stream_body_nt(const uint4* __restrict__ v, int64_t nvec, F&& consume) {
    int64_t i = 0;
    if (nvec >= PF) {
        uint4 a[PF];
#pragma unroll
        for (int k = 0; k < PF; ++k) a[k] = load16_nt(v, (int64_t)k * VSTRIDE);
        for (i = PF; i + PF <= nvec; i += PF) {
            uint4 b[PF];
#pragma unroll
            for (int k = 0; k < PF; ++k) b[k] = load16_nt(v, (i + k) * VSTRIDE);
#pragma unroll
            for (int k = 0; k < PF; ++k) consume(a[k]);
#pragma unroll
            for (int k = 0; k < PF; ++k) a[k] = b[k];
        }
#pragma unroll
        for (int k = 0; k < PF; ++k) consume(a[k]);
    }
    for (; i < nvec; ++i) consume(load16_nt(v, i * VSTRIDE));
}

// Copy `n` bytes (multiple of 16) from global to LDS with the whole block.
__device__ __forceinline__ void stage_table(uint8_t* __restrict__ dst, const uint8_t* __restrict__ src, int n) {
    const uint4* s4 = reinterpret_cast<const uint4*>(src);
    uint4* d4 = reinterpret_cast<uint4*>(dst);
    for (int i = threadIdx.x; i < (n >> 4); i += blockDim.x) d4[i] = s4[i];
}

// ---------------------------------------------------------------------------
// 32-byte symbol windows
// ---------------------------------------------------------------------------

// Records whether any body byte lies outside the aligned window [base, base+32)
// (win4 = base replicated in four bytes), two bit operations per word.
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

// Pair index sym0 | sym1 << 5 of the low / high two bytes of a word (sym = byte & 31).
__device__ __forceinline__ uint32_t pair_lo(uint32_t x) { return (x & 31u) | (((x >> 8) & 31u) << 5); }
__device__ __forceinline__ uint32_t pair_hi(uint32_t x) { return ((x >> 16) & 31u) | (((x >> 24) & 31u) << 5); }

// The eight pair functions (nibble s = successor of s) of a 16-byte vector:
// table reads only, independent of the state.
__device__ __forceinline__ void pair_fns(const uint64_t* __restrict__ f2, const uint4& x, uint64_t (&e)[8]) {
    e[0] = f2[pair_lo(x.x)]; e[1] = f2[pair_hi(x.x)];
    e[2] = f2[pair_lo(x.y)]; e[3] = f2[pair_hi(x.y)];
    e[4] = f2[pair_lo(x.z)]; e[5] = f2[pair_hi(x.z)];
    e[6] = f2[pair_lo(x.w)]; e[7] = f2[pair_hi(x.w)];
}

// PF vectors walked from state s through the pair functions, with the next
// vector's eight table reads issued before the current vector's dependent
// chain (a shift and a mask per pair), so LDS latency overlaps the chain even
// with one wave per SIMD (the compiler's own schedule issued four reads and
// waited on each in turn). Also records the window check of every word.
template <int PF, class WC>
__device__ __forceinline__ uint32_t walk_pairs(const uint64_t* __restrict__ f2, const uint4 (&a)[PF], uint32_t s,
                                               WC& w) {
    uint64_t e[8];
    pair_fns(f2, a[0], e);
#pragma unroll
    for (int k = 0; k < PF; ++k) {
        uint64_t n[8];
        if (k + 1 < PF) {
            // the next vector's reads go out once the previous vector's chain
            // is done (an empty asm ties the words to the state): one vector
            // ahead, not the whole block (which took ~220 VGPRs)
            uint4 x = a[k + 1];
            asm volatile("" : "+v"(x.x), "+v"(x.y), "+v"(x.z), "+v"(x.w) : "v"(s));
            pair_fns(f2, x, n);
        }
        w.add(a[k].x); w.add(a[k].y); w.add(a[k].z); w.add(a[k].w);
#pragma unroll
        for (int j = 0; j < 8; ++j) s = (uint32_t)(e[j] >> (s << 2)) & 15u;
        if (k + 1 < PF) {
#pragma unroll
            for (int j = 0; j < 8; ++j) e[j] = n[j];
        }
    }
    return s;
}

// ---------------------------------------------------------------------------
// 16-state functions
// ---------------------------------------------------------------------------

struct Fn16 {
    uint32_t d[4];   // byte s (s = 0..15) of the 16-byte vector = image of state s
};

__device__ __forceinline__ Fn16 fn_identity() {
    return Fn16{{0x03020100u, 0x07060504u, 0x0b0a0908u, 0x0f0e0d0cu}};
}

__device__ __forceinline__ Fn16 fn_const(uint32_t s) {
    const uint32_t c = s * 0x01010101u;
    return Fn16{{c, c, c, c}};
}

// Bytes g[f.byte[i]] for the four entries packed in `f` (values < 16).
__device__ __forceinline__ uint32_t fn_gather(uint32_t f, const Fn16& g) {
    const uint32_t sel = f & 0x07070707u;
    const uint32_t lo = __builtin_amdgcn_perm(g.d[1], g.d[0], sel);   // g[0..7]
    const uint32_t hi = __builtin_amdgcn_perm(g.d[3], g.d[2], sel);   // g[8..15]
    const uint32_t pick = 0x03020100u | ((f & 0x08080808u) >> 1);     // byte i of hi when bit 3 is set
    return __builtin_amdgcn_perm(hi, lo, pick);
}

// (g o f): apply f, then g.
__device__ __forceinline__ Fn16 fn_then(const Fn16& f, const Fn16& g) {
    return Fn16{{fn_gather(f.d[0], g), fn_gather(f.d[1], g), fn_gather(f.d[2], g), fn_gather(f.d[3], g)}};
}

// The identity on states 0 .. n-1 with the slots of absent states (s >= n,
// whose table rows are padded with the identity) started at state 0: every
// slot then always holds the image of a real state, so the function is
// constant on the real states exactly when all 16 slots are equal.
__device__ __forceinline__ Fn16 fn_identity_on(uint32_t n) {
    Fn16 f = fn_identity();
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int k = (int)n - 4 * i;   // real slots in word i
        f.d[i] &= k >= 4 ? 0xffffffffu : (k <= 0 ? 0u : (1u << (8 * k)) - 1u);
    }
    return f;
}

__device__ __forceinline__ bool fn_is_const(const Fn16& f) {
    const uint32_t c = (f.d[0] & 0xffu) * 0x01010101u;
    return (f.d[0] == c) & (f.d[1] == c) & (f.d[2] == c) & (f.d[3] == c);
}

// f(s): byte s & 7 of each half by a byte permute (selector byte 0x0c reads
// as zero), then the half by a select on the results. Choosing the words
// first, or indexing f.d with s, lets the compiler put f in scratch memory.
__device__ __forceinline__ uint32_t fn_apply(const Fn16& f, uint32_t s) {
    const uint32_t sel = 0x0c0c0c00u | (s & 7u);
    const uint32_t lo = __builtin_amdgcn_perm(f.d[1], f.d[0], sel);
    const uint32_t hi = __builtin_amdgcn_perm(f.d[3], f.d[2], sel);
    return (s & 8u) ? hi : lo;
}

__device__ __forceinline__ uint64_t fn_pack(const Fn16& f) {
    uint64_t out = 0;
#pragma unroll
    for (uint32_t s = 0; s < 16; ++s) out |= (uint64_t)fn_apply(f, s) << (4 * s);
    return out;
}

__device__ __forceinline__ uint32_t nib(uint64_t f, uint32_t s) { return (uint32_t)(f >> (4 * s)) & 15u; }

// (g o f) on 64-bit nibble functions.
__device__ __forceinline__ uint64_t nib_then(uint64_t f, uint64_t g) {
    uint64_t out = 0;
#pragma unroll
    for (uint32_t s = 0; s < 16; ++s) out |= (uint64_t)nib(g, nib(f, s)) << (4 * s);
    return out;
}

__device__ __forceinline__ Fn16 load_row(const uint8_t* __restrict__ table, uint32_t row) {
    const uint4 v = *reinterpret_cast<const uint4*>(table + row * 16u);
    return Fn16{{v.x, v.y, v.z, v.w}};
}

// ---------------------------------------------------------------------------
// Cross-wave publication (agent scope)
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t load_acquire(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
// For spin loops: a relaxed agent-scope load sees other agents' releases
// without invalidating caches on every iteration; once the awaited value is
// seen, one fence_acquire orders the reads that follow.
__device__ __forceinline__ uint32_t load_relaxed(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void fence_acquire() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }
__device__ __forceinline__ void store_release(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
// For a status word that carries everything its readers use (a prefix state):
// no release, so no write-back of the whole L2 ahead of the store (which an
// agent-scope release costs on every publishing wave).
__device__ __forceinline__ void store_relaxed(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A wait bounded twice: by a count of unsuccessful polls (max_spins, 0 = no
// count limit) and by elapsed time on the device's constant-rate wall clock
// (max_ticks ticks of wall_clock64, whose rate the host reads as
// hipDeviceAttributeWallClockRate). The clock starts at the first
// unsuccessful poll, so a wait that never has to poll costs no clock read.
struct BoundedWait {
    uint32_t spins;
    uint32_t max_spins;
    uint64_t t0;
    uint64_t max_ticks;
    __device__ __forceinline__ BoundedWait(uint32_t ms, uint64_t mt) : spins(0), max_spins(ms), t0(0), max_ticks(mt) {}
    // Call after each unsuccessful poll: true once either bound is passed.
    __device__ __forceinline__ bool expired() {
        const uint64_t now = (uint64_t)wall_clock64();
        if (spins == 0u) t0 = now;
        ++spins;
        if (max_spins != 0u && spins > max_spins) return true;
        return now - t0 > max_ticks;
    }
};

// ---------------------------------------------------------------------------
// Boolean NFA-64 step over nibble-sliced successor tables
// ---------------------------------------------------------------------------

// One step through the table of class k (256 uint64 entries, entry (p, v) =
// union of the successors of the states 4p + i for the bits i of v), placed
// in LDS as the host lays it out: entry (p, v) at row p ^ h, slot v ^ x with
// h = (k >> 4) & 1, x = k & 15, so that lanes looking up the same nibble value
// in different classes hit different banks. XOR-ing every nibble of the state
// with x gives the slot directly; rows of even and odd p move by +h and -h.
// Nibble n of w as one v_bfe_u32 (the compiler splits (w >> 4n) & 15 into a
// shift and an and).
template <int N>
__device__ __forceinline__ uint32_t nibble(uint32_t w) {
    if constexpr (N == 0) {
        return w & 15u;
    } else {
        uint32_t r;
        asm("v_bfe_u32 %0, %1, %2, 4" : "=v"(r) : "v"(w), "n"(4 * N));
        return r;
    }
}

__device__ __forceinline__ uint64_t nfa_step(const uint64_t* __restrict__ tab, uint32_t k, uint64_t s) {
    const uint32_t x = (k & 15u) * 0x11111111u;
    const uint32_t lo = (uint32_t)s ^ x, hi = (uint32_t)(s >> 32) ^ x;
    // Row 0 of the even rows (p + h) and row 1 of the odd rows (p + 1 - h):
    // both at or after the class's table, so row p's offset p * 16 entries
    // goes into the read's immediate and a lookup is an extract and a
    // shift-add.
    const uint32_t h = ((k >> 4) & 1u) * 16u;
    const uint64_t* te = tab + k * 256u + h;
    const uint64_t* to = tab + k * 256u + (16u - h);
    uint64_t r0 = 0, r1 = 0;
    r0 |= te[0 * 16 + nibble<0>(lo)];
    r1 |= to[0 * 16 + nibble<1>(lo)];
    r0 |= te[2 * 16 + nibble<2>(lo)];
    r1 |= to[2 * 16 + nibble<3>(lo)];
    r0 |= te[4 * 16 + nibble<4>(lo)];
    r1 |= to[4 * 16 + nibble<5>(lo)];
    r0 |= te[6 * 16 + nibble<6>(lo)];
    r1 |= to[6 * 16 + nibble<7>(lo)];
    r0 |= te[8 * 16 + nibble<0>(hi)];
    r1 |= to[8 * 16 + nibble<1>(hi)];
    r0 |= te[10 * 16 + nibble<2>(hi)];
    r1 |= to[10 * 16 + nibble<3>(hi)];
    r0 |= te[12 * 16 + nibble<4>(hi)];
    r1 |= to[12 * 16 + nibble<5>(hi)];
    r0 |= te[14 * 16 + nibble<6>(hi)];
    r1 |= to[14 * 16 + nibble<7>(hi)];
    return r0 | r1;
}

// ---------------------------------------------------------------------------
// 5-bit slices: 13 lookups per byte instead of 16.
//
//     F[c][r][v] = union of succ[c][5r + i] over the bits i of v   (uint64)
//
// for rows r = 0..12 (row 12 holds states 60..63, four bits). Only the 32
// window classes are held (106,752 bytes of LDS): head and tail bytes enter
// the window check like the body and step with class byte & 31. Class k's
// table starts at byte k * kClass5 with kClass5 = 13 * 256 + 8, so entry
// (r, v) of class k falls in bank pair (v + k) mod 32: lanes reading the same
// slot value in different classes hit different banks, with no per-lookup
// cost. Row r's offset r * 256 goes into the read's immediate.
// ---------------------------------------------------------------------------
constexpr uint32_t kClass5 = 13u * 256u + 8u;
constexpr uint32_t kTab5Bytes = 32u * kClass5;
// Bits [OFF, OFF + 5) of w as one v_bfe_u32 (bits beyond 31 read as zero).
template <int OFF>
__device__ __forceinline__ uint32_t field5(uint32_t w) {
    if constexpr (OFF == 0) {
        return w & 31u;
    } else {
        uint32_t r;
        asm("v_bfe_u32 %0, %1, %2, 5" : "=v"(r) : "v"(w), "n"(OFF));
        return r;
    }
}

__device__ __forceinline__ uint64_t nfa5_step(const char* __restrict__ tb, uint32_t k, uint64_t s) {
    const uint32_t lo = (uint32_t)s, hi = (uint32_t)(s >> 32);
    // the class's byte offset, computed once per step: left visible, the
    // compiler folds k * kClass5 into every lookup's address (one more
    // instruction per lookup)
    uint32_t base = k * kClass5;
    asm("" : "+v"(base));
    const char* t = tb + base;
    auto ld = [t](uint32_t off) { return *reinterpret_cast<const uint64_t*>(t + off); };
    uint64_t r0 = 0, r1 = 0;
    r0 |= ld(0 * 256 + field5<0>(lo) * 8u);
    r1 |= ld(1 * 256 + field5<5>(lo) * 8u);
    r0 |= ld(2 * 256 + field5<10>(lo) * 8u);
    r1 |= ld(3 * 256 + field5<15>(lo) * 8u);
    r0 |= ld(4 * 256 + field5<20>(lo) * 8u);
    r1 |= ld(5 * 256 + field5<25>(lo) * 8u);
    r0 |= ld(6 * 256 + (__builtin_amdgcn_alignbit(hi, lo, 30) & 31u) * 8u);   // bits 30..34
    r1 |= ld(7 * 256 + field5<3>(hi) * 8u);
    r0 |= ld(8 * 256 + field5<8>(hi) * 8u);
    r1 |= ld(9 * 256 + field5<13>(hi) * 8u);
    r0 |= ld(10 * 256 + field5<18>(hi) * 8u);
    r1 |= ld(11 * 256 + field5<23>(hi) * 8u);
    r0 |= ld(12 * 256 + field5<28>(hi) * 8u);                                   // bits 60..63
    return r0 | r1;
}

// ---------------------------------------------------------------------------
// 6/5-bit slices: 12 lookups per byte. Each 32-bit half of the state is cut
// into fields of 6, 6, 5, 5, 5 and 5 bits (no field crosses the halves):
//
//     G[c][r][v] = union of succ[c][b_r + i] over the bits i of v
//
// for the 12 fields r with first bit b_r. Rows of 6-bit fields hold 64
// entries (512 bytes), rows of 5-bit fields 32 (256 bytes): 4,096 bytes per
// class; class k's table starts at byte k * kClass6 with kClass6 = 4096 + 8,
// so entry v of any row falls in bank pair (v + k + row offset / 8) mod 32.
// ---------------------------------------------------------------------------
constexpr uint32_t kClass6 = 4096u + 8u;
constexpr uint32_t kTab6Bytes = 32u * kClass6;

template <int OFF, int W>
__device__ __forceinline__ uint32_t field(uint32_t w) {
    if constexpr (OFF == 0) {
        return w & ((1u << W) - 1u);
    } else if constexpr (OFF + W == 32) {
        return w >> OFF;
    } else {
        uint32_t r;
        asm("v_bfe_u32 %0, %1, %2, %3" : "=v"(r) : "v"(w), "n"(OFF), "n"(W));
        return r;
    }
}

__device__ __forceinline__ uint64_t nfa6_step(const char* __restrict__ tb, uint32_t k, uint64_t s) {
    const uint32_t lo = (uint32_t)s, hi = (uint32_t)(s >> 32);
    uint32_t base = k * kClass6;   // kept in a register: see nfa5_step
    asm("" : "+v"(base));
    const char* t = tb + base;
    auto ld = [t](uint32_t off) { return *reinterpret_cast<const uint64_t*>(t + off); };
    // row byte offsets: 0, 512, 1024, 1280, 1536, 1792 for a half, + 2048 for hi
    uint64_t r0 = 0, r1 = 0;
    r0 |= ld(0 + field<0, 6>(lo) * 8u);
    r1 |= ld(512 + field<6, 6>(lo) * 8u);
    r0 |= ld(1024 + field<12, 5>(lo) * 8u);
    r1 |= ld(1280 + field<17, 5>(lo) * 8u);
    r0 |= ld(1536 + field<22, 5>(lo) * 8u);
    r1 |= ld(1792 + field<27, 5>(lo) * 8u);
    r0 |= ld(2048 + field<0, 6>(hi) * 8u);
    r1 |= ld(2560 + field<6, 6>(hi) * 8u);
    r0 |= ld(3072 + field<12, 5>(hi) * 8u);
    r1 |= ld(3328 + field<17, 5>(hi) * 8u);
    r0 |= ld(3584 + field<22, 5>(hi) * 8u);
    r1 |= ld(3840 + field<27, 5>(hi) * 8u);
    return r0 | r1;
}

}  // namespace trexdev
