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
//    to walking that single state (the slots of absent states, s >= n_states,
//    start at state 0 so that they never hold F from being seen constant).
//    The chunk then publishes AGGREGATE (its function), or PREFIX (its final
//    state) for a first chunk and for any chunk whose function is constant
//    (its final state then does not depend on the chunks before it).
//  Phase 2 (look-back, cuda/monoid_scan.cu:128-222 restated over functions):
//    walk back over predecessors, composing their aggregates, until one with
//    PREFIX is found; the chunk's prefix state is F(acc(p)). Waiting is only
//    for predecessors' phase-1 results, which never wait on anything.
//
// Forward progress: chunks are claimed 256 at a time (one workgroup, 64 per
// wave) from an atomic ticket in index order, and a chunk's predecessor
// always has a smaller index. A predecessor in the same wave is taken from
// registers (shuffles), so every chunk waited on belongs to an earlier wave of
// the same claim or to an earlier claim, held by a running wave that
// publishes without waiting on later chunks (the reference's
// blockIdx.x order, cuda/monoid_scan.cu:141, gives no such guarantee). Waits
// are bounded by a poll count (`max_spins`, 0 = none) and by wall-clock ticks
// (`max_ticks`); on expiry the kernel sets `error` and moves on, so every wave
// always exits.
//
// status[c]: 0 = not ready, 1 = aggregate in agg[c], 2 | state << 8 = prefix.
// status, the ticket and the error word are cleared by fill_u32 before each
// launch (a kernel, not a memset, so the sequence can be graph-captured).

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "device_common.h"

#define TREX_BLOCK 256

namespace {

using namespace trexdev;

constexpr uint32_t kStatusAggregate = 1u;
constexpr uint32_t kStatusPrefix = 2u;
constexpr int kScanPF = 8;   // 16-byte vectors in flight per lane

// The per-lane scanner: either walking one state or composing a function.
// The walk reads nibble-packed transition functions (nibble s of f1[b] /
// f2[pair] = successor of s), so the LDS address of a step does not depend on
// the state: the reads of a whole vector issue together and only a shift and
// a mask per step wait on the previous state. Composition gathers the 16-byte
// rows of t1 / t2.
template <bool WINDOW>
struct Scanner {
    const uint8_t* t2;    // pair rows [1024][16] (WINDOW)
    const uint8_t* t1;    // byte rows [256][16]
    const uint64_t* f2;   // pair functions [1024] (WINDOW)
    const uint64_t* f1;   // byte functions [256]
    bool walking;
    uint32_t s;
    Fn16 f;
    WinCheck w;

    __device__ __forceinline__ void byte(uint32_t b) {
        if (walking) s = (uint32_t)(f1[b] >> (s << 2)) & 15u;
        else f = fn_then(f, load_row(t1, b));
    }
    __device__ __forceinline__ void word(uint32_t x) {
        if constexpr (WINDOW) {
            w.add(x);
            if (walking) {
                const uint64_t a = f2[pair_lo(x)], b = f2[pair_hi(x)];
                s = (uint32_t)(a >> (s << 2)) & 15u;
                s = (uint32_t)(b >> (s << 2)) & 15u;
            } else {
                f = fn_then(f, load_row(t2, pair_lo(x)));
                f = fn_then(f, load_row(t2, pair_hi(x)));
            }
        } else {
            byte(x & 255u); byte((x >> 8) & 255u); byte((x >> 16) & 255u); byte(x >> 24);
        }
    }
    // After a vector: a constant function becomes a single walked state.
    __device__ __forceinline__ void settle() {
        if (!walking && fn_is_const(f)) {
            walking = true;
            s = f.d[0] & 0xffu;
        }
    }
    // One walked word, no mode test: the lookups depend only on the bytes.
    __device__ __forceinline__ void walk_word(uint32_t x) {
        if constexpr (WINDOW) {
            w.add(x);
            const uint64_t a = f2[pair_lo(x)], b = f2[pair_hi(x)];
            s = (uint32_t)(a >> (s << 2)) & 15u;
            s = (uint32_t)(b >> (s << 2)) & 15u;
        } else {
            const uint64_t a = f1[x & 255u], b = f1[(x >> 8) & 255u], c = f1[(x >> 16) & 255u], d = f1[x >> 24];
            s = (uint32_t)(a >> (s << 2)) & 15u;
            s = (uint32_t)(b >> (s << 2)) & 15u;
            s = (uint32_t)(c >> (s << 2)) & 15u;
            s = (uint32_t)(d >> (s << 2)) & 15u;
        }
    }
    __device__ __forceinline__ void walk_vector(const uint4& a) {
        walk_word(a.x); walk_word(a.y); walk_word(a.z); walk_word(a.w);
    }
    // A block of PF vectors: the mode test is made once for the block and the
    // whole wave, so a walking wave runs the block as straight-line code.
    template <int PF>
    __device__ __forceinline__ void block(const uint4 (&a)[PF]) {
        if (__ballot(!walking) == 0ull) {
            if constexpr (WINDOW) {
                s = walk_pairs<PF>(f2, a, s, w);
            } else {
#pragma unroll
                for (int k = 0; k < PF; ++k) walk_vector(a[k]);
            }
        } else {
#pragma unroll
            for (int k = 0; k < PF; ++k) vector(a[k]);
        }
    }
    // One 16-byte vector. The mode test is made once per vector and for the
    // whole wave: when every active lane walks, the branch-free path lets the
    // vector's table reads issue together (a per-lane test around each word
    // made every word wait for its own reads).
    __device__ __forceinline__ void vector(const uint4& a) {
        if (__ballot(!walking) == 0ull) {
            walk_word(a.x); walk_word(a.y); walk_word(a.z); walk_word(a.w);
        } else {
            word(a.x); word(a.y); word(a.z); word(a.w);
            settle();
        }
    }
    // A vector of this scanner and one of `o`, a second scanner over later
    // bytes of the same chunk: when both walk in every lane, the two walks
    // are interleaved word by word (two independent dependent chains).
    __device__ __forceinline__ void vector2(const uint4& a, Scanner& o, const uint4& b) {
        if (__ballot(!walking || !o.walking) == 0ull) {
            walk_word(a.x); o.walk_word(b.x);
            walk_word(a.y); o.walk_word(b.y);
            walk_word(a.z); o.walk_word(b.z);
            walk_word(a.w); o.walk_word(b.w);
        } else {
            vector(a);
            o.vector(b);
        }
    }
    // This scanner followed by `o` (which started composing from the identity
    // on the bytes right after this one's): o's constant function, or o's
    // function applied to this one's state or composed after its function.
    __device__ __forceinline__ void join(const Scanner& o) {
        if (o.walking) {
            walking = true;
            s = o.s;
        } else if (walking) {
            s = fn_apply(o.f, s);
        } else {
            f = fn_then(f, o.f);
        }
        w.any_hi |= o.w.any_hi;
        w.all &= o.w.all;
    }
};

// Always inlined into its kernels, at 8 vectors in flight (<= 106 VGPRs, no
// stack). A build with 16 in flight compiled the byte kernels' copy as a
// called function (256 VGPRs, generic pointers, a private segment) and its
// first GPU run did not finish; that build also let a lane wait on a chunk of
// its own wave through memory, which the in-wave chains below no longer do.
// The deeper prefetch measured no faster, so it stays off.
// IMAGE: `bytes` is the chunk-interleaved image of the plan (tile t = the
// wave's 64 chunks at byte tile_off[t], vector i of lane k's chunk at
// tile_off[t] + (64 i + k) * 16, so one wave load reads 1 KiB contiguously);
// otherwise the bytes the plan's ranges index.
template <bool WINDOW, bool IMAGE, int CHAINS>
__device__ __forceinline__ void chunk_scan(const uint8_t* __restrict__ bytes, const int64_t* __restrict__ tile_off,
                                           const int64_t* __restrict__ cbegin,
                                           const int64_t* __restrict__ cend, const int32_t* __restrict__ csid,
                                           const int32_t* __restrict__ cpred, const uint32_t* __restrict__ cflags,
                                           int64_t n_chunks, const uint8_t* __restrict__ t2g,
                                           const uint8_t* __restrict__ t1g, const uint64_t* __restrict__ f2g,
                                           const uint8_t* __restrict__ accg, uint32_t win4, uint32_t start,
                                           uint32_t dead, uint32_t n_states, uint32_t* __restrict__ status,
                                           uint64_t* __restrict__ agg, int32_t* __restrict__ jump,
                                           uint32_t* __restrict__ ticket, uint32_t* __restrict__ error,
                                           uint32_t max_spins, uint64_t max_ticks, uint32_t hold,
                                           uint8_t* __restrict__ out_state,
                                           uint8_t* __restrict__ out_accept)
{
    __shared__ __attribute__((aligned(16))) uint8_t t2[WINDOW ? 1024 * 16 : 16];
    __shared__ __attribute__((aligned(16))) uint8_t t1[256 * 16];
    __shared__ uint64_t f2[WINDOW ? 1024 : 1];
    __shared__ uint64_t f1[256];
    __shared__ __attribute__((aligned(16))) uint8_t acc[16];
    {
        const uint4* s4 = reinterpret_cast<const uint4*>(t1g);
        for (int i = threadIdx.x; i < 256; i += blockDim.x) {
            const uint4 row = s4[i];
            reinterpret_cast<uint4*>(t1)[i] = row;
            // nibble-pack the row (entries are states < 16)
            const uint32_t w[4] = {row.x, row.y, row.z, row.w};
            uint64_t fn = 0;
#pragma unroll
            for (int k = 0; k < 16; ++k) fn |= (uint64_t)((w[k >> 2] >> ((k & 3) * 8)) & 15u) << (4 * k);
            f1[i] = fn;
        }
        if constexpr (WINDOW) {
            const uint4* p4 = reinterpret_cast<const uint4*>(t2g);
            for (int i = threadIdx.x; i < 1024; i += blockDim.x) {
                reinterpret_cast<uint4*>(t2)[i] = p4[i];
                f2[i] = f2g[i];
            }
        }
        if (threadIdx.x < 16) acc[threadIdx.x] = accg[threadIdx.x];
    }
    __syncthreads();

    const uint32_t lane = threadIdx.x & 63u;
    __shared__ uint32_t group_base;
    while (true) {
        // one ticket per workgroup: TREX_BLOCK consecutive chunks, 64 per
        // wave. One device-scope atomic on one address per workgroup rather
        // than per wave: such atomics are served one after another, and a
        // ticket per wave kept the last waves of a launch waiting for theirs
        // (about 70 ns each; 0.28 ms for 4,096 waves). Chunks of earlier
        // waves of the same workgroup are held by waves that run alongside.
        __syncthreads();
        if (threadIdx.x == 0) group_base = atomicAdd(ticket, (uint32_t)TREX_BLOCK);
        __syncthreads();
        const uint32_t gb = group_base;
        if ((int64_t)gb >= n_chunks) break;   // the same for the whole workgroup
        const uint32_t base = gb + (threadIdx.x & ~63u);
        if ((int64_t)base >= n_chunks) continue;   // this wave's share is past the end
        const int64_t c = (int64_t)base + lane;
        const bool active = c < n_chunks;

        Scanner<WINDOW> sc;
        sc.t2 = t2;
        sc.t1 = t1;
        sc.f2 = f2;
        sc.f1 = f1;
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
        sc.f = fn_identity_on(n_states);

        // ---- phase 1: walk or compose this chunk's bytes ----
        // CHAINS == 2: the body's first half goes through `sc`, its second
        // half through `sb`, which composes from the identity (and walks once
        // its function is constant), side by side; then sb is joined to sc.
        Scanner<WINDOW> sb;
        sb.t2 = t2;
        sb.t1 = t1;
        sb.f2 = f2;
        sb.f1 = f1;
        sb.w.init();
        sb.walking = false;
        sb.s = start;
        sb.f = fn_identity_on(n_states);
        auto vector = [&](const uint4& a) { sc.vector(a); };
        auto vector2 = [&](const uint4& a, const uint4& b) { sc.vector2(a, sb, b); };
        // the body (nvec vectors, VSTRIDE apart) through one or two chains
        auto body = [&](const uint4* v, int64_t nvec, auto stride_tag) {
            constexpr int VS = decltype(stride_tag)::value;
            if constexpr (CHAINS == 2) {
                const int64_t na = (nvec + 1) >> 1, nb = nvec - na;
                stream_body2<kScanPF / 2, VS>(v, v + na * VS, nb, na > nb, vector2, vector);
                sc.join(sb);
            } else {
                // ordinary loads on the image too: non-temporal ones measured
                // slower on a synchronising automaton (faster on one that keeps
                // composing)
                stream_blocks<kScanPF, VS>(v, nvec, [&](const uint4 (&a)[kScanPF]) { sc.block(a); }, vector);
            }
        };
        if constexpr (IMAGE) {
            // vectors of this lane's chunk are 1 KiB apart; the tail bytes sit
            // at the start of the next vector slot
            const int64_t len = end - p;
            const uint8_t* mine = bytes + tile_off[base >> 6] + 16u * lane;
            const int64_t nvec = len >> 4;
            body(reinterpret_cast<const uint4*>(mine), nvec, std::integral_constant<int, 64>{});
            const uint8_t* tail = mine + nvec * 1024;
            for (int64_t k = 0; k < (len & 15); ++k) sc.byte(tail[k]);
        } else {
            const int64_t head_end =
                min(end, p + (int64_t)((16u - ((uint32_t)(uintptr_t)(bytes + p) & 15u)) & 15u));
            for (; p < head_end; ++p) sc.byte(bytes[p]);
            // body: 16-byte vectors, kScanPF of them in flight while the
            // previous ones are walked (one lane's loads are its own chunk's
            // next bytes)
            const int64_t nvec = (end - p) >> 4;
            body(reinterpret_cast<const uint4*>(bytes + p), nvec, std::integral_constant<int, 1>{});
            p += nvec << 4;
            for (; p < end; ++p) sc.byte(bytes[p]);
        }
        sc.settle();
        if constexpr (WINDOW) {
            if (sc.w.outside(win4)) { sc.walking = true; sc.s = dead; }
        }

        // ---- in-wave chains, resolved in registers ----
        // A lane whose predecessor chunk lies in the same wave (pred >= base)
        // takes that lane's function by shuffle instead of waiting for its
        // publication: pointer jumping over the in-wave predecessor links
        // (six rounds cover a chain of 64 lanes) leaves in g the composition
        // of the chunks from the chain's root lane up to this one, and in
        // `root` the root lane. A root's predecessor is a chunk of an earlier
        // ticket (or none), so the look-back below only ever waits on other
        // waves' chunks: no lane waits on a store of its own wave, whatever
        // order the compiler gives the wave's publish stores and its loop. A
        // first chunk, or any chunk whose function turned constant, enters as
        // a constant function, so everything after it on the chain comes out
        // constant: its final state is known here.
        Fn16 g = (pred < 0 || sc.walking) ? fn_const(sc.s) : sc.f;
        const bool inwave = active && pred >= (int32_t)base;
        uint32_t root = inwave ? (uint32_t)(pred - (int32_t)base) : lane;
        bool closed = !inwave;   // g reaches back to the chain's root
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            Fn16 up;
            up.d[0] = __shfl(g.d[0], (int)root, 64);
            up.d[1] = __shfl(g.d[1], (int)root, 64);
            up.d[2] = __shfl(g.d[2], (int)root, 64);
            up.d[3] = __shfl(g.d[3], (int)root, 64);
            const bool up_closed = __shfl((int)closed, (int)root, 64) != 0;
            const uint32_t up_root = (uint32_t)__shfl((int)root, (int)root, 64);
            if (!closed) {
                g = fn_then(up, g);   // the earlier lanes' chunks apply first
                closed = up_closed;
                root = up_root;
            }
        }
        const int32_t root_pred = __shfl(pred, (int)root, 64);

        // ---- publish ----
        // A constant g publishes PREFIX (its final state). Otherwise agg[c] =
        // g, the function of the chain's chunks from its root up to c, and
        // jump[c] = the root's predecessor: a look-back that meets c composes
        // g and continues at jump[c], one step per chain instead of per chunk.
        // test hook: hold back this tile's publication, tiles later in each
        // group of 8 less, so successors find their predecessors unpublished
        if (hold != 0u) {
            const uint32_t n = hold * (7u - ((base >> 6) & 7u));
            for (uint32_t k = 0; k < n; ++k) __builtin_amdgcn_s_sleep(8);
        }
        const bool known = fn_is_const(g);
        const uint32_t kstate = g.d[0] & 0xffu;
        const uint64_t my_fn = known ? 0ull : fn_pack(g);
        if (active) {
            if (known) {
                store_relaxed(&status[c], kStatusPrefix | (kstate << 8));   // the word says it all
            } else {
                agg[c] = my_fn;
                jump[c] = root_pred;
                store_release(&status[c], kStatusAggregate);   // agg and jump first
            }
        }

        // ---- phase 2: the root of each unresolved chain looks back ----
        // (a chain is unresolved only if its root is: a constant root makes
        // every later lane of the chain constant). Every chunk waited on
        // belongs to an earlier ticket, whose wave publishes without waiting
        // on this one. Waits are bounded by the spin count and by wall-clock
        // ticks; on expiry the error word is set and the wave moves on.
        uint32_t incoming = 0;
        if (active && !known && root == lane) {
            uint64_t acc_fn = 0xfedcba9876543210ull;   // identity
            int32_t j = pred;
            BoundedWait wait(max_spins, max_ticks);
            while (true) {
                const uint32_t st = load_relaxed(&status[j]);
                if (st == 0u) {
                    if (wait.expired()) {
                        atomicOr(error, 1u);
                        incoming = dead;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                    continue;
                }
                if ((st & 0xffu) == kStatusPrefix) {   // the word holds the state
                    incoming = nib(acc_fn, (st >> 8) & 15u);
                    break;
                }
                fence_acquire();   // agg[j] and jump[j] were written before status[j] was released
                acc_fn = nib_then(agg[j], acc_fn);   // earlier chunks apply first
                j = jump[j];
                if (j < 0) {   // cannot happen: a string's first chunk publishes a prefix
                    atomicOr(error, 2u);
                    incoming = dead;
                    break;
                }
            }
        }
        // the state before each chain, from its root, then each lane's own
        incoming = __shfl(incoming, (int)root, 64);
        const uint32_t state = known ? kstate : fn_apply(g, incoming);
        if (active && !known) store_relaxed(&status[c], kStatusPrefix | (state << 8));
        if (active && (flags & 2u)) {
            const int32_t sid = csid[c];
            out_state[sid] = (uint8_t)state;
            out_accept[sid] = acc[state];
        }
    }
}

}  // namespace

#define TREX_CHUNK_PARAMS                                                                                 \
    const uint8_t *__restrict__ bytes, const int64_t *__restrict__ tile_off, const int64_t *__restrict__ cbegin, \
        const int64_t *__restrict__ cend, const int32_t *__restrict__ csid, const int32_t *__restrict__ cpred,  \
        const uint32_t *__restrict__ cflags, int64_t n_chunks, const uint8_t *__restrict__ t2g,               \
        const uint8_t *__restrict__ t1g, const uint64_t *__restrict__ f2g, const uint8_t *__restrict__ accg,   \
        uint32_t win4, uint32_t start, uint32_t dead, uint32_t n_states, uint32_t *__restrict__ status,        \
        uint64_t *__restrict__ agg, int32_t *__restrict__ jump, uint32_t *__restrict__ ticket,                 \
        uint32_t *__restrict__ error, uint32_t max_spins, uint64_t max_ticks, uint32_t hold,                 \
        uint8_t *__restrict__ out_state,                                                                      \
        uint8_t *__restrict__ out_accept
#define TREX_CHUNK_ARGS                                                                                     \
    bytes, tile_off, cbegin, cend, csid, cpred, cflags, n_chunks, t2g, t1g, f2g, accg, win4, start, dead,   \
        n_states, status, agg, jump, ticket, error, max_spins, max_ticks, hold, out_state, out_accept

// Chunks of any string, bytes as the plan's ranges index them (tile_off
// unused); _c2: each chunk's body walked as two chains.
#define TREX_CHUNK_KERNELS(SUFFIX, CHAINS)                                                                   \
    extern "C" __global__ void __launch_bounds__(TREX_BLOCK) dfa16_chunks_window##SUFFIX(TREX_CHUNK_PARAMS) { \
        chunk_scan<true, false, CHAINS>(TREX_CHUNK_ARGS);                                                    \
    }                                                                                                        \
    extern "C" __global__ void __launch_bounds__(TREX_BLOCK) dfa16_chunks_bytes##SUFFIX(TREX_CHUNK_PARAMS) {  \
        chunk_scan<false, false, CHAINS>(TREX_CHUNK_ARGS);                                                   \
    }                                                                                                        \
    /* the same over the plan's chunk-interleaved image */                                                  \
    extern "C" __global__ void __launch_bounds__(TREX_BLOCK)                                                 \
        dfa16_chunks_window_img##SUFFIX(TREX_CHUNK_PARAMS) {                                                 \
        chunk_scan<true, true, CHAINS>(TREX_CHUNK_ARGS);                                                     \
    }                                                                                                        \
    extern "C" __global__ void __launch_bounds__(TREX_BLOCK)                                                 \
        dfa16_chunks_bytes_img##SUFFIX(TREX_CHUNK_PARAMS) {                                                  \
        chunk_scan<false, true, CHAINS>(TREX_CHUNK_ARGS);                                                    \
    }
TREX_CHUNK_KERNELS(, 1)
TREX_CHUNK_KERNELS(_c2, 2)

// Sets dst[0 .. n) = value (status words, tickets, error flags).
extern "C" __global__ void __launch_bounds__(TREX_BLOCK)
fill_u32(uint32_t* __restrict__ dst, int64_t n, uint32_t value)
{
    const int64_t stride = (int64_t)gridDim.x * TREX_BLOCK;
    for (int64_t i = (int64_t)blockIdx.x * TREX_BLOCK + threadIdx.x; i < n; i += stride) dst[i] = value;
}

// ---------------------------------------------------------------------------
// Chunk table built on the device, for look-back launches whose string ranges
// are in device memory (the allocation-free entry point): the same chunks as
// trex_plan_create cuts (chunk_len bytes, the last one shorter, one empty
// chunk for an empty string), in string order, so each chunk's predecessor is
// the chunk just before it.
// ---------------------------------------------------------------------------

// error bits of the table kernels (the look-back's own: bit 0 a wait
// expired, bit 1 a chain without a first chunk)
constexpr uint32_t kTableBadRange = 4u;    // a range with begin < 0, end < begin or end > bytes_len
constexpr uint32_t kTableOverBound = 8u;   // more chunks than the launch was sized for

// One workgroup of 1024 threads: first[s] = index of string s's first chunk,
// first[n_strings] = the chunk count; clears the ticket and sets the error word
// (to 0, or to the bits above).
extern "C" __global__ void __launch_bounds__(1024)
chunk_table_count(const int64_t* __restrict__ begin, const int64_t* __restrict__ end, int64_t n_strings,
                  int64_t bytes_len, int64_t chunk_len, int64_t n_bound, int64_t* __restrict__ first,
                  uint32_t* __restrict__ ticket, uint32_t* __restrict__ error)
{
    __shared__ int64_t wave_sum[16];
    __shared__ int64_t carry_s;
    __shared__ uint32_t bad_s;
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    if (t == 0) { carry_s = 0; bad_s = 0; *ticket = 0; }
    __syncthreads();
    for (int64_t s0 = 0; s0 < n_strings; s0 += 1024) {
        const int64_t s = s0 + t;
        int64_t cnt = 0;
        if (s < n_strings) {
            const int64_t b = begin[s], e = end[s];
            if (b < 0 || e < b || e > bytes_len) {
                atomicOr(&bad_s, kTableBadRange);
                cnt = 1;
            } else {
                const int64_t len = e - b;
                cnt = len <= chunk_len ? 1 : (len + chunk_len - 1) / chunk_len;
            }
        }
        // inclusive scan within the wave, then across the 16 waves
        int64_t x = cnt;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int64_t y = __shfl_up(x, d, 64);
            if ((int)lane >= d) x += y;
        }
        if (lane == 63u) wave_sum[wave] = x;
        __syncthreads();
        int64_t before = carry_s;
        for (uint32_t w = 0; w < wave; ++w) before += wave_sum[w];
        if (s < n_strings) first[s] = before + x - cnt;
        __syncthreads();   // every thread has read carry_s and wave_sum
        if (t == 1023u) carry_s = before + x;
        __syncthreads();
    }
    if (t == 0) {
        const int64_t total = carry_s;
        first[n_strings] = total;
        uint32_t bits = bad_s;
        if (total > n_bound) bits |= kTableOverBound;
        *error = bits;
    }
}

// Chunk slots 0 .. n_bound: the chunks of the table above, then empty chunks
// that publish at once and write no result (the launch is sized by a bound
// computed on the host). Also clears the look-back status. After an error of
// the count kernel every slot is empty, so nothing is matched.
extern "C" __global__ void __launch_bounds__(TREX_BLOCK)
chunk_table_fill(const int64_t* __restrict__ begin, const int64_t* __restrict__ end, int64_t n_strings,
                 int64_t chunk_len, int64_t n_bound, const int64_t* __restrict__ first,
                 const uint32_t* __restrict__ error, int64_t* __restrict__ cbegin, int64_t* __restrict__ cend,
                 int32_t* __restrict__ csid, int32_t* __restrict__ cpred, uint32_t* __restrict__ cflags,
                 uint32_t* __restrict__ status)
{
    const bool ok = (*error & (kTableBadRange | kTableOverBound)) == 0u;
    const int64_t total = ok ? first[n_strings] : 0;
    const int64_t stride = (int64_t)gridDim.x * TREX_BLOCK;
    for (int64_t c = (int64_t)blockIdx.x * TREX_BLOCK + threadIdx.x; c < n_bound; c += stride) {
        int64_t b = 0, e = 0;
        int32_t sid = 0, pred = -1;
        uint32_t fl = 0;
        if (c < total) {
            // the last string whose first chunk is <= c (first[] strictly increases)
            int64_t lo = 0, hi = n_strings - 1;
            while (lo < hi) {
                const int64_t mid = (lo + hi + 1) >> 1;
                if (first[mid] <= c) lo = mid; else hi = mid - 1;
            }
            const int64_t k = c - first[lo], cnt = first[lo + 1] - first[lo];
            b = begin[lo] + k * chunk_len;
            e = min(end[lo], b + chunk_len);
            sid = (int32_t)lo;
            pred = k == 0 ? -1 : (int32_t)(c - 1);
            fl = (k == 0 ? 1u : 0u) | (k == cnt - 1 ? 2u : 0u);
        }
        cbegin[c] = b;
        cend[c] = e;
        csid[c] = sid;
        cpred[c] = pred;
        cflags[c] = fl;
        status[c] = 0u;
    }
}
