// One persistent launch over a mixed batch of DFA-16 and Boolean NFA-64
// patterns and strings of uneven length (gfx950, wave64).
//
// The host (mix.cpp) turns the batch into a queue of work batches. A work
// batch holds up to kMixBlock lane items of ONE pattern and one role, so the
// workgroup that claims it stages that pattern's tables in LDS once (skipped
// when the previous batch had the same pattern). Workgroups claim batches from
// an atomic ticket until the queue is empty. Roles:
//
//   DFA_WHOLE  a string walked from the start state: pair-function table f2
//              (two bytes per LDS lookup) with the 32-byte symbol window check
//   DFA_CHUNK  one chunk of a long string, chained by the decoupled look-back
//              over 16-state functions (the scheme of scan_kernels.hip)
//   NFA_WHOLE  a string walked by the bit-parallel NFA-64 step
//   NFA_SPEC   one chunk of a long string: the first chunk of the string is
//              walked from the start set; every other chunk is walked from
//              the set of all states (a guess) and records its state set
//              every kCk bytes (checkpoints) and at its end
//   NFA_FIXUP  one lane per long string, after all its NFA_SPEC chunks are
//              done: walks each chunk again from the true entry set until the
//              set equals the chunk's checkpoint at the same position, after
//              which both walks coincide, so the chunk's exit is the recorded
//              one. A chunk that never meets its checkpoints is walked to its
//              end, and an empty set ends the fix-up (the empty set is
//              absorbing). The result is exact either way.
//
// Forward progress: the queue puts every batch that is waited on before the
// batches that wait on it (DFA chunks after their predecessors, fix-ups after
// all NFA_SPEC batches), so every waited-on item has been claimed by a running
// workgroup, and no waited-on work waits itself. A DFA chunk never waits on
// a chunk of its own wave (those come by shuffle). Waits are bounded by a
// poll count (max_spins) and by wall-clock ticks (max_ticks); on expiry the
// error word is set and the wave moves on.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.h"

#define MIX_BLOCK 512
#define MIX_BLOCK_SLICE 1024

namespace {

using namespace trexdev;

constexpr uint32_t kRoleDfaWhole = 0, kRoleDfaChunk = 1, kRoleNfaWhole = 2, kRoleNfaSpec = 3, kRoleNfaFixup = 4;
constexpr uint32_t kStatusAggregate = 1u, kStatusPrefix = 2u;
constexpr int kNfaClasses = 34;   // window bytes 0..31, reject, skip

// Pattern descriptor (host layout in mix.cpp).
struct MixPattern {
    uint32_t kind;        // 0 DFA, 1 NFA
    uint32_t win4;        // window base byte x 0x01010101
    uint32_t start;       // DFA start state
    uint32_t dead;        // DFA dead state
    uint32_t nstates;     // DFA state count (<= 16)
    uint32_t reserved;
    uint64_t nstart;      // NFA start set
    uint64_t naccept;     // NFA accept set
    const uint8_t* t1;    // DFA byte table [256][16]
    const uint8_t* t2;    // DFA pair rows [1024][16]
    const uint8_t* f2;    // DFA pair functions [1024] x uint64
    const uint8_t* acc;   // DFA accept [16]
    const uint8_t* ntab;  // NFA window tables: [34][256] x uint64 (bank-swizzled) or 6/5-bit slices
    const uint8_t* ncls;  // NFA byte -> class [256]
};

struct MixArgs {
    const uint8_t* bytes;
    const MixPattern* patterns;
    const int32_t* bfirst;    // batch: first item
    const int32_t* bcount;    // batch: item count
    const int32_t* bpattern;  // batch: pattern
    const int32_t* brole;     // batch: role
    int32_t n_batches;
    int32_t ck;               // NFA checkpoint spacing in bytes (multiple of 16)
    const int64_t* ibeg;      // item: byte range (FIXUP: first NFA_SPEC item)
    const int64_t* iend;      //       (FIXUP: one past the last NFA_SPEC item)
    const int32_t* ia;        // item: DFA_CHUNK predecessor item / NFA string slot
    const int32_t* ib;        // item: output string, -1 if none
    const int32_t* ic;        // item: NFA_SPEC checkpoint ordinal
    const uint32_t* iflags;   // item: bit 0 first chunk of its string
    int32_t n_ck;             // checkpoints per NFA_SPEC item
    uint32_t* ticket;
    uint32_t* status;         // per item (DFA_CHUNK)
    uint64_t* agg;            // per item (DFA_CHUNK)
    uint32_t* done;           // per NFA string slot: finished NFA_SPEC chunks
    uint64_t* ckpt;           // [n_spec][n_ck]
    uint64_t* spec_exit;      // [n_spec]
    uint32_t* error;
    uint32_t* converged;      // fix-up chunks that met a checkpoint
    uint32_t max_spins;       // wait bound in polls (0: none)
    uint32_t reserved;
    uint64_t max_ticks;       // wait bound in wall-clock ticks
    uint64_t* out_mask;       // per string: DFA final state / NFA final set
    uint8_t* out_accept;      // per string
};

// ---------------------------------------------------------------------------
// DFA roles
// ---------------------------------------------------------------------------

struct DfaLds {
    const uint8_t* t1;     // [256][16]
    const uint8_t* t2;     // [1024][16]
    const uint64_t* f2;    // [1024]
    const uint8_t* acc;    // [16]
};

__device__ __forceinline__ uint32_t k2f_word(const uint64_t* __restrict__ f2, uint32_t s, uint32_t x) {
    const uint64_t a = f2[pair_lo(x)];
    const uint64_t b = f2[pair_hi(x)];
    s = (uint32_t)(a >> (s << 2)) & 15u;
    return (uint32_t)(b >> (s << 2)) & 15u;
}

__device__ uint32_t dfa_walk(const uint8_t* __restrict__ bytes, int64_t p, int64_t end, const DfaLds& L,
                             const MixPattern& P) {
    uint32_t s = P.start;
    const int64_t head_end = min(end, p + head_len(bytes, p));
    for (; p < head_end; ++p) s = L.t1[((uint32_t)bytes[p] << 4) | s];
    WinCheck w;
    w.init();
    const int64_t nvec = (end - p) >> 4;
    stream_body<8>(reinterpret_cast<const uint4*>(bytes + p), nvec, [&](uint4 v) {
        w.add(v.x); w.add(v.y); w.add(v.z); w.add(v.w);
        s = k2f_word(L.f2, s, v.x);
        s = k2f_word(L.f2, s, v.y);
        s = k2f_word(L.f2, s, v.z);
        s = k2f_word(L.f2, s, v.w);
    });
    p += nvec << 4;
    for (; p < end; ++p) s = L.t1[((uint32_t)bytes[p] << 4) | s];
    if (w.outside(P.win4)) s = P.dead;
    return s;
}

// Walk one state or compose the function of a chunk (window pair rows).
struct ChunkScanner {
    const DfaLds* L;
    bool walking;
    uint32_t s;
    Fn16 f;
    WinCheck w;
    __device__ __forceinline__ void byte(uint32_t b) {
        if (walking) s = L->t1[(b << 4) | s];
        else f = fn_then(f, load_row(L->t1, b));
    }
    __device__ __forceinline__ void word(uint32_t x) {
        w.add(x);
        if (walking) {
            s = k2f_word(L->f2, s, x);
        } else {
            f = fn_then(f, load_row(L->t2, pair_lo(x)));
            f = fn_then(f, load_row(L->t2, pair_hi(x)));
        }
    }
    __device__ __forceinline__ void settle() {
        if (!walking && fn_is_const(f)) {
            walking = true;
            s = f.d[0] & 0xffu;
        }
    }
    // One 16-byte vector, walked or composed as decided once for the whole
    // wave: when every active lane walks, the branch-free path issues the
    // vector's pair-function reads together (a per-lane test around each
    // word made every word wait for its own reads).
    __device__ __forceinline__ void vector(const uint4& a) {
        if (__ballot(!walking) == 0ull) {
            w.add(a.x); w.add(a.y); w.add(a.z); w.add(a.w);
            s = k2f_word(L->f2, s, a.x);
            s = k2f_word(L->f2, s, a.y);
            s = k2f_word(L->f2, s, a.z);
            s = k2f_word(L->f2, s, a.w);
        } else {
            word(a.x); word(a.y); word(a.z); word(a.w);
            settle();
        }
    }
};

// One chunk per lane; every lane of the workgroup calls it (inactive lanes
// take part in the wave's shuffles). Always inlined into the persistent loop.
__device__ __forceinline__ void dfa_chunk(const MixArgs& A, const MixPattern& P, const DfaLds& L, int64_t item,
                                          bool active) {
    ChunkScanner sc;
    sc.L = &L;
    sc.w.init();
    int32_t pred = -1;
    int64_t p = 0, end = 0;
    if (active) {
        pred = (A.iflags[item] & 1u) ? -1 : A.ia[item];
        p = A.ibeg[item];
        end = A.iend[item];
    }
    sc.walking = pred < 0;
    sc.s = P.start;
    sc.f = fn_identity_on(P.nstates);
    const uint8_t* bytes = A.bytes;
    const int64_t head_end = min(end, p + head_len(bytes, p));
    for (; p < head_end; ++p) sc.byte(bytes[p]);
    const int64_t nvec = (end - p) >> 4;
    stream_body<4>(reinterpret_cast<const uint4*>(bytes + p), nvec, [&](const uint4& a) { sc.vector(a); });
    p += nvec << 4;
    for (; p < end; ++p) sc.byte(bytes[p]);
    sc.settle();
    if (sc.w.outside(P.win4)) { sc.walking = true; sc.s = P.dead; }

    // In-wave chains in registers (as in scan_kernels.hip): a lane whose
    // predecessor item lies in its own wave takes that lane's function by
    // shuffle; pointer jumping leaves in g the composition from the chain's
    // root lane up to this one. Only roots look back through memory, and a
    // root's predecessor belongs to another wave (of this workgroup or of an
    // earlier batch), which publishes without waiting on this one.
    const uint32_t lane = threadIdx.x & 63u;
    const int64_t wave_first = item - (int64_t)lane;
    const bool own_known = pred < 0 || sc.walking;
    const Fn16 own = own_known ? fn_const(sc.s) : sc.f;
    Fn16 g = own;
    const bool inwave = active && (int64_t)pred >= wave_first;
    uint32_t root = inwave ? (uint32_t)((int64_t)pred - wave_first) : lane;
    bool closed = !inwave;
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

    // A chunk whose chain function is constant knows its final state and
    // publishes PREFIX at once; any other publishes its own chunk's function
    // (a look-back from another wave steps chunk by chunk over A.ia).
    const bool known = fn_is_const(g);
    const uint32_t kstate = g.d[0] & 0xffu;
    if (active) {
        if (known) {
            store_relaxed(&A.status[item], kStatusPrefix | (kstate << 8));   // the word says it all
        } else {
            A.agg[item] = fn_pack(own);
            store_release(&A.status[item], kStatusAggregate);   // agg first
        }
    }
    uint32_t incoming = 0;
    if (active && !known && root == lane) {
        uint64_t acc_fn = 0xfedcba9876543210ull;   // identity
        int32_t j = pred;
        BoundedWait wait(A.max_spins, A.max_ticks);
        while (true) {
            const uint32_t st = load_relaxed(&A.status[j]);
            if (st == 0u) {
                if (wait.expired()) {
                    atomicOr(A.error, 1u);
                    incoming = P.dead;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
                continue;
            }
            if ((st & 0xffu) == kStatusPrefix) {   // the word holds the state
                incoming = nib(acc_fn, (st >> 8) & 15u);
                break;
            }
            fence_acquire();   // agg[j] was written before status[j] was released
            acc_fn = nib_then(A.agg[j], acc_fn);   // the earlier chunk applies first
            j = (A.iflags[j] & 1u) ? -1 : A.ia[j];
            if (j < 0) {   // cannot happen: a string's first chunk publishes a prefix
                atomicOr(A.error, 2u);
                incoming = P.dead;
                break;
            }
        }
    }
    incoming = __shfl(incoming, (int)root, 64);
    const uint32_t state = known ? kstate : fn_apply(g, incoming);
    if (active && !known) store_relaxed(&A.status[item], kStatusPrefix | (state << 8));
    if (active && A.ib[item] >= 0) {
        A.out_mask[A.ib[item]] = state;
        A.out_accept[A.ib[item]] = L.acc[state];
    }
}

// ---------------------------------------------------------------------------
// NFA roles
// ---------------------------------------------------------------------------

// The NFA step over the tables of the launch's form: nibble tables (34
// classes, reject and skip included, bank-swizzled) with a byte -> class map,
// or 6/5-bit slice tables of the 32 window classes, where every byte enters the
// window check and steps with class byte & 31.
template <bool SLICE>
struct NfaWalker {
    const void* tab;
    const uint8_t* cls;
    uint64_t s;
    WinCheck w;
    __device__ __forceinline__ uint64_t step(uint32_t k, uint64_t x) const {
        if constexpr (SLICE) return nfa6_step(static_cast<const char*>(tab), k, x);
        else return nfa_step(static_cast<const uint64_t*>(tab), k, x);
    }
    __device__ __forceinline__ void byte(uint32_t b) {
        if constexpr (SLICE) {
            w.add(b * 0x01010101u);
            s = step(b & 31u, s);
        } else {
            s = step(cls[b], s);
        }
    }
    __device__ __forceinline__ void word(uint32_t x) {
        w.add(x);
        s = step(x & 31u, s);
        s = step((x >> 8) & 31u, s);
        s = step((x >> 16) & 31u, s);
        s = step((x >> 24) & 31u, s);
    }
    __device__ __forceinline__ void vec(uint4 v) { word(v.x); word(v.y); word(v.z); word(v.w); }
    // A byte outside the window rejects: once seen, the set is empty for good.
    // Applied before every set that is compared or stored, so both forms give
    // the same sets there.
    __device__ __forceinline__ void settle(uint32_t win4) {
        if (w.outside(win4)) s = 0;
        w.init();
    }
};

template <bool SLICE>
__device__ uint64_t nfa_walk(const uint8_t* __restrict__ bytes, int64_t p, int64_t end, uint64_t s,
                             const void* tab, const uint8_t* cls, uint32_t win4) {
    NfaWalker<SLICE> nw{tab, cls, s, {}};
    nw.w.init();
    const int64_t head_end = min(end, p + head_len(bytes, p));
    for (; p < head_end; ++p) nw.byte(bytes[p]);
    const int64_t nvec = (end - p) >> 4;
    stream_body<2>(reinterpret_cast<const uint4*>(bytes + p), nvec, [&](uint4 x) { nw.vec(x); });
    nw.settle(win4);
    p += nvec << 4;
    for (; p < end; ++p) nw.byte(bytes[p]);
    nw.settle(win4);
    return nw.s;
}

// Walks [p, end) from set s. When `record`, stores the set at every absolute
// position that is a multiple of ck (ck a multiple of 16) into ck_out[0..n_ck).
// When `compare`, stops at the first such position whose set equals
// ck_in[idx] and reports it through `met`.
template <bool SLICE, bool RECORD, bool COMPARE>
__device__ uint64_t nfa_walk_ck(const uint8_t* __restrict__ bytes, int64_t p, int64_t end, uint64_t s,
                                const void* tab, const uint8_t* cls, uint32_t win4, int32_t ck,
                                uint64_t* __restrict__ ck_out, const uint64_t* __restrict__ ck_in, int32_t n_ck,
                                bool& met) {
    NfaWalker<SLICE> nw{tab, cls, s, {}};
    nw.w.init();
    met = false;
    const int64_t head_end = min(end, p + head_len(bytes, p));
    for (; p < head_end; ++p) nw.byte(bytes[p]);
    int32_t idx = 0;
    const uint4* v = reinterpret_cast<const uint4*>(bytes + p);
    const int64_t nvec = (end - p) >> 4;
    // two vectors in flight ahead of the one being walked (a walk that stops
    // at a checkpoint drops them)
    uint4 n0 = make_uint4(0u, 0u, 0u, 0u), n1 = n0;
    if (nvec > 0) n0 = v[0];
    if (nvec > 1) n1 = v[1];
    // the first multiple of ck after p; positions advance by 16 and ck is a
    // multiple of 16, so pos meets every later one exactly (no division per
    // vector)
    int64_t next_ck = (p + 16 + ck - 1) / ck * ck;
    for (int64_t i = 0; i < nvec; ++i) {
        const uint4 cur = n0;
        n0 = n1;
        if (i + 2 < nvec) n1 = v[i + 2];
        nw.vec(cur);
        const int64_t pos = p + ((i + 1) << 4);
        if (pos != next_ck) continue;
        next_ck += ck;
        if (idx < n_ck) {
            nw.settle(win4);
            if constexpr (RECORD) ck_out[idx] = nw.s;
            if constexpr (COMPARE) {
                if (nw.s == ck_in[idx]) {
                    met = true;
                    return nw.s;
                }
            }
            ++idx;
        }
    }
    nw.settle(win4);
    p += nvec << 4;
    for (; p < end; ++p) nw.byte(bytes[p]);
    nw.settle(win4);
    return nw.s;
}

// One persistent workgroup's loop. Dynamic LDS: the NFA tables of the form
// (nibble: kNfaClasses * 2048 + 256 bytes; 6/5-bit slices: kTab6Bytes); the
// DFA tables use the first 29 KB of the same space.
template <bool SLICE>
__device__ __forceinline__ void mix_loop(const MixArgs& A)
{
    extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
    __shared__ int32_t s_batch;
    __shared__ int32_t s_loaded;
    if (threadIdx.x == 0) s_loaded = -1;

    DfaLds dl;
    dl.t1 = lds;                                              // 4 KB
    dl.t2 = lds + 4096;                                       // 16 KB
    dl.f2 = reinterpret_cast<const uint64_t*>(lds + 20480);   // 8 KB
    dl.acc = lds + 28672;                                     // 16 B
    const void* ntab = lds;
    const uint8_t* ncls = SLICE ? nullptr : lds + kNfaClasses * 2048;

    while (true) {
        __syncthreads();   // the previous batch is finished with LDS and s_batch
        if (threadIdx.x == 0) s_batch = (int32_t)atomicAdd(A.ticket, 1u);
        __syncthreads();
        const int32_t b = s_batch;
        if (b >= A.n_batches) break;
        const int32_t pat = A.bpattern[b];
        const uint32_t role = (uint32_t)A.brole[b];
        const MixPattern& P = A.patterns[pat];
        if (pat != s_loaded) {
            if (P.kind == 0) {
                stage_table(lds, P.t1, 4096);
                stage_table(lds + 4096, P.t2, 16384);
                stage_table(lds + 20480, P.f2, 8192);
                stage_table(lds + 28672, P.acc, 16);
            } else if constexpr (SLICE) {
                stage_table(lds, P.ntab, (int)kTab6Bytes);
            } else {
                stage_table(lds, P.ntab, kNfaClasses * 2048);
                stage_table(lds + kNfaClasses * 2048, P.ncls, 256);
            }
            __syncthreads();
            if (threadIdx.x == 0) s_loaded = pat;
        }
        const int32_t count = A.bcount[b];
        const bool active = (int32_t)threadIdx.x < count;
        const int64_t item = (int64_t)A.bfirst[b] + threadIdx.x;

        if (role == kRoleDfaChunk) {
            // every lane takes part: the look-back publishes before it waits
            dfa_chunk(A, P, dl, item, active);
            continue;
        }
        if (!active) continue;
        if (role == kRoleDfaWhole) {
            const uint32_t s = dfa_walk(A.bytes, A.ibeg[item], A.iend[item], dl, P);
            A.out_mask[A.ib[item]] = s;
            A.out_accept[A.ib[item]] = dl.acc[s];
        } else if (role == kRoleNfaWhole) {
            const uint64_t s = nfa_walk<SLICE>(A.bytes, A.ibeg[item], A.iend[item], P.nstart, ntab, ncls, P.win4);
            A.out_mask[A.ib[item]] = s;
            A.out_accept[A.ib[item]] = (s & P.naccept) != 0 ? 1 : 0;
        } else if (role == kRoleNfaSpec) {
            const bool first = (A.iflags[item] & 1u) != 0;
            const int32_t o = A.ic[item];
            bool met;
            const uint64_t s = nfa_walk_ck<SLICE, true, false>(A.bytes, A.ibeg[item], A.iend[item],
                                                            first ? P.nstart : ~0ull, ntab, ncls, P.win4, A.ck,
                                                            A.ckpt + (int64_t)o * A.n_ck, nullptr, A.n_ck, met);
            A.spec_exit[o] = s;
            __threadfence();
            atomicAdd(&A.done[A.ia[item]], 1u);
        } else if (role == kRoleNfaFixup) {
            const int64_t c0 = A.ibeg[item], c1 = A.iend[item];
            const uint32_t need = (uint32_t)(c1 - c0);
            BoundedWait wait(A.max_spins, A.max_ticks);
            bool ok = true;
            while (load_relaxed(&A.done[A.ia[item]]) < need) {
                if (wait.expired()) {
                    atomicOr(A.error, 4u);
                    ok = false;
                    break;
                }
                __builtin_amdgcn_s_sleep(8);
            }
            fence_acquire();   // checkpoints and exits were written before each chunk counted itself
            uint64_t s = 0;
            if (ok) {
                s = A.spec_exit[A.ic[c0]];   // the first chunk was walked from the start set
                // the empty set is absorbing: once the true set is empty, so are all later exits
                for (int64_t c = c0 + 1; c < c1 && s != 0; ++c) {
                    const int32_t o = A.ic[c];
                    bool met;
                    const uint64_t t = nfa_walk_ck<SLICE, false, true>(A.bytes, A.ibeg[c], A.iend[c], s, ntab, ncls,
                                                                    P.win4, A.ck, nullptr,
                                                                    A.ckpt + (int64_t)o * A.n_ck, A.n_ck, met);
                    s = met ? A.spec_exit[o] : t;
                    if (met) atomicAdd(A.converged, 1u);
                }
            }
            A.out_mask[A.ib[item]] = s;
            A.out_accept[A.ib[item]] = (s & P.naccept) != 0 ? 1 : 0;
        }
    }
}

}  // namespace

// Nibble NFA tables: two 512-thread workgroups per CU (70 KB of LDS each).
extern "C" __global__ void __launch_bounds__(MIX_BLOCK) __attribute__((amdgpu_waves_per_eu(4)))
mix_persistent(MixArgs A)
{
    mix_loop<false>(A);
}

// 6/5-bit slice NFA tables (12 reads per byte instead of 16): one
// 1024-thread workgroup per CU (128 KB of LDS).
extern "C" __global__ void __launch_bounds__(MIX_BLOCK_SLICE) __attribute__((amdgpu_waves_per_eu(4)))
mix_persistent_slice(MixArgs A)
{
    mix_loop<true>(A);
}
