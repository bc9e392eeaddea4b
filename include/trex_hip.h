/*
 * trex_hip.h -- C ABI of libtrexhip.so, the MI355X (gfx950) automaton-matching
 * engine.
 *
 * Two layers of entry points:
 *
 *  1. Reference-compatible entry points with the exact names, argument lists
 *     and return conventions of the reference's ctypes-bound C ABI, so its
 *     Python bridges can load this library instead of build/lib*.so. Each
 *     cites the reference declaration it replaces. Like the reference, they
 *     keep one process-global engine per family.
 *
 *  2. Handle-based trex_* entry points used by the trex_amd Python package:
 *     many resident patterns, device-pointer dispatch on caller streams,
 *     int64 CSR offsets, per-string final state + accept bit.
 *
 * Conventions: every function returns 0 (TREX_OK) on success or a negative
 * TREX_ERR_* code; none of them exits the process. trex_last_error() copies
 * the message of the calling thread's last failure. A string batch is CSR:
 * bytes[offsets[i] .. offsets[i+1]) is string i, offsets has n_strings + 1
 * entries. "d_" arguments are device pointers, other pointers are host memory.
 * `stream` is a hipStream_t (NULL = the default stream).
 */
#ifndef TREX_HIP_H
#define TREX_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TREX_OK 0
#define TREX_ERR_NO_DEVICE (-1)  /* same meaning as the reference's -1 */
#define TREX_ERR_ARCH (-2)       /* same meaning as the reference's -2 */
#define TREX_ERR_INVALID (-3)
#define TREX_ERR_CAPACITY (-4)
#define TREX_ERR_HIP (-5)
#define TREX_ERR_NOT_INIT (-6)
#define TREX_ERR_TIMEOUT (-7)

/* ------------------------------------------------------------------------
 * Device and errors
 * ------------------------------------------------------------------------ */

/* 0 if a gfx950 device is present; TREX_ERR_NO_DEVICE / TREX_ERR_ARCH otherwise.
 * Replaces monoid_batch_engine_device_check (cuda/monoid_batch.cu:601-609). */
int trex_device_check(void);

/* Copies the calling thread's last error message into buf[0..n); returns its length. */
int trex_last_error(char* buf, int n);

/* Time bound, in microseconds of device wall-clock time, of every cross-wave
 * wait in the launches that follow (look-back plans and mixed queues); 0
 * restores the default of 100 ms. Process-wide. A wait that expires sets its
 * launch's error word and the launch still ends. No reference counterpart: the
 * reference spins without a bound (cuda/monoid_scan.cu:188-192). */
int trex_set_wait_bound_us(uint32_t us);

/* Architecture name (e.g. gfx950:sramecc+:xnack-), compute units and HBM bytes of the current device. */
int trex_device_info(char* name, int n, int* cu_count, int64_t* hbm_bytes);

/* ------------------------------------------------------------------------
 * Memory and stream plumbing (no reference counterpart: the reference copies
 * synchronously inside each dispatch, cuda/monoid_batch.cu:504-553).
 * ------------------------------------------------------------------------ */
int trex_malloc(void** d_ptr, size_t bytes);
int trex_free(void* d_ptr);
int trex_host_alloc(void** h_ptr, size_t bytes); /* pinned host memory */
int trex_host_free(void* h_ptr);
int trex_memcpy_h2d(void* d_dst, const void* src, size_t n, void* stream);
int trex_memcpy_d2h(void* dst, const void* d_src, size_t n, void* stream);
int trex_memcpy_d2d(void* d_dst, const void* d_src, size_t n, void* stream);
/* n bytes copied by a kernel on `stream` rather than a copy engine, so that a
 * short copy does not queue behind long ones on the engines. dst: device
 * memory; src: device memory or pinned host memory (trex_host_alloc); both
 * 16-byte aligned. No reference counterpart (its string uploads are
 * synchronous cudaMemcpy calls, cuda/monoid_batch.cu:506-507). */
int trex_memcpy_kernel(void* d_dst, const void* src, int64_t n, void* stream);
int trex_stream_create(void** stream);
int trex_stream_destroy(void* stream);
int trex_stream_sync(void* stream);
int trex_device_sync(void);
int trex_event_create(void** ev);
int trex_event_destroy(void* ev);
int trex_event_record(void* ev, void* stream);
int trex_event_elapsed_ms(void* start, void* stop, float* ms);
/* 0 when every command before the event has completed, 1 when not yet, < 0 on error. */
int trex_event_query(void* ev);
int trex_event_sync(void* ev);
/* Later work on `stream` waits for `ev` (no host wait). */
int trex_stream_wait_event(void* stream, void* ev);

/* HIP graphs of captured launches (no reference counterpart: it launches each
 * kernel from the host, cuda/monoid_batch.cu:647-658). trex_graph_begin starts
 * capturing `stream` (not the default stream); every asynchronous trex_* call
 * on it until trex_graph_end is recorded instead of run, and trex_graph_end
 * instantiates the recording. trex_graph_launch replays it on any stream with
 * the same device pointers and parameters. Entry points that copy from host
 * memory, synchronise or allocate (the *_host forms, the first call of a
 * look-back plan or mixed queue that sizes its buffers) cannot be captured:
 * run them once before capturing, or use trex_dfa_match_chunked_static, which
 * allocates nothing. trex_graph_abort ends a capture that failed part way and
 * drops what it recorded. */
typedef struct trex_graph trex_graph;
int trex_graph_begin(void* stream);
int trex_graph_end(void* stream, trex_graph** out);
int trex_graph_abort(void* stream);
int trex_graph_launch(trex_graph* g, void* stream);
int trex_graph_destroy(trex_graph* g);

/* Synthetic input in HBM (no reference counterpart): bytes [0, n_bytes) of
 * d_out (16-byte aligned, n_bytes a multiple of 16) become lo + (r & (span-1)),
 * r = byte (i % 8) of splitmix64(seed ^ ((i / 8) * 0x9E3779B97F4A7C15)). */
int trex_gen_symbols(void* d_out, int64_t n_bytes, uint64_t seed, uint32_t lo, uint32_t span, void* stream);

/* ------------------------------------------------------------------------
 * DFA byte-table engine (hot path (a)).
 * Replaces the engine behind monoid_batch_engine_init/dispatch
 * (cuda/monoid_batch.cu:611-658) and prefix_engine_init/dispatch
 * (cuda/prefix_compose.cu:425-470) with a handle per resident pattern.
 * ------------------------------------------------------------------------ */
typedef struct trex_dfa trex_dfa;

/* delta: [n_states][256] uint8 successor table, accept: [n_states] (0/1),
 * 1 <= n_states <= 256, start < n_states. The table is taken as given (it
 * fixes what bytes outside the alphabet do). Tables with <= 16 states use the
 * DFA-16 kernels; a 16-state table whose live bytes lie in one aligned
 * 32-byte window uses the two-bytes-per-lookup kernel. */
int trex_dfa_create(const uint8_t* delta, const uint8_t* accept, int n_states, int start, trex_dfa** out);
int trex_dfa_destroy(trex_dfa* h);

#define TREX_DFA_AUTO 0
#define TREX_DFA_K1_GENERIC 1 /* any table, one LDS lookup per byte */
#define TREX_DFA_K1_16 2      /* <= 16 states, one lookup per byte */
#define TREX_DFA_K2_WINDOW 3   /* <= 16 states, window bytes, one lookup per byte pair */
#define TREX_DFA_K2_WINDOW_F 5 /* as 3, pair table of 64-bit nibble functions */
#define TREX_DFA_K2_WINDOW_FR 6 /* as 5, four interleaved copies of the table in LDS (fewer bank conflicts) */

/* Reports the kernel variant chosen at create time, the window base byte
 * (-1 if none) and the dead state (-1 if none). */
int trex_dfa_info(const trex_dfa* h, int* variant, int* window_base, int* dead_state);
/* Forces a variant (TREX_DFA_AUTO restores the default); fails if the table
 * does not allow it. */
int trex_dfa_set_variant(trex_dfa* h, int variant);

/* 16-byte vectors each lane keeps in flight while streaming a string (8 or 16). */
int trex_dfa_set_prefetch(trex_dfa* h, int vectors);

/* Read order of contiguous strings (trex_dfa_match / _ranges with the
 * two-bytes-per-lookup function table, 8 vectors in flight): with rotation,
 * the lanes of wave w read each string's 16-byte vectors from
 * (16 x (w mod 16)) mod n onwards, then the first ones, composing the
 * transition function of the part read first until it turns constant. Equal
 * strings packed back to back are then not all read at the same offset
 * modulo 4 KiB at once, which crowds a few HBM channels. Same results.
 * mode -1 (default): on when the table synchronises (some word over its
 * window bytes sends every state that can occur to one state); 0 off; 1 on.
 * No counterpart in the reference (its kernels read strings as packed,
 * cuda/monoid_batch.cu:58-95). trex_dfa_rotation reports whether launches
 * use it and whether the table synchronises. */
int trex_dfa_set_rotation(trex_dfa* h, int mode);
int trex_dfa_rotation(const trex_dfa* h, int* active, int* synchronizes);

/* Device-pointer dispatch, asynchronous on `stream`: final state and accept
 * bit per string. d_offsets: int64[n_strings + 1]. */
int trex_dfa_match(trex_dfa* h, const uint8_t* d_bytes, const int64_t* d_offsets, int64_t n_strings,
                   uint8_t* d_final_state, uint8_t* d_accept, void* stream);

/* Same, with string i = d_bytes[d_begin[i] .. d_end[i]) (ranges may leave
 * gaps, as in the staggered HBM layout). */
int trex_dfa_match_ranges(trex_dfa* h, const uint8_t* d_bytes, const int64_t* d_begin, const int64_t* d_end,
                          int64_t n_strings, uint8_t* d_final_state, uint8_t* d_accept, void* stream);

/* Same over the interleaved layout (no reference counterpart; it replaces
 * the packed CSR upload of cuda/monoid_batch.cu:506-507 for uniform batches):
 * the strings of a group of 64 consecutive strings are stored as rows of
 * 64 x 16 bytes, vector k of the group's lane j at group base + (64 k + j) * 16.
 * d_begin[i] = address of string i's vector 0 (16-byte aligned),
 * d_end[i] - d_begin[i] = its length; its last len % 16 bytes open vector
 * len / 16. One wave load then reads one contiguous 1 KiB row. */
int trex_dfa_match_interleaved(trex_dfa* h, const uint8_t* d_bytes, const int64_t* d_begin, const int64_t* d_end,
                               int64_t n_strings, uint8_t* d_final_state, uint8_t* d_accept, void* stream);

/* Host-pointer dispatch (synchronous): stages the batch through engine-owned
 * device buffers. kernel_ms / total_ms (nullable) receive the kernel time and
 * the copy+kernel+copy time, like the reference's dispatch outputs. */
int trex_dfa_match_host(trex_dfa* h, const uint8_t* bytes, const int64_t* offsets, int64_t n_strings,
                        uint8_t* final_state, uint8_t* accept, float* kernel_ms, float* total_ms);

/* ------------------------------------------------------------------------
 * Several DFA-16 patterns in one pass over the bytes. The reference packs
 * several DFAs block-diagonally into one state vector and evolves it through
 * its batched matrix kernels (PackedEngine, src/packed_engine.py:25-126,
 * batched_engine_dispatch_multi); these entry points replace that with
 * per-pattern pair tables in LDS. Each launch reads the batch once and advances up to `group`
 * patterns, so matching P patterns reads the bytes ceil(P / group) times
 * instead of P.
 * ------------------------------------------------------------------------ */
typedef struct trex_dfa_multi trex_dfa_multi;

/* patterns: n_patterns (1..256) handles of DFA-16 tables whose live bytes lie
 * in one aligned 32-byte window each (trex_dfa_info reports window_base >= 0;
 * the windows may differ). The tables are copied: the handles may be
 * destroyed afterwards. Any other table is refused (TREX_ERR_INVALID, naming
 * the pattern). */
int trex_dfa_multi_create(trex_dfa* const* patterns, int n_patterns, trex_dfa_multi** out);
int trex_dfa_multi_destroy(trex_dfa_multi* m);
/* Table form: TREX_DFA_MULTI_K2F_PAIRED (the pair functions of two
 * patterns side by side, one 16-byte LDS read for both; rebuilt when the
 * group changes), TREX_DFA_MULTI_K2F_COMPACT (patterns of at most 4 states
 * as one byte per pair entry, eight to a 64-bit word, beside the others'
 * full functions: fewer table reads per byte pair; rebuilt when the group
 * changes), TREX_DFA_K2_WINDOW_F (8 KB of LDS per pattern) or
 * TREX_DFA_K2_WINDOW (16 KB per pattern). TREX_DFA_AUTO (and a new handle):
 * compact when at least two patterns have at most 4 states, else paired. */
#define TREX_DFA_MULTI_K2F_PAIRED 7
#define TREX_DFA_MULTI_K2F_COMPACT 8
int trex_dfa_multi_set_form(trex_dfa_multi* m, int form);
/* Patterns advanced per launch at most, 1..8 (a new handle: 8 when it
 * starts in the compact form, else 4); the patterns are split evenly over
 * ceil(n_patterns / group) launches. */
int trex_dfa_multi_set_group(trex_dfa_multi* m, int group);
int trex_dfa_multi_info(const trex_dfa_multi* m, int* n_patterns, int* form, int* group);
/* Final state and accept bit of every (pattern, string), pattern-major:
 * d_final_state[p * n_strings + i], d_accept[p * n_strings + i] (each
 * n_patterns * n_strings bytes). Strings as in trex_dfa_match_ranges /
 * trex_dfa_match_interleaved; results equal those of each pattern's own
 * handle. Asynchronous on `stream`. */
int trex_dfa_multi_match_ranges(trex_dfa_multi* m, const uint8_t* d_bytes, const int64_t* d_begin,
                                const int64_t* d_end, int64_t n_strings, uint8_t* d_final_state, uint8_t* d_accept,
                                void* stream);
int trex_dfa_multi_match_interleaved(trex_dfa_multi* m, const uint8_t* d_bytes, const int64_t* d_begin,
                                     const int64_t* d_end, int64_t n_strings, uint8_t* d_final_state,
                                     uint8_t* d_accept, void* stream);

/* ------------------------------------------------------------------------
 * Boolean NFA-64, bit-parallel (hot path (c)).
 * Replaces nfa_tc_engine_init/dispatch with the V4 bit-parallel kernel
 * (cuda/nfa_tc_evolution.cu:683-745, 943-1013), which handles two symbols
 * only; this engine handles up to 62 symbols with nibble-sliced successor
 * tables in LDS.
 * ------------------------------------------------------------------------ */
typedef struct trex_nfa trex_nfa;

#define TREX_NFA_BYTES 1   /* any byte -> class map */
#define TREX_NFA_WINDOW 2  /* every byte outside one aligned 32-byte window rejects */
#define TREX_NFA_WINDOW5 3 /* as 2, 5-bit state slices: 13 table reads per byte instead of 16 */
#define TREX_NFA_WINDOW6 4 /* as 2, 6/5-bit state slices: 12 table reads per byte */

/* succ: [n_symbols][64] uint64, bit d of succ[c][s] = edge s -> d on symbol c
 * (epsilon closure already folded in, as in NFAMatrices). byte_to_sym[256]:
 * symbol index, -1 = reject (the state set becomes empty), -2 = skip (the byte
 * leaves the set unchanged). start / accept: state masks. n_symbols <= 62. */
int trex_nfa_create(const uint64_t* succ, int n_symbols, const int16_t* byte_to_sym, uint64_t start,
                    uint64_t accept, trex_nfa** out);
int trex_nfa_destroy(trex_nfa* h);
/* The variant chosen at create time: TREX_NFA_WINDOW6 when every byte outside
 * one aligned 32-byte window rejects, else TREX_NFA_BYTES; set_variant forces
 * another the automaton allows. */
int trex_nfa_info(const trex_nfa* h, int* variant, int* window_base);
int trex_nfa_set_variant(trex_nfa* h, int variant);
/* Final state mask and accept bit ((mask & accept) != 0) per string. */
int trex_nfa_match(trex_nfa* h, const uint8_t* d_bytes, const int64_t* d_offsets, int64_t n_strings,
                   uint64_t* d_final_mask, uint8_t* d_accept, void* stream);
int trex_nfa_match_ranges(trex_nfa* h, const uint8_t* d_bytes, const int64_t* d_begin, const int64_t* d_end,
                          int64_t n_strings, uint64_t* d_final_mask, uint8_t* d_accept, void* stream);
int trex_nfa_match_host(trex_nfa* h, const uint8_t* bytes, const int64_t* offsets, int64_t n_strings,
                        uint64_t* final_mask, uint8_t* accept, float* kernel_ms, float* total_ms);

/* cuda/nfa_tc_evolution.cu:943-953. T_matrices: [sigma][n][n] row-major
 * [dst][src], entries > 0.5 are edges; n <= 64. Symbols >= sigma in the
 * dispatched input are skipped, as in the reference's simulate_nfa_sequential
 * (:1029-1054). */
int nfa_tc_engine_init(const float* T_matrices, const float* accept_mask, const float* start_vec, int n, int sigma,
                       int max_B, int max_L);
/* cuda/nfa_tc_evolution.cu:955-1009. raw_concat holds symbol indices. */
int nfa_tc_engine_dispatch(const uint8_t* raw_concat, const int* offsets, int* results, int B, int total_chars,
                           float* kernel_ms, float* total_ms);
/* cuda/nfa_tc_evolution.cu:1011-1013 */
void nfa_tc_engine_destroy(void);

/* ------------------------------------------------------------------------
 * Probabilistic NFA-64 on MFMA.
 * Replaces the V5 probabilistic evolution nfa_prob_v5_tc_kernel and its
 * launch through nfa_tc_engine_init/dispatch (cuda/nfa_tc_evolution.cu:451-561,
 * 880-895), which runs s' = T[c] s in FP16 with an FP16 accumulator, no
 * rescaling, and returns the bit (a . s_L) > 0.5. Here every sequence gets
 *   LL      = log(a . T[c_L] ... T[c_1] s_0)
 *   logmass = log(1 . T[c_L] ... T[c_1] s_0)
 * in FP64, from FP16/BF16 MFMA products with FP32 accumulation, a power-of-two
 * rescale per step into an integer log2-scale, and the last step recomputed in
 * FP32 from the unrounded T. For a nonnegative model every product's mass is
 * checked against sum_k colsum(round T)_k x_k computed on the vector ALUs and
 * restored (the matrix cores' FP32 accumulation drifts one way). A symbol
 * >= n_symbols empties the state (LL = -inf).
 * ------------------------------------------------------------------------ */
typedef struct trex_pnfa trex_pnfa;

#define TREX_PNFA_F16X2 1   /* FP16 T, state split into two FP16 halves (two MFMAs) */
#define TREX_PNFA_BF16X2 2  /* BF16 T, state split into two BF16 halves */
#define TREX_PNFA_F16 3     /* FP16 T and state (one MFMA) */
#define TREX_PNFA_BF16 4    /* BF16 T and state */

/* T: [n_symbols][64][64] FP32 row-major [dst][src] (the reference's layout);
 * start, accept: [64]. n_symbols in 1..32. The default mode is F16X2, or
 * BF16X2 when some |T| exceeds the FP16 range. */
int trex_pnfa_create(const float* T, int n_symbols, const float* start, const float* accept, trex_pnfa** out);
int trex_pnfa_destroy(trex_pnfa* h);
int trex_pnfa_info(const trex_pnfa* h, int* mode, int* n_symbols, int* f16_ok, int* waves);
int trex_pnfa_set_mode(trex_pnfa* h, int mode);
/* Waves per 64-sequence workgroup: 4 (eight symbols' matrices per wave) or 8
 * (four per wave, two waves per SIMD). */
int trex_pnfa_set_waves(trex_pnfa* h, int waves);
/* Most workgroups per kernel launch (0, the default: one per group of 64
 * sequences). With fewer, the groups go out in consecutive launches on the
 * stream, which bounds the compute units a long call holds while other work
 * shares the device; put long sequences first. No counterpart in the
 * reference (it has no probabilistic kernel). */
int trex_pnfa_set_grid(trex_pnfa* h, int64_t workgroups);
/* The operator a mode applies on every step but the last: the rounded T as FP32
 * [n_symbols][64][64] and the column mass factors [n_symbols][64] (FP64: the
 * kernels apply each as an FP32 high part plus its FP32 rest), so that
 * s' = T_rounded (col_scale[c] * s). Either output may be null. */
int trex_pnfa_rounded(const trex_pnfa* h, int mode, float* t_rounded, double* col_scale);
/* d_symbols: [n_seq][length] uint8 symbol indices; d_ll, d_logmass: [n_seq]
 * FP64, only written (pinned host memory works too). Asynchronous on
 * `stream`. length <= 2^31 - 1. */
int trex_pnfa_loglik(trex_pnfa* h, const uint8_t* d_symbols, int64_t n_seq, int64_t length, double* d_ll,
                     double* d_logmass, void* stream);
/* Ragged sequences: sequence j = d_symbols[d_begin[j] .. d_end[j]) (lengths
 * up to 2^31 - 1; sequences of similar length in the same group of 64 waste
 * fewer steps, so callers sort by length). */
int trex_pnfa_loglik_ranges(trex_pnfa* h, const uint8_t* d_symbols, const int64_t* d_begin, const int64_t* d_end,
                            int64_t n_seq, double* d_ll, double* d_logmass, void* stream);
/* Host-pointer form (synchronous), with kernel and end-to-end times. */
/* Parallel in time, for long sequences: sequence j = d_symbols[begin[j] ..
 * end[j]) with begin / end in HOST memory (the chunk table is built on the
 * host). Each sequence's rounded steps are cut into chunks of chunk_len
 * symbols (0: about four chunks per compute unit over the call, at least 256
 * symbols each: the error grows with the number of chunks) whose
 * 64 x 64 products run concurrently on MFMA, then chained per sequence in
 * FP64; the same operator and last FP32 step as trex_pnfa_loglik_ranges.
 * About 64 times the arithmetic of the batch kernel, in exchange for a
 * latency set by chunk_len rather than by the sequence length. Workspace
 * (16.3 KB per chunk) is allocated and freed in stream order on `stream`.
 * No reference counterpart. */
int trex_pnfa_loglik_chunked(trex_pnfa* h, const uint8_t* d_symbols, const int64_t* begin, const int64_t* end,
                             int64_t n_seq, int64_t chunk_len, double* d_ll, double* d_logmass, void* stream);
int trex_pnfa_loglik_host(trex_pnfa* h, const uint8_t* symbols, int64_t n_seq, int64_t length, double* ll,
                          double* logmass, float* kernel_ms, float* total_ms);

/* ------------------------------------------------------------------------
 * Mixed batches: many DFA-16 and Boolean NFA-64 patterns and strings of
 * uneven length in ONE persistent launch fed by a length-sorted work queue.
 * Replaces the reference's one-engine-per-dispatch model (one global engine
 * per library and one synchronous launch per batch and pattern,
 * cuda/monoid_batch.cu:596-671, cuda/nfa_tc_evolution.cu:943-1013); long DFA
 * strings are chained by the decoupled look-back of monoid_r3_kernel
 * (cuda/monoid_scan.cu:369-451), long NFA strings by speculative chunks with
 * checkpoints and an exact fix-up pass.
 * ------------------------------------------------------------------------ */
typedef struct trex_mix trex_mix;

#define TREX_MIX_DFA 0
#define TREX_MIX_NFA 1

/* kinds[i]: TREX_MIX_DFA (handles[i] is a trex_dfa* of <= 16 states with a
 * 32-byte symbol window) or TREX_MIX_NFA (a trex_nfa* with a window).
 * handles may be NULL: the object then only plans (no device needed). */
int trex_mix_create(const int32_t* kinds, const void* const* handles, int32_t n_patterns, trex_mix** out);
int trex_mix_destroy(trex_mix* m);
/* String i = bytes[begin[i] .. end[i]) is matched against pattern
 * pattern_of[i]. Strings longer than dfa_split / nfa_split bytes are cut into
 * chunks of dfa_chunk (multiple of 16) / nfa_chunk (multiple of ck) bytes;
 * NFA chunks record their state set every ck bytes (ck multiple of 16).
 * Host work only: the plan is staged in pinned memory and copied to the
 * device, asynchronously and by a copy kernel, by the next trex_mix_run or
 * trex_mix_capture on that call's stream (device buffers grow there too, in
 * stream order). A handle holds one plan on the device: runs of one handle on
 * different streams must be ordered by the caller. */
int trex_mix_plan(trex_mix* m, const int64_t* begin, const int64_t* end, const int32_t* pattern_of, int64_t n_strings,
                  int64_t dfa_split, int64_t dfa_chunk, int64_t nfa_split, int64_t nfa_chunk, int32_t ck);
int trex_mix_info(const trex_mix* m, int64_t* n_items, int64_t* n_batches, int64_t* n_chunks, int64_t* n_fixups);
/* Copies the plan (items, then batches in queue order); any pointer may be NULL. */
int trex_mix_layout(const trex_mix* m, int64_t* ibeg, int64_t* iend, int32_t* ia, int32_t* ib, uint32_t* iflags,
                    int32_t* bfirst, int32_t* bcount, int32_t* bpattern, int32_t* brole);
/* One launch sequence on `stream` (the plan's copy after a new plan, a fill
 * kernel for the queue counters, then the persistent kernel). The outputs are
 * only written, so they may be pinned host memory (trex_host_alloc): the
 * results then need no copy back. d_out_mask[i] = DFA final state or NFA final set,
 * d_out_accept[i] = accept bit, for every planned string. Every wait of the
 * launch (a DFA chunk on an earlier chunk of another wave, a fix-up on its
 * string's speculative chunks) ends after max_spins polls (0: no count limit)
 * or 100 ms of device wall-clock time, whichever comes first; on expiry the
 * launch still ends and trex_mix_error reports it. */
int trex_mix_run(trex_mix* m, const uint8_t* d_bytes, uint64_t* d_out_mask, uint8_t* d_out_accept, uint32_t max_spins,
                 void* stream);
/* Persistent workgroups of later runs (0: two per compute unit). Fewer leave
 * compute units to kernels on other streams. */
int trex_mix_set_grid(trex_mix* m, int64_t workgroups);
/* NFA table form of the launch: TREX_MIX_FORM_NIBBLE (default; 512-thread
 * workgroups, two per CU) or TREX_MIX_FORM_SLICE (6/5-bit state slices, 12 LDS
 * reads per NFA byte instead of 16; 1024-thread workgroups, one per CU; every
 * NFA pattern must have a window). Batches are cut to the workgroup size, so
 * call it before trex_mix_plan; a run whose plan was made for the other form
 * fails. Results are the same in both forms. */
#define TREX_MIX_FORM_NIBBLE 0
#define TREX_MIX_FORM_SLICE 1
int trex_mix_set_form(trex_mix* m, int form);
/* Enqueues on `stream`, behind the last run, a copy of the run's first four
 * queue words (ticket, error word, converged count, one flag) to dst:
 * 16 bytes, 16-byte aligned, device or pinned host memory, copied by a kernel
 * so the check never waits behind long copies on the copy engines. With no
 * planned work nothing is enqueued. No reference counterpart. */
int trex_mix_status(trex_mix* m, void* dst, void* stream);
/* Error word of the last run (0: no wait expired). Synchronises `stream`. */
int trex_mix_error(trex_mix* m, uint32_t* error, void* stream);
/* Speculative NFA chunks of the last run whose fix-up met a checkpoint (the
 * rest were walked to their end). Synchronises `stream`. */
int trex_mix_stats(trex_mix* m, uint32_t* converged, void* stream);
/* Captures trex_mix_run into a HIP graph (non-default stream), replayed by
 * trex_mix_replay with the same buffers. */
int trex_mix_capture(trex_mix* m, const uint8_t* d_bytes, uint64_t* d_out_mask, uint8_t* d_out_accept,
                     uint32_t max_spins, void* stream);
int trex_mix_replay(trex_mix* m, void* stream);

/* ------------------------------------------------------------------------
 * Long strings: chunk plans and the decoupled look-back scan (hot path (b)).
 * Replaces monoid_engine_dispatch_batch / dispatch_adaptive with the
 * monoid_r3_kernel look-back (cuda/monoid_scan.cu:369-451, 468-504) and
 * monoid_batch_engine_dispatch_prefix (cuda/monoid_batch.cu:660-671): one
 * launch covers a whole batch, every string split into chunks of chunk_len
 * bytes whose results are chained by a decoupled look-back over DFA functions.
 * ------------------------------------------------------------------------ */
typedef struct trex_plan trex_plan;

/* begin/end: host int64 byte ranges of n_strings strings (in the device byte
 * buffer the plan will be used with). chunk_len: multiple of 16, >= 64; the
 * launch is fastest with about 65,536 chunks in all (one wave of 64 chunk
 * lanes per SIMD), e.g. 16 KiB for 1 GiB of long strings. */
int trex_plan_create(const int64_t* begin, const int64_t* end, int64_t n_strings, int64_t chunk_len,
                     trex_plan** out);
int trex_plan_destroy(trex_plan* plan);
int trex_plan_info(const trex_plan* plan, int64_t* n_chunks, int64_t* n_strings);
/* Final state and accept bit of every string of the plan (DFA of <= 16
 * states). A plan holds the look-back state of one launch: launches that use
 * the same plan must be ordered on one stream. A chunk whose predecessor is
 * in its own wave takes it by shuffle; any other waits for its predecessor's
 * publication, at most max_spins polls (0: no count limit) and at most 100 ms
 * of device wall-clock time; on expiry the launch still ends and
 * trex_plan_error reports it. */
int trex_dfa_match_plan(trex_dfa* h, trex_plan* plan, const uint8_t* d_bytes, uint8_t* d_final_state,
                        uint8_t* d_accept, uint32_t max_spins, void* stream);

/* The plan's chunk-interleaved image: the same bytes laid out so that one
 * wave load reads 1 KiB contiguously. Tile t holds the plan's chunks 64t ..
 * 64t+63 (processing order); vector i (16 bytes) of the tile's k-th chunk
 * lies at tile_off[t] + (64 i + k) * 16, a chunk's last partial vector
 * zero-padded, each tile as long as its longest chunk and followed by a gap
 * (256 bytes by default, trex_plan_set_image_gap) so that the tiles the
 * resident waves stream at once start on different HBM channels. The image form of
 * the HBM layout for long strings, as interleaved rows are for batches of
 * similar lengths (same matching as monoid_r3_kernel, cuda/monoid_scan.cu:
 * 128-222; no counterpart of the layout in the reference).
 * trex_plan_image_size: bytes of the image. trex_plan_build_image: fill it on
 * the host from `src`, the host copy of what the plan's ranges index
 * (src_len bytes); dst holds at least the image size. */
int trex_plan_image_size(const trex_plan* plan, int64_t* bytes);
int trex_plan_build_image(const trex_plan* plan, const uint8_t* src, int64_t src_len, uint8_t* dst,
                          int64_t dst_len);
/* trex_dfa_match_plan over the plan's image in device memory (d_image, built
 * by trex_plan_build_image for this plan and uploaded; image_bytes = the
 * buffer's size, refused when smaller than trex_plan_image_size). */
int trex_dfa_match_plan_image(trex_dfa* h, trex_plan* plan, const uint8_t* d_image, int64_t image_bytes,
                              uint8_t* d_final_state, uint8_t* d_accept, uint32_t max_spins, void* stream);
/* Phase 1 of the plan's later launches with 1 or 2 dependent chains per lane:
 * 2 walks the two halves of each chunk side by side (the second half from
 * the identity function) and joins them with one composition. Same results;
 * a tuning choice. No counterpart in the reference. */
int trex_plan_set_chains(trex_plan* plan, int chains);
/* Bytes left after each tile of the plan's image (a multiple of 16, at most
 * 65536; 256 by default). Changes trex_plan_image_size and the layout of
 * images built afterwards; an image must be built under the gap it is
 * matched with. Waits for the device to finish queued work first (launches
 * in flight read the old layout). A tuning choice with no counterpart in the
 * reference. */
int trex_plan_set_image_gap(trex_plan* plan, int64_t gap);
/* Error word of the plan's last launch (0 = none; bit 0 = a wait expired),
 * read synchronously after the work queued on `stream`. */
int trex_plan_error(trex_plan* plan, uint32_t* error, void* stream);
/* Test hook, no counterpart in the reference (its tests ask for the
 * look-back under an adversarial tile order, which in-order tickets rule
 * out): later launches of the plan hold back each tile's publication by
 * hold x (7 - t mod 8) sleeps of 512 clocks, so tiles publish in reverse
 * order within every group of 8 and successors wait for and walk back over
 * unpublished predecessors (at most about 3 ms per tile). 0 (the default) =
 * off; at most 2000. */
int trex_plan_set_hold(trex_plan* plan, uint32_t hold);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): trex_plan_create + trex_dfa_match_plan in one call that
 * allocates nothing. The string ranges are in DEVICE memory (d_begin/d_end:
 * n_strings int64 offsets into d_bytes, which holds bytes_len bytes); the
 * chunk table (chunks of chunk_len bytes, cut as trex_plan_create cuts them)
 * is built on the device, and it and the look-back state live in a static
 * device array of TREX_STATIC_SCRATCH_BYTES that each GPU the library runs on
 * holds until the process ends. Limits, refused with TREX_ERR_CAPACITY: at
 * most TREX_STATIC_MAX_STRINGS strings, and n_strings + ceil(bytes_len /
 * chunk_len) at most TREX_STATIC_MAX_CHUNKS. All calls on one GPU share its
 * array: calls that overlap on one GPU race. d_error: one device word, set by
 * the call to 0 or to bits: 1 a look-back wait expired, 2 a chain without a
 * first chunk, 4 a range with begin < 0, end < begin or end > bytes_len, 8
 * more chunks than the bound (overlapping ranges); with bit 4 or 8 set no
 * string is matched and the outputs are left unwritten. Asynchronous on
 * `stream`. */
#define TREX_STATIC_MAX_STRINGS 65536
#define TREX_STATIC_MAX_CHUNKS 131072
#define TREX_STATIC_SCRATCH_BYTES 6291968
int trex_dfa_match_chunked_static(trex_dfa* h, const uint8_t* d_bytes, int64_t bytes_len, const int64_t* d_begin,
                                  const int64_t* d_end, int64_t n_strings, int64_t chunk_len,
                                  uint8_t* d_final_state, uint8_t* d_accept, uint32_t* d_error, uint32_t max_spins,
                                  void* stream);

/* Plain device copy with the engine's streaming-load kernel (16-byte aligned
 * buffers, size a multiple of 16): the reference point for HBM traffic
 * counters and the achievable-bandwidth probe. No reference counterpart. */
int trex_copy(void* d_dst, const void* d_src, size_t n_bytes, void* stream);

/* ------------------------------------------------------------------------
 * Reference-compatible DFA entry points (process-global engines).
 * ------------------------------------------------------------------------ */

/* cuda/prefix_compose.cu:425-433 */
int prefix_engine_device_check(void);
/* cuda/prefix_compose.cu:435-450. tmap[byte * N + s], unknown bytes as the
 * caller encodes them (the reference's tables use the identity). K is
 * ignored, as in the reference. */
int prefix_engine_init(const uint8_t* tmap, const uint8_t* accept, int start_state, int N, int K,
                       int max_total_chars, int max_batch);
/* cuda/prefix_compose.cu:452-463. results[i] = accept bit (0/1). Fails with
 * TREX_ERR_CAPACITY instead of overrunning when B or total_chars exceed the
 * init capacities. */
int prefix_engine_dispatch(const uint8_t* raw_concat, const int* offsets, int* results, int B,
                           int total_chars, float* kernel_ms, float* total_ms);
/* cuda/prefix_compose.cu:465-470 */
void prefix_engine_destroy(void);

/* cuda/monoid_batch.cu:601-609 */
int monoid_batch_engine_device_check(void);
/* cuda/monoid_batch.cu:611-628. The monoid index walk is run as a byte-table
 * walk over M states: delta[m][byte] = char_compose[m * sigma_ext + raw_char_map[byte]]. */
int monoid_batch_engine_init(int M, int sigma_ext, int identity, const uint8_t* char_compose,
                             const uint8_t* raw_char_map, const uint8_t* accept, const uint8_t* monoid_compose,
                             int max_total_chars, int max_batch);
/* cuda/monoid_batch.cu:630-638. Accepted for compatibility; the byte-table
 * walk needs no separate K-gram table. */
int monoid_batch_engine_set_kgram(const uint8_t* fused_compose, const uint8_t* kgram_compose, int kgram_k,
                                  int sigma_k, int sigma);
/* cuda/monoid_batch.cu:647-658 */
int monoid_batch_engine_dispatch(const uint8_t* raw_concat, const int* offsets, int* results, int B,
                                 int total_chars, float* kernel_ms, float* total_ms);
/* cuda/monoid_batch.cu:660-671 (long strings): with M <= 16 every string is
 * cut into chunks chained by the decoupled look-back (trex_dfa_match_plan, one
 * launch); wider monoids walk each string on one lane. */
int monoid_batch_engine_dispatch_prefix(const uint8_t* raw_concat, const int* offsets, int* results, int B,
                                        int total_chars, float* kernel_ms, float* total_ms);
/* cuda/monoid_batch.cu:640-645 */
void monoid_batch_engine_destroy(void);

#ifdef __cplusplus
}
#endif

#endif /* TREX_HIP_H */
