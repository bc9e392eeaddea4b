/*
 * cpu_engine.c -- multithreaded C restatement of the matching loops, used as
 * the full-size checker and the timed CPU baseline. TEST INFRASTRUCTURE ONLY:
 * nothing in trex_amd links or calls it.
 *
 *  oracle_dfa_walk: s = delta[s * 256 + byte] per byte, final state and
 *    accept bit per string -- the loop of cpu_prefix_compose
 *    (cuda/prefix_compose.cu:488-516) and cpu_monoid_batch
 *    (cuda/monoid_batch.cu:679-699), with the table indexed [state][byte].
 *  oracle_nfa_walk: set-of-states walk, new set = union of succ[sym][j] over
 *    the members j of the current set (the bit-parallel step of
 *    compose_bitpar_host, cuda/nfa_tc_evolution.cu:623-634, applied to a
 *    state vector). byte_to_sym: symbol index, -1 = reject (empty set),
 *    -2 = skip (identity).
 *
 * Strings are split over `threads` pthreads in contiguous ranges.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>

typedef struct {
    const uint8_t* delta;
    const uint8_t* accept;
    int start;
    const uint8_t* bytes;
    const int64_t* offsets;
    int64_t lo, hi;
    uint8_t* final_state;
    uint8_t* acc;
} dfa_job;

static void* dfa_worker(void* arg) {
    dfa_job* j = (dfa_job*)arg;
    const uint8_t* delta = j->delta;
    for (int64_t i = j->lo; i < j->hi; ++i) {
        const uint8_t* p = j->bytes + j->offsets[i];
        const uint8_t* e = j->bytes + j->offsets[i + 1];
        unsigned s = (unsigned)j->start;
        while (p < e) s = delta[s * 256u + *p++];
        j->final_state[i] = (uint8_t)s;
        j->acc[i] = j->accept[s] ? 1 : 0;
    }
    return NULL;
}

int oracle_dfa_walk(const uint8_t* delta, const uint8_t* accept, int n_states, int start,
                    const uint8_t* bytes, const int64_t* offsets, int64_t n_strings,
                    uint8_t* final_state, uint8_t* acc, int threads) {
    if (n_states < 1 || n_states > 256 || start < 0 || start >= n_states || n_strings < 0) return -3;
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t tid[256];
    dfa_job jobs[256];
    int64_t per = (n_strings + threads - 1) / threads;
    int launched = 0;
    for (int t = 0; t < threads; ++t) {
        int64_t lo = t * per, hi = lo + per;
        if (hi > n_strings) hi = n_strings;
        if (lo >= hi) break;
        dfa_job jb = {delta, accept, start, bytes, offsets, lo, hi, final_state, acc};
        jobs[t] = jb;
        if (pthread_create(&tid[t], NULL, dfa_worker, &jobs[t]) != 0) { dfa_worker(&jobs[t]); tid[t] = 0; }
        launched = t + 1;
    }
    for (int t = 0; t < launched; ++t)
        if (tid[t]) pthread_join(tid[t], NULL);
    return 0;
}

typedef struct {
    const uint64_t* succ;     /* [n_sym][64] */
    const int16_t* byte_to_sym;
    uint64_t start, accept;
    const uint8_t* bytes;
    const int64_t* offsets;
    int64_t lo, hi;
    uint64_t* final_mask;
    uint8_t* acc;
} nfa_job;

static void* nfa_worker(void* arg) {
    nfa_job* j = (nfa_job*)arg;
    for (int64_t i = j->lo; i < j->hi; ++i) {
        const uint8_t* p = j->bytes + j->offsets[i];
        const uint8_t* e = j->bytes + j->offsets[i + 1];
        uint64_t s = j->start;
        for (; p < e; ++p) {
            int sym = j->byte_to_sym[*p];
            if (sym == -2) continue;
            if (sym < 0) { s = 0; continue; }
            const uint64_t* row = j->succ + (size_t)sym * 64;
            uint64_t nxt = 0, m = s;
            while (m) {
                nxt |= row[__builtin_ctzll(m)];
                m &= m - 1;
            }
            s = nxt;
        }
        j->final_mask[i] = s;
        j->acc[i] = (s & j->accept) ? 1 : 0;
    }
    return NULL;
}

int oracle_nfa_walk(const uint64_t* succ, int n_sym, const int16_t* byte_to_sym, uint64_t start,
                    uint64_t accept, const uint8_t* bytes, const int64_t* offsets, int64_t n_strings,
                    uint64_t* final_mask, uint8_t* acc, int threads) {
    if (n_sym < 0 || n_strings < 0) return -3;
    for (int b = 0; b < 256; ++b)
        if (byte_to_sym[b] >= n_sym || byte_to_sym[b] < -2) return -3;
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t tid[256];
    nfa_job jobs[256];
    int64_t per = (n_strings + threads - 1) / threads;
    int launched = 0;
    for (int t = 0; t < threads; ++t) {
        int64_t lo = t * per, hi = lo + per;
        if (hi > n_strings) hi = n_strings;
        if (lo >= hi) break;
        nfa_job jb = {succ, byte_to_sym, start, accept, bytes, offsets, lo, hi, final_mask, acc};
        jobs[t] = jb;
        if (pthread_create(&tid[t], NULL, nfa_worker, &jobs[t]) != 0) { nfa_worker(&jobs[t]); tid[t] = 0; }
        launched = t + 1;
    }
    for (int t = 0; t < launched; ++t)
        if (tid[t]) pthread_join(tid[t], NULL);
    return 0;
}
