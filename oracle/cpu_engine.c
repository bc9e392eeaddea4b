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
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

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

/*
 *  oracle_pnfa_forward: FP64 forward pass of the probabilistic NFA-64, the
 *    restatement of oracle/prob.py:forward_loglik (s <- T[c] (scale[c] * s)
 *    per step, the last step with T_last and no scale; exact power-of-two
 *    renormalisation into an integer exponent; symbol >= sigma empties the
 *    state). Tt / Tt_last are the matrices TRANSPOSED, [sigma][src][dst], so
 *    the product is a sum of scaled columns (s'[i] += T[i][j] s[j], j outer).
 */
typedef struct {
    const double* Tt;
    const double* Tt_last;
    const double* scale;
    int sigma;
    const double* s0;
    int e0;
    const double* a;
    const uint8_t* sym;
    int64_t L;
    int64_t lo, hi;
    double* ll;
    double* lm;
} pnfa_job;

static void* pnfa_worker(void* arg) {
    pnfa_job* j = (pnfa_job*)arg;
    double s[64], x[64], nx[64];
    for (int64_t q = j->lo; q < j->hi; ++q) {
        memcpy(s, j->s0, sizeof(s));
        long long expo = j->e0;
        const uint8_t* sy = j->sym + q * j->L;
        for (int64_t t = 0; t < j->L; ++t) {
            const int c = sy[t];
            const int last = t == j->L - 1;
            if (c >= j->sigma) {
                memset(s, 0, sizeof(s));
                continue;
            }
            const double* M = (last ? j->Tt_last : j->Tt) + (size_t)c * 4096;
            for (int k = 0; k < 64; ++k) x[k] = (j->scale && !last) ? s[k] * j->scale[c * 64 + k] : s[k];
            for (int i = 0; i < 64; ++i) nx[i] = 0.0;
            for (int k = 0; k < 64; ++k) {
                const double xk = x[k];
                const double* col = M + k * 64;
                for (int i = 0; i < 64; ++i) nx[i] += col[i] * xk;
            }
            double m = 0.0;
            for (int i = 0; i < 64; ++i) m = fmax(m, fabs(nx[i]));
            int e = 0;
            if (m > 0.0) frexp(m, &e);
            for (int i = 0; i < 64; ++i) s[i] = ldexp(nx[i], -e);
            expo += e;
        }
        double da = 0.0, dm = 0.0;
        for (int i = 0; i < 64; ++i) {
            da += s[i] * j->a[i];
            dm += s[i];
        }
        const double sc = (double)expo * 0.69314718055994530942;
        j->ll[q] = (da > 0.0 ? log(da) : (da == 0.0 ? -INFINITY : NAN)) + sc;
        j->lm[q] = (dm > 0.0 ? log(dm) : (dm == 0.0 ? -INFINITY : NAN)) + sc;
    }
    return NULL;
}

int oracle_pnfa_forward(const double* Tt, const double* Tt_last, const double* scale, int sigma, const double* s0,
                        int e0, const double* a, const uint8_t* sym, int64_t n_seq, int64_t L, double* ll,
                        double* lm, int threads) {
    if (sigma < 1 || n_seq < 0 || L < 0) return -3;
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t tid[256];
    pnfa_job jobs[256];
    int64_t per = (n_seq + threads - 1) / threads;
    int launched = 0;
    for (int t = 0; t < threads; ++t) {
        int64_t lo = t * per, hi = lo + per;
        if (hi > n_seq) hi = n_seq;
        if (lo >= hi) break;
        pnfa_job jb = {Tt, Tt_last, scale, sigma, s0, e0, a, sym, L, lo, hi, ll, lm};
        jobs[t] = jb;
        if (pthread_create(&tid[t], NULL, pnfa_worker, &jobs[t]) != 0) { pnfa_worker(&jobs[t]); tid[t] = 0; }
        launched = t + 1;
    }
    for (int t = 0; t < launched; ++t)
        if (tid[t]) pthread_join(tid[t], NULL);
    return 0;
}

/*
 *  oracle_gen_symbols: bytes [start, start + n) of the synthetic input that
 *    trex_gen_symbols writes on the device: byte i = lo + (r & (span - 1)),
 *    r = byte (i % 8) of splitmix64(seed ^ ((i / 8) * 0x9E3779B97F4A7C15)).
 */
static uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

void oracle_gen_symbols(uint8_t* out, int64_t start, int64_t n, uint64_t seed, uint32_t lo, uint32_t span) {
    int64_t i = start;
    const int64_t stop = start + n;
    while (i < stop) {
        const uint64_t w = (uint64_t)i >> 3;
        const uint64_t h = splitmix64(seed ^ (w * 0x9E3779B97F4A7C15ull));
        for (int k = (int)(i & 7); k < 8 && i < stop; ++k, ++i)
            out[i - start] = (uint8_t)(lo + ((h >> (8 * k)) & (span - 1)));
    }
}
