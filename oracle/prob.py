"""FP64 oracle of the probabilistic NFA-64 forward pass (test infrastructure
only; see ``oracle/__init__.py``).

The recurrence is the one the reference's V5/V6 probabilistic kernels evolve,
``s' = T[c] s`` with ``T[c]`` stored row-major as [dst][src]
(cuda/nfa_tc_evolution.cu:451-561 and the scalar V6, :572-610). The
reference returns only ``(a . s_L) > 0.5``; the quantity defined here is the
log-likelihood ``LL = log(a . s_L)`` and the mass ``log(1 . s_L)``, which the
reference has no counterpart for: parity unpinned against the reference,
pinned instead by closed forms in tests/test_oracle.py (identity and
permutation matrices, rank-one matrices, exact scaling, a single symbol's
stationary limit, and a brute-force matrix product on small cases).
"""

from __future__ import annotations

import numpy as np

N = 64


def forward_loglik(T: np.ndarray, start: np.ndarray, accept: np.ndarray, symbols: np.ndarray,
                   col_scale: np.ndarray | None = None, T_last: np.ndarray | None = None):
    """FP64 forward pass over a batch of equal-length sequences.

    ``T``: [σ, n, n], ``symbols``: [n_seq, L] integer symbol indices. Each step
    ``s <- T[c] (col_scale[c] * s)`` (``col_scale`` defaults to ones), except
    the last, which applies ``T_last[c]`` (defaults to ``T``) with no column
    scale. After every step the state is divided by the power of two that
    brings its largest magnitude into [0.5, 1) and the exponent is added to an
    integer scale, so nothing over- or underflows. A symbol >= σ empties the
    state. Returns (LL, logmass), float64 [n_seq]; -inf where the product is 0.
    """
    T = np.asarray(T, dtype=np.float64)
    sigma, n = T.shape[0], T.shape[1]
    T_last = T if T_last is None else np.asarray(T_last, dtype=np.float64)
    symbols = np.asarray(symbols)
    n_seq, L = symbols.shape
    s = np.tile(np.asarray(start, dtype=np.float64), (n_seq, 1))
    expo = np.zeros(n_seq, dtype=np.int64)
    s, e = _normalise(s)
    expo += e
    for t in range(L):
        last = t == L - 1
        col = symbols[:, t]
        nxt = np.zeros_like(s)
        for c in np.unique(col):
            if c >= sigma:
                continue
            rows = col == c
            x = s[rows]
            if col_scale is not None and not last:
                x = x * np.asarray(col_scale, dtype=np.float64)[c]
            nxt[rows] = x @ (T_last[c] if last else T[c]).T
        s, e = _normalise(nxt)
        expo += e
    scale = expo.astype(np.float64) * np.log(2.0)
    a = np.asarray(accept, dtype=np.float64)
    with np.errstate(divide="ignore"):
        ll = np.log(s @ a) + scale
        lm = np.log(s.sum(axis=1)) + scale
    return ll, lm


def _normalise(s: np.ndarray):
    m = np.abs(s).max(axis=1)
    _, e = np.frexp(m)                 # m = f * 2^e, f in [0.5, 1); e = 0 for m = 0
    return np.ldexp(s, -e[:, None]), e.astype(np.int64)


def loglik_bruteforce(T: np.ndarray, start: np.ndarray, accept: np.ndarray, seq) -> tuple[float, float]:
    """One short sequence by forming the matrix product P = T[c_L] ... T[c_1]
    with plain loops (no rescaling; for small L only), then a . (P s0)."""
    T = np.asarray(T, dtype=np.float64)
    n = T.shape[1]
    P = [[1.0 if i == j else 0.0 for j in range(n)] for i in range(n)]
    for c in seq:
        if c >= T.shape[0]:
            P = [[0.0] * n for _ in range(n)]
            continue
        Tc = T[c]
        P = [[sum(Tc[i, k] * P[k][j] for k in range(n)) for j in range(n)] for i in range(n)]
    v = [sum(P[i][j] * float(start[j]) for j in range(n)) for i in range(n)]
    a_dot = sum(v[i] * float(accept[i]) for i in range(n))
    mass = sum(v)
    ll = float(np.log(a_dot)) if a_dot > 0 else -np.inf
    lm = float(np.log(mass)) if mass > 0 else -np.inf
    return ll, lm
