"""Shared oracle comparisons for the GPU tests of served and restored results.

Every result of a DFA / NFA engine is checked bit-exact against the C oracle's
walk of the same tables; every probabilistic log-likelihood against the FP64
oracle within the stated relative tolerance.
"""

from __future__ import annotations

import numpy as np

from oracle import cpu as ocpu

# relative log-likelihood error allowed against the FP64 oracle
LL_REL_TOL = 1e-3


def assert_walk_matches_oracle(engine, data, offsets, final, accept):
    """``final`` / ``accept`` of a DfaEngine or NfaEngine equal the oracle's walk."""
    from trex_amd.engine import DfaEngine, NfaEngine
    data = np.ascontiguousarray(data, dtype=np.uint8)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    if isinstance(engine, DfaEngine):
        t = engine.table
        rf, ra = ocpu.dfa_walk(t.delta, t.accept, t.start, data, offsets)
    elif isinstance(engine, NfaEngine):
        rf, ra = ocpu.nfa_walk(engine.succ, engine.byte_to_sym, engine.start, engine.accept, data, offsets)
    else:
        raise TypeError(type(engine).__name__)
    final = np.asarray(final).astype(rf.dtype)
    bad = np.flatnonzero((final != rf) | (np.asarray(accept) != ra))
    assert bad.size == 0, f"{bad.size} strings differ from the oracle, first {bad[:5]}"


def oracle_loglik(engine, data, offsets):
    """FP64 LL and log-mass of every sequence of a ProbNfaEngine's ragged
    batch (sequences of one length go to the C oracle together)."""
    data = np.asarray(data, dtype=np.uint8)
    offsets = np.asarray(offsets, dtype=np.int64)
    lengths = np.diff(offsets)
    ll = np.empty(len(lengths))
    lm = np.empty(len(lengths))
    for L in np.unique(lengths):
        idx = np.flatnonzero(lengths == L)
        sy = np.stack([data[offsets[i]:offsets[i] + L] for i in idx]) if L else np.zeros((len(idx), 0), np.uint8)
        a, b = ocpu.pnfa_forward(engine.T, engine.start, engine.accept, sy)
        ll[idx], lm[idx] = a, b
    return ll, lm


def assert_loglik_matches_oracle(engine, data, offsets, ll, tol: float = LL_REL_TOL):
    ref, _ = oracle_loglik(engine, data, offsets)
    ll = np.asarray(ll, dtype=np.float64)
    both_inf = np.isneginf(ref) & np.isneginf(ll)
    err = np.where(both_inf, 0.0, np.abs(ll - ref) / np.maximum(np.abs(ref), 1e-300))
    assert np.all(err <= tol), f"max relative LL error {np.nanmax(err):.3e} > {tol} (FP64 oracle)"
