"""Boolean NFA oracles (test infrastructure only; see ``oracle/__init__.py``)."""

from __future__ import annotations

import numpy as np


def simulate_nfa(nm, s: str) -> bool:
    """Restates ``simulate_nfa`` (src/nfa_matrices.py:133-158):
    ``state = min(T[c] @ state, 1)`` per character; a character outside the
    alphabet rejects; the empty string checks the start vector."""
    n_raw = nm.n_states_raw
    if not s:
        return bool(np.any(nm.start_vec[:n_raw] & nm.accept_mask[:n_raw]))
    state = nm.start_vec.astype(np.int32)
    for ch in s:
        idx = nm.char_to_idx.get(ch)
        if idx is None:
            return False
        state = nm.matrix_stack[idx].astype(np.int32) @ state
        np.minimum(state, 1, out=state)
    return bool(np.any(state[:n_raw].astype(np.int8) & nm.accept_mask[:n_raw]))


def set_walk_one(succ: np.ndarray, byte_to_sym: np.ndarray, start: int, accept: int,
                 data: bytes) -> tuple[int, int]:
    """Set-of-states walk over packed successor masks: the new set is the union
    of ``succ[sym][j]`` over members j of the current set. ``byte_to_sym`` maps
    a byte to a symbol index, -1 (reject: empty set) or -2 (skip: identity,
    the behaviour of ``simulate_nfa_sequential``, cuda/nfa_tc_evolution.cu:1029-1054)."""
    state = int(start)
    for b in data:
        sym = int(byte_to_sym[b])
        if sym == -2:
            continue
        if sym == -1:
            state = 0
            continue
        nxt = 0
        row = succ[sym]
        j = state
        while j:
            low = j & -j
            nxt |= int(row[low.bit_length() - 1])
            j ^= low
        state = nxt
    return state, int((state & accept) != 0)
