"""DFA oracles (test infrastructure only; see ``oracle/__init__.py``)."""

from __future__ import annotations

import numpy as np


def dfa_simulate(dfa, s: str) -> bool:
    """Restates ``DFA.simulate`` (src/regex_to_dfa.py:224-233): walk the
    transition dict; a character without a transition goes to the dead state,
    or rejects when the DFA has none."""
    state = dfa.start
    for ch in s:
        row = dfa.transitions.get(state, {})
        if ch in row:
            state = row[ch]
        elif dfa.dead_state is not None:
            state = dfa.dead_state
        else:
            return False
    return state in dfa.accept_states


def dfa_final_state(dfa, s: str, dead: int) -> int:
    """Final state of the dict walk, with characters outside the alphabet going
    to ``dead`` (the table semantics of trex_amd.tables.dfa_byte_table)."""
    state = dfa.start
    for ch in s:
        row = dfa.transitions.get(state, {})
        state = row[ch] if ch in row else dead
    return state


def table_walk_one(delta: np.ndarray, accept: np.ndarray, start: int, data: bytes) -> tuple[int, int]:
    """Byte-table walk of one string: ``s = delta[s][byte]`` (the loop of
    ``cpu_prefix_compose``, cuda/prefix_compose.cu:488-516, with the table
    indexed [state][byte])."""
    s = int(start)
    for b in data:
        s = int(delta[s, b])
    return s, int(accept[s])


def table_walk(delta: np.ndarray, accept: np.ndarray, start: int, data: np.ndarray,
               offsets: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Byte-table walk of a CSR batch, vectorised across strings (numpy):
    position t advances every string still longer than t."""
    data = np.asarray(data, dtype=np.uint8)
    offsets = np.asarray(offsets, dtype=np.int64)
    n = len(offsets) - 1
    lengths = np.diff(offsets)
    state = np.full(n, start, dtype=np.int64)
    order = np.argsort(-lengths, kind="stable")
    sorted_len = lengths[order]
    max_len = int(sorted_len[0]) if n else 0
    neg_sorted = -sorted_len                      # ascending
    for t in range(max_len):
        k = int(np.searchsorted(neg_sorted, -t, side="left"))  # strings longer than t
        idx = order[:k]
        state[idx] = delta[state[idx], data[offsets[idx] + t]]
    final = state.astype(np.uint8)
    return final, np.asarray(accept, dtype=np.uint8)[final]


def simulate_monoid(md, s: str) -> bool:
    """Restates ``simulate_monoid`` (src/monoid.py:174-192): accumulate one
    monoid index, newer character on the left; characters outside the
    alphabet are the identity."""
    acc = md.identity_idx
    for ch in s:
        acc = int(md.compose_table[md.char_to_monoid.get(ch, md.identity_idx), acc])
    return bool(md.accept_table[acc])


def simulate_kgram_monoid(kg, md, s: str) -> bool:
    """Restates ``simulate_kgram_monoid`` (src/kgram.py:195-232): whole k-grams
    by table lookup, the tail character by character."""
    if not s:
        return bool(md.accept_table[md.identity_idx])
    acc = md.identity_idx
    k = kg.k
    pos = 0
    while pos + k <= len(s):
        acc = int(md.compose_table[kg.lookup(tuple(s[pos:pos + k])), acc])
        pos += k
    while pos < len(s):
        acc = int(md.compose_table[md.char_to_monoid.get(s[pos], md.identity_idx), acc])
        pos += 1
    return bool(md.accept_table[acc])
