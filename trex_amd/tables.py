"""Byte-level transition tables handed to the GPU engine.

A DFA is given to the engine as a full byte table ``delta[state][byte]``
(uint8, ``n_states`` rows of 256) plus ``accept[state]`` and a start state.
The table fixes every semantic choice, so the GPU result and the CPU oracle
agree bit-for-bit as long as both walk the same table:

* :func:`dfa_byte_table` (the engine's default) sends every byte outside the
  DFA alphabet to the dead state, which is absorbing and non-accepting. This
  is ``DFA.simulate``'s behaviour (``src/regex_to_dfa.py:224-233``): reject
  once an unknown character is read. A dead state is added when the DFA has
  none.
* :func:`tmap_byte_table` wraps the reference's ``precompute_tmap`` layout
  (``tmap[byte * N + s]``), where unknown bytes are the identity
  (``src/simulation.py:274-297``); it is used by the reference-compatible
  ``prefix+gpu`` entry points.

Padded states (beyond ``n_states``) never occur: no transition leads to them.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from trex_amd.regex_to_dfa import DFA

MAX_TABLE_STATES = 256   # the generic byte-walk kernel indexes states with uint8
DFA16_STATES = 16        # the K-gram fused kernels hold 16 states


@dataclass
class DfaByteTable:
    delta: np.ndarray      # uint8 [n_states, 256]
    accept: np.ndarray     # uint8 [n_states]
    start: int
    dead: int              # absorbing non-accepting state, or -1
    n_states: int

    def __post_init__(self):
        self.delta = np.ascontiguousarray(self.delta, dtype=np.uint8)
        self.accept = np.ascontiguousarray(self.accept, dtype=np.uint8)
        if self.delta.shape != (self.n_states, 256):
            raise ValueError(f"delta must be [{self.n_states}, 256], got {self.delta.shape}")
        if not 1 <= self.n_states <= MAX_TABLE_STATES:
            raise ValueError(f"n_states={self.n_states} outside 1..{MAX_TABLE_STATES}")
        if int(self.delta.max(initial=0)) >= self.n_states:
            raise ValueError("delta names a state >= n_states")
        if not 0 <= self.start < self.n_states:
            raise ValueError("start state out of range")


def dfa_byte_table(dfa: DFA) -> DfaByteTable:
    """Byte table of a complete DFA with unknown bytes -> dead state."""
    n = dfa.n_states
    dead = dfa.dead_state
    if dead is None:
        dead = n
        n += 1
    delta = np.full((n, 256), dead, dtype=np.uint8) if n <= 256 else None
    if delta is None:
        raise ValueError(f"DFA needs {n} states with its dead state; at most "
                         f"{MAX_TABLE_STATES} fit a byte table")
    for s in range(dfa.n_states):
        row = dfa.transitions.get(s, {})
        for ch, dst in row.items():
            b = ord(ch)
            if b < 256:
                delta[s, b] = dst
    accept = np.zeros(n, dtype=np.uint8)
    for s in dfa.accept_states:
        accept[s] = 1
    return DfaByteTable(delta=delta, accept=accept, start=dfa.start, dead=dead, n_states=n)


def tmap_byte_table(tmap: np.ndarray, accept, start: int, n: int) -> DfaByteTable:
    """Wrap a reference ``tmap[byte * n + s]`` table (identity for unknown bytes)."""
    tmap = np.asarray(tmap, dtype=np.uint8).reshape(256, n)
    acc = np.zeros(n, dtype=np.uint8)
    acc[: len(accept)] = np.asarray(accept, dtype=np.uint8)[:n]
    return DfaByteTable(delta=np.ascontiguousarray(tmap.T), accept=acc, start=int(start),
                        dead=-1, n_states=n)
