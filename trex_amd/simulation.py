"""DFA -> transition matrices and byte tables (host-side construction).

``DFAMatrices`` keeps the reference's matrix convention
(``src/simulation.py:20-101``): ``T[c][dst, src] = 1`` iff ``delta(src, c) = dst``,
states padded to a multiple of 16 with padded states looping on themselves.

``precompute_tmap`` is the reference's fused byte table
(``src/simulation.py:274-297``): ``tmap[byte * N + s]``, with bytes outside the
alphabet acting as the identity. The engine's own table, which sends such
bytes to the dead state, is :func:`trex_amd.tables.dfa_byte_table`.
"""

from __future__ import annotations

from typing import Optional

import numpy as np

from trex_amd.regex_to_dfa import DFA


class DFAMatrices:
    def __init__(self, dfa: DFA, pad_to: Optional[int] = None):
        self.dfa = dfa
        self.alphabet = sorted(dfa.alphabet)
        self.char_to_idx = {c: i for i, c in enumerate(self.alphabet)}
        self.n_states_raw = dfa.n_states
        if pad_to is not None:
            self.n_states = max(pad_to, dfa.n_states)
        else:
            self.n_states = -(-dfa.n_states // 16) * 16
        n = self.n_states
        stack = np.zeros((len(self.alphabet), n, n), dtype=np.int8)
        pad = np.arange(self.n_states_raw, n)
        for i, ch in enumerate(self.alphabet):
            for src in range(self.n_states_raw):
                dst = dfa.transitions.get(src, {}).get(ch)
                if dst is not None:
                    stack[i, dst, src] = 1
            stack[i, pad, pad] = 1
        self.matrix_stack = stack
        self.matrices = {ch: stack[i] for i, ch in enumerate(self.alphabet)}
        self.start_vec = np.zeros(n, dtype=np.int8)
        self.start_vec[dfa.start] = 1
        self.accept_mask = np.zeros(n, dtype=np.int8)
        for s in dfa.accept_states:
            self.accept_mask[s] = 1

    def get_matrix_for_char(self, ch: str):
        return self.matrices.get(ch)

    def get_matrix_sequence(self, input_str: str) -> np.ndarray:
        """``[L, N, N]`` stack of the string's transition matrices; a character
        outside the alphabet contributes the identity (``src/simulation.py:80-93``)."""
        n = self.n_states
        idx = np.fromiter((self.char_to_idx.get(ch, -1) for ch in input_str), dtype=np.int64,
                          count=len(input_str))
        seq = np.zeros((len(input_str), n, n), dtype=np.int8)
        known = idx >= 0
        seq[known] = self.matrix_stack[idx[known]]
        unknown = np.flatnonzero(~known)
        if unknown.size:
            diag = np.arange(n)
            seq[unknown[:, None], diag[None, :], diag[None, :]] = 1
        return seq

    def check_accept(self, state_vec: np.ndarray) -> bool:
        """Whether any real (unpadded) state set in ``state_vec`` accepts
        (``src/simulation.py:95-97``)."""
        raw = self.n_states_raw
        vec = np.asarray(state_vec)[:raw].astype(np.int8)
        return bool(np.any(vec & self.accept_mask[:raw]))

    def identity_matrix(self) -> np.ndarray:
        return np.eye(self.n_states, dtype=np.int8)

    def functions(self) -> np.ndarray:
        """``f[c][src] = dst`` for every symbol (padded states map to themselves)."""
        return np.ascontiguousarray(np.argmax(self.matrix_stack, axis=1).astype(np.uint8))


def precompute_tmap(dm: DFAMatrices) -> np.ndarray:
    """``tmap[byte * N + s]`` = successor of ``s`` on ``byte``; bytes outside the
    alphabet leave the state unchanged (``src/simulation.py:274-297``)."""
    n = dm.n_states
    tmap = np.tile(np.arange(n, dtype=np.uint8), 256).reshape(256, n)
    funcs = dm.functions()
    for i, ch in enumerate(dm.alphabet):
        b = ord(ch)
        if b < 256:
            tmap[b, :dm.n_states_raw] = funcs[i, :dm.n_states_raw]
    return np.ascontiguousarray(tmap.reshape(-1))
