"""NFA -> per-symbol Boolean transition matrices (host-side construction).

Convention as in the reference (``src/nfa_matrices.py:40-113``):
``T[c][dst, src] = 1`` iff ``src`` reaches ``dst`` by reading symbol ``c``
followed by any number of epsilon moves, so one step is
``state = min(T[c] @ state, 1)``. The start vector is the epsilon closure of
the NFA start state. ``n_states`` is padded to a multiple of 16 unless
``pad_to`` is given.

Besides regex-derived NFAs, :meth:`NFAMatrices.from_matrices` wraps an
arbitrary Boolean automaton (used for random 64-state NFAs). The GPU engine
consumes the same object through :meth:`NFAMatrices.successor_masks`, which
packs ``T`` into one 64-bit successor mask per (symbol, source state).
"""

from __future__ import annotations

from typing import Optional, Sequence

import numpy as np

from trex_amd.regex_to_dfa import NFA, RegexParser, epsilon_closure


class NFAMatrices:
    def __init__(self, nfa: Optional[NFA], pad_to: Optional[int] = None):
        self.nfa = nfa
        if nfa is None:          # filled in by from_matrices
            return
        self.alphabet = sorted(nfa.alphabet)
        self.char_to_idx = {c: i for i, c in enumerate(self.alphabet)}
        self.n_states_raw = len(nfa.states)
        if pad_to is not None:
            self.n_states = max(pad_to, self.n_states_raw)
        else:
            self.n_states = -(-self.n_states_raw // 16) * 16
        n = self.n_states
        stack = np.zeros((len(self.alphabet), n, n), dtype=np.int8)
        for src, st in nfa.states.items():
            for ch, dsts in st.transitions.items():
                if ch is None:
                    continue
                for dst in epsilon_closure(nfa, dsts):
                    stack[self.char_to_idx[ch], dst, src] = 1
        self.matrix_stack = stack
        self.matrices = {ch: stack[i] for i, ch in enumerate(self.alphabet)}
        self.start_vec = np.zeros(n, dtype=np.int8)
        for s in epsilon_closure(nfa, [nfa.start]):
            self.start_vec[s] = 1
        self.accept_mask = np.zeros(n, dtype=np.int8)
        for sid, st in nfa.states.items():
            if st.is_accept:
                self.accept_mask[sid] = 1

    @classmethod
    def from_matrices(cls, matrix_stack: np.ndarray, start_vec: Sequence[int],
                      accept_mask: Sequence[int], alphabet: Sequence[str]) -> "NFAMatrices":
        """Wrap Boolean matrices ``matrix_stack[c][dst, src]`` directly."""
        stack = np.ascontiguousarray(np.asarray(matrix_stack) != 0, dtype=np.int8)
        sig, n, n2 = stack.shape
        if n != n2 or sig != len(alphabet):
            raise ValueError("matrix_stack must be [len(alphabet)][N][N]")
        obj = cls(None)
        obj.alphabet = list(alphabet)
        obj.char_to_idx = {c: i for i, c in enumerate(obj.alphabet)}
        obj.n_states_raw = n
        obj.n_states = n
        obj.matrix_stack = stack
        obj.matrices = {ch: stack[i] for i, ch in enumerate(obj.alphabet)}
        obj.start_vec = (np.asarray(start_vec) != 0).astype(np.int8)
        obj.accept_mask = (np.asarray(accept_mask) != 0).astype(np.int8)
        return obj

    def check_accept(self, state_vec: np.ndarray) -> bool:
        """Whether any real (unpadded) state set in ``state_vec`` accepts
        (``src/nfa_matrices.py:117-122``)."""
        raw = self.n_states_raw
        vec = np.asarray(state_vec)[:raw].astype(np.int8)
        return bool(np.any(vec & self.accept_mask[:raw]))

    # -- packed forms for the GPU --------------------------------------------
    def successor_masks(self) -> np.ndarray:
        """``succ[c][src]`` = 64-bit mask of destinations (bit d = state d),
        shape [n_symbols, 64] (sources beyond the NFA's states have no edges)."""
        if self.n_states_raw > 64:
            raise ValueError(f"NFA has {self.n_states_raw} states; the bit-parallel "
                             "engine holds at most 64")
        n = min(self.n_states, 64)
        stack = self.matrix_stack[:, :n, :n].astype(np.uint64)
        weights = (np.uint64(1) << np.arange(n, dtype=np.uint64))
        out = np.zeros((stack.shape[0], 64), dtype=np.uint64)
        # succ[c, src] = sum_dst T[c, dst, src] << dst
        out[:, :n] = (stack * weights[None, :, None]).sum(axis=1, dtype=np.uint64)
        return out

    def start_mask(self) -> int:
        return _bits_to_int(self.start_vec[:64])

    def accept_bits(self) -> int:
        return _bits_to_int(self.accept_mask[:self.n_states_raw][:64])


def _bits_to_int(vec) -> int:
    out = 0
    for i, v in enumerate(np.asarray(vec).tolist()):
        if v:
            out |= 1 << i
    return out


def compile_nfa_matrices(regex: str, pad_to: Optional[int] = None) -> NFAMatrices:
    """Parse ``regex`` into a Thompson NFA and wrap it in :class:`NFAMatrices`
    (``src/nfa_matrices.py:127-130``)."""
    return NFAMatrices(RegexParser(regex).parse(), pad_to=pad_to)
