"""Transition monoid of a DFA and the monoid-index tables (host construction).

Same public names and conventions as the reference (``src/monoid.py:30-287``):

* ``compose_table[i, j]`` is the index of ``elements[i] @ elements[j]`` --
  element ``i`` is the newer one (applied after ``j``);
* ``accept_table[k]`` says whether ``elements[k]`` maps the start state to an
  accepting state; ``identity_idx`` is the identity element;
* ``compute_monoid`` returns None when the monoid exceeds ``max_size``.

The closure is computed on transition *functions* (one successor per state)
rather than on N x N matrices: every element is a word over the per-character
functions, so breadth-first extension by one character on the left reaches
the whole monoid. Element order is this BFS order (the reference's order is
its own pairwise-product BFS, ``src/monoid.py:107-143``).

The engine runs a monoid as a byte-table walk whose states are the monoid
elements (``delta[m][byte] = index(char(byte) o m)``, see
:func:`monoid_byte_table`); ``precompute_batch_tables`` reproduces the
reference's table set for its monoid-batch ABI.
"""

from __future__ import annotations

from collections import deque
from dataclasses import dataclass
from typing import Optional

import numpy as np

from trex_amd.simulation import DFAMatrices
from trex_amd.tables import DfaByteTable


@dataclass
class MonoidData:
    functions: np.ndarray        # [M, N] uint8, element k maps state s to functions[k, s]
    compose_table: np.ndarray    # [M, M] uint16 (or uint32 when M > 65535)
    char_to_monoid: dict
    accept_table: np.ndarray     # [M] bool
    identity_idx: int
    size: int

    @property
    def elements(self) -> list[np.ndarray]:
        """The elements as N x N one-hot int8 matrices ``E[dst, src]``."""
        n = self.functions.shape[1]
        out = []
        for f in self.functions:
            e = np.zeros((n, n), np.int8)
            e[f, np.arange(n)] = 1
            out.append(e)
        return out


def _keys(funcs: np.ndarray) -> np.ndarray:
    """Hashable byte-string keys of function rows."""
    return np.ascontiguousarray(funcs).view(np.dtype((np.void, funcs.shape[1])))[:, 0]


def compute_monoid(dm: DFAMatrices, max_size: int = 65536) -> Optional[MonoidData]:
    """BFS closure of the per-character transition functions
    (``src/monoid.py:57-169``); None if it exceeds ``max_size`` elements."""
    n = dm.n_states
    gens = dm.functions()                      # [sigma, N]
    ident = np.arange(n, dtype=np.uint8)
    index: dict[bytes, int] = {}
    funcs: list[np.ndarray] = []

    def add(f: np.ndarray) -> int:
        key = f.tobytes()
        k = index.get(key)
        if k is None:
            k = len(funcs)
            index[key] = k
            funcs.append(f)
        return k

    identity_idx = add(ident)
    char_to_monoid = {ch: add(gens[i]) for i, ch in enumerate(dm.alphabet)}
    if len(funcs) > max_size:
        return None
    queue = deque(range(len(funcs)))
    while queue:
        f = funcs[queue.popleft()]
        for g in gens:
            before = len(funcs)
            add(g[f])                          # character after the word: g o f
            if len(funcs) > before:
                if len(funcs) > max_size:
                    return None
                queue.append(len(funcs) - 1)
    table = np.stack(funcs)                    # [M, N]
    m = len(funcs)
    # compose[i, j] = index(f_i o f_j) = index(f_i[f_j])
    keys = _keys(table)
    order = np.argsort(keys)
    sorted_keys = keys[order]
    dtype = np.uint16 if m <= 65535 else np.uint32
    compose = np.empty((m, m), dtype=dtype)
    for i in range(m):
        prod = table[i][table]                 # row j: f_i o f_j
        pos = np.searchsorted(sorted_keys, _keys(prod))
        compose[i] = order[pos]
    final = table[:, dm.dfa.start]
    accept_table = dm.accept_mask[final].astype(bool)
    return MonoidData(functions=table, compose_table=compose, char_to_monoid=char_to_monoid,
                      accept_table=accept_table, identity_idx=identity_idx, size=m)


def monoid_byte_table(md: MonoidData, dm: DFAMatrices, unknown: str = "identity") -> DfaByteTable:
    """Byte table over monoid elements: ``delta[m][byte] = compose[char(byte), m]``.
    Bytes outside the alphabet are the identity (``simulate_monoid``'s rule,
    ``src/monoid.py:185-190``)."""
    if md.size > 256:
        raise ValueError(f"monoid has {md.size} elements; a byte table holds at most 256")
    m = md.size
    cols = np.tile(np.arange(m, dtype=np.uint8)[:, None], (1, 256))
    for ch, k in md.char_to_monoid.items():
        b = ord(ch)
        if b < 256:
            cols[:, b] = md.compose_table[k, :m]
    return DfaByteTable(delta=cols, accept=md.accept_table.astype(np.uint8), start=md.identity_idx,
                        dead=-1, n_states=m)


def precompute_batch_tables(md: MonoidData, dm: DFAMatrices) -> dict:
    """The reference's monoid-batch GPU tables (``src/monoid.py:195-287``)."""
    m = md.size
    if m > 255:
        raise ValueError(f"Monoid size {m} exceeds uint8 limit (255)")
    sigma = len(dm.alphabet)
    char_compose = np.zeros((m, sigma + 1), np.uint8)
    for ch, ci in dm.char_to_idx.items():
        char_compose[:, ci] = md.compose_table[md.char_to_monoid[ch], :m]
    char_compose[:, sigma] = np.arange(m)
    raw_char_map = np.full(256, sigma, np.uint8)
    for ch, ci in dm.char_to_idx.items():
        if ord(ch) < 256:
            raw_char_map[ord(ch)] = ci
    fused = np.ascontiguousarray(char_compose[:, raw_char_map]).reshape(-1)
    if sigma >= 2:
        k = 0
        while sigma ** (k + 1) <= 256:
            k += 1
        sigma_k = sigma ** k
    else:
        k, sigma_k = 1, 1
    kgram = None
    if sigma >= 2 and k >= 2:
        # gram index: first character most significant
        digits = np.array([[(g // sigma ** (k - 1 - p)) % sigma for p in range(k)] for g in range(sigma_k)])
        cur = np.tile(np.arange(m)[:, None], (1, sigma_k))
        for p in range(k):
            cur = char_compose[cur, digits[None, :, p]]
        kgram = np.ascontiguousarray(cur.astype(np.uint8).reshape(-1))
    return {
        "char_compose": np.ascontiguousarray(char_compose.reshape(-1)),
        "raw_char_map": raw_char_map,
        "accept": md.accept_table.astype(np.uint8),
        "monoid_compose": np.ascontiguousarray(md.compose_table.astype(np.uint8).reshape(-1)),
        "fused_compose": fused,
        "kgram_compose": kgram,
        "M": m, "sigma": sigma, "sigma_ext": sigma + 1, "sigma_k": sigma_k,
        "kgram_k": k, "identity_idx": md.identity_idx,
    }
