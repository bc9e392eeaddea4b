"""k-gram precomputation over DFA monoids or matrices (host construction).

Same names and conventions as the reference (``src/kgram.py:29-190``): the key
of a k-gram ``(c0, ..., c_{k-1})`` is mixed-radix with the first character
most significant, ``key = ((c0 * sigma) + c1) * sigma + ...``, and composition
runs left to right with the newer character on the left.
"""

from __future__ import annotations

import itertools
from typing import Optional

import numpy as np

from trex_amd.monoid import MonoidData
from trex_amd.simulation import DFAMatrices


def auto_k(alphabet_size: int, max_entries: int = 65536) -> int:
    """Largest k with ``alphabet_size ** k <= max_entries`` (1 for alphabets of size <= 1)."""
    if alphabet_size <= 1:
        return 1
    k = 1
    while alphabet_size ** (k + 1) <= max_entries:
        k += 1
    return k


def auto_k_for_gpu(alphabet_size: int, n_states: int, max_table_bytes: int = 48_000_000) -> int:
    """Largest k with ``alphabet_size ** k * n_states ** 2 <= max_table_bytes``
    (the reference stops at k = 1 even if that exceeds the budget)."""
    if alphabet_size <= 1:
        return 1
    k = 1
    while alphabet_size ** (k + 1) * n_states * n_states <= max_table_bytes:
        k += 1
    return k


class KGramTable:
    def __init__(self, k: int, alphabet: list[str], monoid_table: Optional[dict] = None,
                 matrix_table: Optional[dict] = None):
        self.k = k
        self.alphabet = list(alphabet)
        self.sigma = len(self.alphabet)
        self.char_to_idx = {ch: i for i, ch in enumerate(self.alphabet)}
        self._monoid_table = monoid_table
        self._matrix_table = matrix_table

    def _gram_to_key(self, gram) -> int:
        key = 0
        for ch in gram:
            key = key * self.sigma + self.char_to_idx[ch]
        return key

    def lookup(self, gram) -> int:
        return self._monoid_table[self._gram_to_key(gram)]

    def lookup_matrix(self, gram) -> np.ndarray:
        return self._matrix_table[self._gram_to_key(gram)]


def precompute_kgrams(dm: DFAMatrices, k: int, monoid: Optional[MonoidData] = None) -> KGramTable:
    """All k-grams composed to a monoid index (``monoid`` given) or to a
    matrix (``src/kgram.py:110-190``)."""
    sigma = len(dm.alphabet)
    if monoid is not None:
        chars = [monoid.char_to_monoid[ch] for ch in dm.alphabet]
        table = {}
        for key, gram in enumerate(itertools.product(range(sigma), repeat=k)):
            acc = monoid.identity_idx
            for c in gram:
                acc = int(monoid.compose_table[chars[c], acc])
            table[key] = acc
        return KGramTable(k, dm.alphabet, monoid_table=table)
    funcs = dm.functions()
    n = dm.n_states
    table = {}
    for key, gram in enumerate(itertools.product(range(sigma), repeat=k)):
        f = np.arange(n)
        for c in gram:
            f = funcs[c][f]
        mat = np.zeros((n, n), np.int8)
        mat[f, np.arange(n)] = 1
        table[key] = mat
    return KGramTable(k, dm.alphabet, matrix_table=table)
