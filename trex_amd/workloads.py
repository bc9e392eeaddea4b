"""Seeded synthetic workloads (numpy PCG64): random patterns and string batches.

* :func:`random_regex` draws a regex AST over an alphabet (concatenation,
  alternation, star; depth <= 4) and renders it as pattern text, escaping the
  characters the parser treats as syntax (the 0x40-0x5F alphabet contains
  ``[ \\ ]``).
* :func:`random_dfa_pattern` redraws until the minimal complete DFA, with its
  dead state for bytes outside the alphabet, has at most ``max_states`` states.
* :func:`random_nfa64` draws a Boolean NFA with each per-symbol edge present
  with probability ``p``, start state 0 and ``n_accept`` random accept states.
* :func:`uniform_batch` / :func:`loguniform_batch` build CSR byte batches.
* :func:`plain_dfa_workload` is the CPU-path plain DFA workload: one seeded
  regex over ``abcd`` and 1,024 strings of 1,024 bytes over that alphabet.
"""

from __future__ import annotations

import numpy as np

from trex_amd.nfa_matrices import NFAMatrices
from trex_amd.regex_to_dfa import DFA, compile_regex
from trex_amd.tables import dfa_byte_table

SYNTAX_CHARS = set("[\\]()|*+?.")
ALPHABET_ABCD = "abcd"
ALPHABET_32 = "".join(chr(c) for c in range(0x40, 0x60))


def _lit(ch: str) -> str:
    return "\\" + ch if ch in SYNTAX_CHARS else ch


def _draw(rng: np.random.Generator, alphabet: str, depth: int, max_depth: int) -> str:
    if depth >= max_depth or rng.random() < 0.3:
        return _lit(alphabet[int(rng.integers(len(alphabet)))])
    op = int(rng.integers(3))
    if op == 0:
        parts = [_draw(rng, alphabet, depth + 1, max_depth) for _ in range(int(rng.integers(2, 4)))]
        return "".join(f"({p})" for p in parts)
    if op == 1:
        parts = [_draw(rng, alphabet, depth + 1, max_depth) for _ in range(int(rng.integers(2, 4)))]
        return "(" + "|".join(parts) + ")"
    return "(" + _draw(rng, alphabet, depth + 1, max_depth) + ")*"


def random_regex(rng: np.random.Generator, alphabet: str, max_depth: int = 4) -> str:
    return _draw(rng, alphabet, 0, max_depth)


def random_dfa_pattern(rng: np.random.Generator, alphabet: str, max_states: int = 16,
                       min_states: int = 3, max_depth: int = 4, tries: int = 10000) -> tuple[str, DFA]:
    """Draw patterns until the byte table (live states + dead) has between
    ``min_states`` and ``max_states`` states."""
    for _ in range(tries):
        pat = random_regex(rng, alphabet, max_depth)
        dfa = compile_regex(pat)
        n = dfa_byte_table(dfa).n_states
        if min_states <= n <= max_states:
            return pat, dfa
    raise RuntimeError("no pattern within the state bounds")


def random_nfa64(rng: np.random.Generator, n_symbols: int = 32, n_states: int = 64, p: float = 0.06,
                 n_accept: int = 4, alphabet: str = ALPHABET_32) -> NFAMatrices:
    stack = (rng.random((n_symbols, n_states, n_states)) < p).astype(np.int8)   # [c, dst, src]
    start = np.zeros(n_states, np.int8)
    start[0] = 1
    accept = np.zeros(n_states, np.int8)
    accept[rng.choice(n_states, size=n_accept, replace=False)] = 1
    return NFAMatrices.from_matrices(stack, start, accept, list(alphabet[:n_symbols]))


def dirichlet_pnfa(rng: np.random.Generator, n_symbols: int = 32, n_states: int = 64, alpha: float = 0.1,
                   n_accept: int = 16):
    """Probabilistic NFA: column-stochastic T[c] ([dst][src]) whose columns are
    drawn from Dirichlet(alpha), a uniform start vector and a 0/1 accept vector
    with ``n_accept`` ones. Returns (T [σ, n, n] float32, start, accept)."""
    T = np.stack([rng.dirichlet(np.full(n_states, alpha), size=n_states).T for _ in range(n_symbols)])
    start = np.full(n_states, 1.0 / n_states, dtype=np.float32)
    accept = np.zeros(n_states, dtype=np.float32)
    accept[rng.choice(n_states, size=n_accept, replace=False)] = 1.0
    return T.astype(np.float32), start, accept


def uniform_batch(rng: np.random.Generator, n_strings: int, length: int, lo: int, hi: int):
    """``n_strings`` strings of ``length`` bytes uniform over [lo, hi)."""
    data = rng.integers(lo, hi, size=n_strings * length, dtype=np.uint8)
    offsets = np.arange(n_strings + 1, dtype=np.int64) * length
    return data, offsets


def ragged_batch(rng: np.random.Generator, lengths: np.ndarray, lo: int, hi: int):
    lengths = np.asarray(lengths, dtype=np.int64)
    offsets = np.zeros(len(lengths) + 1, dtype=np.int64)
    np.cumsum(lengths, out=offsets[1:])
    data = rng.integers(lo, hi, size=int(offsets[-1]), dtype=np.uint8)
    return data, offsets


def loguniform_lengths(rng: np.random.Generator, n: int, lo: int, hi: int) -> np.ndarray:
    return np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), size=n))).astype(np.int64)


PLAIN_DFA_SEED = 1


def plain_dfa_workload(seed: int = PLAIN_DFA_SEED, n_strings: int = 1024, length: int = 1024):
    """The plain DFA workload of the reference's CPU path: a seeded regex over
    ``abcd`` (concatenation / alternation / star, depth <= 4) whose minimal
    complete DFA has at most 16 states (bytes outside the alphabet go to the
    dead state), and ``n_strings`` strings of ``length`` bytes uniform over
    the alphabet. Returns (pattern, dfa, data, offsets)."""
    rng = np.random.default_rng(seed)
    pattern, dfa = random_dfa_pattern(rng, ALPHABET_ABCD)
    data, offsets = uniform_batch(rng, n_strings, length, ord("a"), ord("d") + 1)
    return pattern, dfa, data, offsets


def live_walk_batch(table, rng: np.random.Generator, n_strings: int, max_len: int, alphabet: str,
                    stray: float = 0.03):
    """Strings that keep a DFA byte table (``DfaByteTable``) out of its dead
    state: each byte is drawn among the alphabet bytes that lead from the
    current state to a live one (any alphabet byte with probability
    ``stray``, or when none is live), lengths uniform in [0, max_len]. Uniform
    random text dies within a few bytes for most patterns; these strings reach
    the table's other states and its accepting ones. Returns (data, offsets)."""
    alpha = np.frombuffer(alphabet.encode("latin-1"), np.uint8)
    out = []
    for _ in range(n_strings):
        s = table.start
        chars = []
        for _ in range(int(rng.integers(0, max_len + 1))):
            live = [b for b in alpha if table.delta[s, b] != table.dead]
            b = int(rng.choice(live)) if live and rng.random() >= stray else int(rng.choice(alpha))
            chars.append(b)
            s = int(table.delta[s, b])
        out.append(bytes(chars))
    offsets = np.zeros(n_strings + 1, np.int64)
    np.cumsum([len(x) for x in out], out=offsets[1:])
    return np.frombuffer(b"".join(out), np.uint8).copy(), offsets
