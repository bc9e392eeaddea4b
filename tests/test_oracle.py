"""Pins the CPU oracle before it is trusted as the checker.

Each oracle function is checked against the reference's own golden vectors
(tests/golden/reference_vectors.json), against Python's ``re`` module, against
closed forms, and against an independent formulation, chosen so that a dropped
term, a wrong index or a swapped operand fails at least one check. The
"mutation" tests plant such mistakes in copies of the oracle's inputs and show
the pins catch them.
"""

import json
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import cpu as ocpu
from oracle.dfa import dfa_final_state, dfa_simulate, simulate_monoid, table_walk, table_walk_one
from oracle.nfa import set_walk_one, simulate_nfa
from trex_amd.nfa_matrices import NFAMatrices, compile_nfa_matrices
from trex_amd.regex_to_dfa import compile_regex
from trex_amd.tables import dfa_byte_table
from trex_amd.workloads import ALPHABET_32, ALPHABET_ABCD, random_dfa_pattern, random_regex

GOLD = json.loads((Path(__file__).parent / "golden" / "reference_vectors.json").read_text())


def _golden_strings(entry):
    if entry.get("strings") is not None:
        return entry["strings"]
    return [ch * n for ch, n in entry["repeat"]]


# ---------------------------------------------------------------------------
# DFA.simulate restatement
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("entry", GOLD["regex_accept"], ids=lambda e: e["source"])
def test_dfa_simulate_golden(entry):
    dfa = compile_regex(entry["pattern"])
    got = [dfa_simulate(dfa, s) for s in _golden_strings(entry)]
    assert got == entry["expected"]


def test_dfa_state_count_bound_golden():
    for e in GOLD["dfa_state_count_max"]:
        assert compile_regex(e["pattern"]).n_states <= e["max_states"]


CLOSED_FORMS = [
    ("(b*ab*ab*)*b*", "ab", lambda s: s.count("a") % 2 == 0),
    ("(0|(1(01*0)*1))*", "01", lambda s: s == "" or int(s, 2) % 3 == 0),
    ("(a|b)*abb", "ab", lambda s: s.endswith("abb")),
    ("[a-c]*abc", "abc", lambda s: s.endswith("abc")),
    ("(ab)*", "ab", lambda s: s == "ab" * (len(s) // 2)),
]


@pytest.mark.parametrize("pattern,alpha,truth", CLOSED_FORMS, ids=[c[0] for c in CLOSED_FORMS])
def test_dfa_simulate_closed_forms(pattern, alpha, truth):
    rng = np.random.default_rng(11)
    dfa = compile_regex(pattern)
    for _ in range(400):
        n = int(rng.integers(0, 24))
        s = "".join(alpha[i] for i in rng.integers(0, len(alpha), n))
        assert dfa_simulate(dfa, s) == truth(s), s


def to_python_re(pattern: str) -> str:
    """Our syntax -> Python ``re``: ``^ $ { }`` are literals outside classes
    and a backslash is a literal inside a class."""
    out, i, in_class = [], 0, False
    while i < len(pattern):
        ch = pattern[i]
        if in_class:
            if ch == "]" and not (out and out[-1] in ("[", "[^")):
                in_class = False
                out.append("]")
            elif ch == "\\":
                out.append("\\\\")
            else:
                out.append(ch)
        elif ch == "\\" and i + 1 < len(pattern):
            out.append(pattern[i:i + 2])
            i += 1
        elif ch == "[":
            in_class = True
            if pattern[i + 1:i + 2] == "^":
                out.append("[^")
                i += 1
            else:
                out.append("[")
        elif ch in "^${}":
            out.append("\\" + ch)
        else:
            out.append(ch)
        i += 1
    return "".join(out)


def _re_check(pattern, alphabet, rng, n=150, max_len=12, extra="#"):
    dfa = compile_regex(pattern)
    rx = re.compile(f"(?:{to_python_re(pattern)})", re.DOTALL)
    pool = alphabet + extra
    for _ in range(n):
        k = int(rng.integers(0, max_len))
        s = "".join(pool[i] for i in rng.integers(0, len(pool), k))
        assert dfa_simulate(dfa, s) == (rx.fullmatch(s) is not None), (pattern, s)


@pytest.mark.parametrize("seed", range(6))
def test_dfa_simulate_vs_python_re_random_patterns(seed):
    rng = np.random.default_rng(seed)
    for _ in range(8):
        _re_check(random_regex(rng, ALPHABET_ABCD), ALPHABET_ABCD, rng)
        _re_check(random_regex(rng, ALPHABET_32), ALPHABET_32, rng, extra="a")


@pytest.mark.parametrize("pattern,alpha", [
    ("abc", "abcd"), ("a|b", "abc"), ("(ab)+", "ab"), ("[0-9]+", "0123456789a"),
    ("[a-z]+@[a-z]+\\.[a-z]+", "abz@."), ("0x[0-9a-f]+", "0x19af"), ("[a-z][a-z0-9]*", "az09_"),
    ("(if|else|while|for|return)", "ifelswhrtnu"), ("((ab|cd)(ef|gh))+", "abcdefgh"),
    ("[^abc]x", "abcdxy"), ("a.b", "ab#~"), ("\\d\\w\\s", "0a_ \t"),
])
def test_dfa_simulate_vs_python_re_library(pattern, alpha):
    _re_check(pattern, alpha, np.random.default_rng(5), n=300, max_len=10, extra="")


# ---------------------------------------------------------------------------
# Byte-table walk (the loop every GPU DFA kernel computes)
# ---------------------------------------------------------------------------
def _rand_bytes_over(rng, alphabet, n, p_other=0.05):
    out = bytearray()
    for _ in range(n):
        if rng.random() < p_other:
            out.append(int(rng.integers(0, 256)))
        else:
            out.append(ord(alphabet[int(rng.integers(len(alphabet)))]))
    return bytes(out)


@pytest.mark.parametrize("seed", range(4))
def test_table_walk_matches_dict_walk(seed):
    rng = np.random.default_rng(100 + seed)
    alpha = ALPHABET_ABCD if seed % 2 == 0 else ALPHABET_32
    _, dfa = random_dfa_pattern(rng, alpha)
    tab = dfa_byte_table(dfa)
    strings = [_rand_bytes_over(rng, alpha, int(rng.integers(0, 40))) for _ in range(200)]
    for b in strings:
        s = b.decode("latin-1")
        fs, acc = table_walk_one(tab.delta, tab.accept, tab.start, b)
        assert fs == dfa_final_state(dfa, s, tab.dead)
        assert bool(acc) == dfa_simulate(dfa, s)
    data = np.frombuffer(b"".join(strings), np.uint8)
    offs = np.concatenate([[0], np.cumsum([len(b) for b in strings])]).astype(np.int64)
    f, a = table_walk(tab.delta, tab.accept, tab.start, data, offs)
    ref = [table_walk_one(tab.delta, tab.accept, tab.start, b) for b in strings]
    assert f.tolist() == [r[0] for r in ref] and a.tolist() == [r[1] for r in ref]
    f2, a2 = ocpu.dfa_walk(tab.delta, tab.accept, tab.start, data, offs, threads=3)
    assert np.array_equal(f, f2) and np.array_equal(a, a2)


def _monoid_golden_table(e):
    """Byte table of the reference's even-a monoid test tables (index walk)."""
    m, sx = e["M"], e["sigma_ext"]
    cc = np.array(e["char_compose"], np.uint8).reshape(m, sx)
    cmap = np.full(256, e["unmapped"], np.int64)
    for ch, i in e["char_map"].items():
        cmap[ord(ch)] = i
    return np.ascontiguousarray(cc[:, cmap]), np.array(e["accept"], np.uint8), e["identity"]


def _tmap_golden_table(e):
    n = e["N"]
    delta = np.tile(np.arange(n, dtype=np.uint8)[:, None], (1, 256))
    for ch, row in e["transitions"].items():
        delta[:, ord(ch)] = row
    return np.ascontiguousarray(delta), np.array(e["accept"], np.uint8), e["start"]


def test_table_walk_monoid_and_tmap_golden():
    for e in GOLD["monoid_tables"]:
        delta, acc, start = _monoid_golden_table(e)
        got = [table_walk_one(delta, acc, start, s.encode())[1] for s in e["strings"]]
        assert got == e["expected"]
    for e in GOLD["tmap_tables"]:
        delta, acc, start = _tmap_golden_table(e)
        got = [table_walk_one(delta, acc, start, s.encode())[1] for s in e["strings"]]
        assert got == e["expected"]
        # unknown bytes are the identity in this table
        assert table_walk_one(delta, acc, start, b"aa#")[1] == 1


def test_c_oracle_dfa_golden():
    for e in GOLD["tmap_tables"]:
        delta, acc, start = _tmap_golden_table(e)
        data = np.frombuffer("".join(e["strings"]).encode(), np.uint8)
        offs = np.concatenate([[0], np.cumsum([len(s) for s in e["strings"]])]).astype(np.int64)
        _, a = ocpu.dfa_walk(delta, acc, start, data, offs, threads=4)
        assert a.tolist() == e["expected"]


def test_mutation_dfa_table_is_caught():
    """A single wrong table entry (a planted oracle mistake) fails the goldens."""
    e = GOLD["tmap_tables"][0]
    delta, acc, start = _tmap_golden_table(e)
    # the table recognises (a|b)*a(a|b): second-to-last character is 'a'
    rng = np.random.default_rng(2)
    probes = e["strings"] + ["".join("ab"[i] for i in rng.integers(0, 2, int(rng.integers(0, 12))))
                             for _ in range(200)]
    truth = [int(len(p) >= 2 and p[-2] == "a") for p in probes]
    assert [table_walk_one(delta, acc, start, p.encode())[1] for p in probes] == truth
    caught = 0
    for s_bad in range(e["N"]):
        for ch in "ab":
            bad = delta.copy()
            bad[s_bad, ord(ch)] = (bad[s_bad, ord(ch)] + 1) % e["N"]
            got = [table_walk_one(bad, acc, start, p.encode())[1] for p in probes]
            caught += got != truth
    assert caught == 2 * e["N"]


# ---------------------------------------------------------------------------
# Monoid restatement
# ---------------------------------------------------------------------------
def test_simulate_monoid_matches_dfa():
    from trex_amd.monoid import compute_monoid
    from trex_amd.simulation import DFAMatrices
    rng = np.random.default_rng(3)
    for pat in ["(a|b)*abb", "(b*ab*ab*)*b*", "(ab)*", "[a-c]*abc"]:
        dfa = compile_regex(pat)
        md = compute_monoid(DFAMatrices(dfa))
        for _ in range(100):
            s = "".join("abc"[i] for i in rng.integers(0, 3, int(rng.integers(0, 15))))
            inside = all(c in dfa.alphabet for c in s)
            if inside:
                assert simulate_monoid(md, s) == dfa_simulate(dfa, s)


# ---------------------------------------------------------------------------
# NFA restatements
# ---------------------------------------------------------------------------
def _golden_nfa(e):
    n, sig = e["n"], e["sigma"]
    stack = np.zeros((sig, n, n), np.int8)
    for sym, edges in e["edges"].items():
        for src, dst in edges:
            stack[int(sym), dst, src] = 1
    for s in range(4, n):
        stack[:, s, s] = 1
    start = np.zeros(n, np.int8)
    start[e["start"]] = 1
    acc = np.zeros(n, np.int8)
    acc[e["accept"]] = 1
    return NFAMatrices.from_matrices(stack, start, acc, ["a", "b"])


def test_nfa_golden_three_as():
    e = GOLD["nfa_tables"][0]
    nm = _golden_nfa(e)
    succ = nm.successor_masks()
    b2s = np.full(256, -1, np.int16)
    b2s[ord("a")], b2s[ord("b")] = 0, 1
    for syms, exp in zip(e["symbol_strings"], e["expected"]):
        s = "".join("ab"[c] for c in syms)
        assert int(simulate_nfa(nm, s)) == exp
        assert set_walk_one(succ, b2s, nm.start_mask(), nm.accept_bits(), s.encode())[1] == exp
    data = np.frombuffer("".join("".join("ab"[c] for c in x) for x in e["symbol_strings"]).encode(), np.uint8)
    offs = np.concatenate([[0], np.cumsum([len(x) for x in e["symbol_strings"]])]).astype(np.int64)
    _, a = ocpu.nfa_walk(succ, b2s, nm.start_mask(), nm.accept_bits(), data, offs, threads=2)
    assert a.tolist() == e["expected"]


@pytest.mark.parametrize("pattern", ["(a|b)*abb", "[a-z]+", "(ab)*", "a(b|c)*d", "(b*ab*ab*)*b*"])
def test_simulate_nfa_vs_re_and_set_walk(pattern):
    nm = compile_nfa_matrices(pattern)
    rx = re.compile(f"(?:{pattern})")
    alpha = "".join(sorted(nm.alphabet))[:8] + "#"
    rng = np.random.default_rng(7)
    succ = nm.successor_masks() if nm.n_states_raw <= 64 else None
    b2s = np.full(256, -1, np.int16)
    for i, ch in enumerate(nm.alphabet):
        b2s[ord(ch)] = i
    for _ in range(120):
        s = "".join(alpha[i] for i in rng.integers(0, len(alpha), int(rng.integers(0, 10))))
        want = rx.fullmatch(s) is not None
        assert simulate_nfa(nm, s) == want, s
        if succ is not None:
            assert bool(set_walk_one(succ, b2s, nm.start_mask(), nm.accept_bits(), s.encode())[1]) == want


def test_c_oracle_nfa_matches_python_on_random_nfa():
    from trex_amd.workloads import random_nfa64
    rng = np.random.default_rng(21)
    nm = random_nfa64(rng)
    succ = nm.successor_masks()
    b2s = np.full(256, -1, np.int16)
    b2s[0x40:0x60] = np.arange(32)
    strings = [bytes(rng.integers(0x40, 0x60, int(rng.integers(0, 30)), dtype=np.uint8)) for _ in range(100)]
    strings.append(b"AB\x00CD")      # a rejecting byte
    data = np.frombuffer(b"".join(strings), np.uint8)
    offs = np.concatenate([[0], np.cumsum([len(b) for b in strings])]).astype(np.int64)
    f, a = ocpu.nfa_walk(succ, b2s, nm.start_mask(), nm.accept_bits(), data, offs, threads=4)
    for i, b in enumerate(strings):
        fs, acc = set_walk_one(succ, b2s, nm.start_mask(), nm.accept_bits(), b)
        assert int(f[i]) == fs and int(a[i]) == acc
        assert bool(acc) == simulate_nfa(nm, b.decode("latin-1"))


def test_successor_masks_always_64_columns():
    """The C ABI reads [n_symbols][64]; small NFAs are padded, not truncated."""
    nm = compile_nfa_matrices("(a|b)*abb")
    assert nm.n_states < 64
    succ = nm.successor_masks()
    assert succ.shape == (len(nm.alphabet), 64) and succ.dtype == np.uint64
    assert not succ[:, nm.n_states:].any()
    for c in range(len(nm.alphabet)):
        for src in range(nm.n_states):
            for dst in range(nm.n_states):
                assert bool(int(succ[c, src]) >> dst & 1) == bool(nm.matrix_stack[c, dst, src])


def test_mutation_nfa_successor_is_caught():
    e = GOLD["nfa_tables"][0]
    nm = _golden_nfa(e)
    succ = nm.successor_masks()
    b2s = np.full(256, -1, np.int16)
    b2s[ord("a")], b2s[ord("b")] = 0, 1
    bad = succ.copy()
    bad[0, 2] &= ~np.uint64(1 << 3)       # drop the edge 2 -a-> 3
    got = [set_walk_one(bad, b2s, nm.start_mask(), nm.accept_bits(),
                        "".join("ab"[c] for c in x).encode())[1] for x in e["symbol_strings"]]
    assert got != e["expected"]
