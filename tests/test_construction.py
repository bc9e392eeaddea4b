"""Pins of the host construction API against the reference's own tests.

Covers the k-gram selection and precomputation (``src/kgram.py:29-232``;
reference tests ``tests/test_kgram.py``), the monoid-batch table set
(``src/monoid.py:195-287``; reference tests ``tests/test_monoid.py:449-531``)
and the reference's construction-type methods ``DFA.simulate``,
``DFAMatrices.get_matrix_sequence`` / ``check_accept`` and
``NFAMatrices.check_accept``. Every check compares against something fixed
outside the code under test: the reference's own expected values (the golden
fixture), Python's ``re`` module, closed forms, or an independent sequential
composition written in the test.
"""

import itertools
import json
import re
from pathlib import Path

import numpy as np
import pytest

from oracle.dfa import dfa_simulate, simulate_kgram_monoid, simulate_monoid
from oracle.nfa import simulate_nfa
from trex_amd.kgram import KGramTable, auto_k, auto_k_for_gpu, precompute_kgrams
from trex_amd.monoid import compute_monoid, precompute_batch_tables
from trex_amd.nfa_matrices import compile_nfa_matrices
from trex_amd.regex_to_dfa import compile_regex
from trex_amd.simulation import DFAMatrices

GOLD = json.loads((Path(__file__).parent / "golden" / "reference_vectors.json").read_text())

# The reference's named test patterns (src/generate_data.py:38-100) and the
# alphabets its tests draw from (tests/test_monoid.py:40-49).
PATTERNS = {
    "abb": ("(a|b)*abb", "ab"),
    "even_a": ("(b*ab*ab*)*b*", "ab"),
    "binary_div3": ("(0|(1(01*0)*1))*", "01"),
    "identifier": ("[a-z][a-z0-9]*", "abcdefghijklmnopqrstuvwxyz0123456789"),
    "abc_suffix": ("[a-c]*abc", "abc"),
    "abcd_pairs": ("((ab|cd)(ef|gh))+", "abcdefgh"),
}

# Closed forms of the patterns above, independent of any automaton.
TRUTH = {
    "abb": lambda s: s.endswith("abb"),
    "even_a": lambda s: s.count("a") % 2 == 0,
    "binary_div3": lambda s: s == "" or int(s, 2) % 3 == 0,
    "identifier": lambda s: re.fullmatch("[a-z][a-z0-9]*", s) is not None,
    "abc_suffix": lambda s: s.endswith("abc"),
    "abcd_pairs": lambda s: re.fullmatch("((ab|cd)(ef|gh))+", s) is not None,
}


def _rand_str(rng, alpha, n):
    return "".join(alpha[i] for i in rng.integers(0, len(alpha), n))


def _setup(name):
    regex, alpha = PATTERNS[name]
    dfa = compile_regex(regex)
    dm = DFAMatrices(dfa)
    md = compute_monoid(dm)
    assert md is not None
    return dfa, dm, md, alpha


# ---------------------------------------------------------------------------
# auto_k / auto_k_for_gpu: the reference's expected values
# ---------------------------------------------------------------------------
def _gold_cases(prefix):
    for entry in GOLD["kgram_auto_k"]:
        if entry["source"].startswith(prefix):
            return entry["cases"]
    raise KeyError(prefix)


@pytest.mark.parametrize("sigma,max_entries,k", _gold_cases("tests/test_kgram.py:79-98"))
def test_auto_k_reference_values(sigma, max_entries, k):
    assert auto_k(sigma, max_entries) == k


@pytest.mark.parametrize("sigma,n,budget,k", _gold_cases("tests/test_kgram.py:321-343"))
def test_auto_k_for_gpu_reference_values(sigma, n, budget, k):
    assert auto_k_for_gpu(sigma, n, budget) == k


def test_auto_k_closed_form_and_monotone():
    # largest k with sigma**k <= max_entries, by brute force over k
    for sigma in range(2, 300):
        for cap in (16, 255, 256, 4096, 65536):
            best = max(k for k in range(1, 40) if k == 1 or sigma ** k <= cap)
            assert auto_k(sigma, cap) == best, (sigma, cap)
    ks = [auto_k(s) for s in (2, 4, 8, 16, 32, 64, 128, 256)]
    assert ks == sorted(ks, reverse=True)
    assert auto_k(32) == 3                      # 32**3 = 32768 <= 65536 < 32**4
    for n in (2, 16, 48, 64):
        for sigma in (2, 3, 4, 32, 256):
            k = auto_k_for_gpu(sigma, n)
            assert k == 1 or sigma ** k * n * n <= 48_000_000
            assert sigma ** (k + 1) * n * n > 48_000_000
    ks = [auto_k_for_gpu(2, n) for n in (16, 32, 48, 64)]
    assert ks == sorted(ks, reverse=True)


# ---------------------------------------------------------------------------
# precompute_kgrams, both modes, and the oracle's k-gram walk
# ---------------------------------------------------------------------------
def _seq_compose(md, gram):
    acc = md.identity_idx
    for ch in gram:
        acc = int(md.compose_table[md.char_to_monoid[ch], acc])
    return acc


@pytest.mark.parametrize("k", [2, 3, 4, 8])
def test_kgram_monoid_table_matches_sequential_compose(k):
    _, dm, md, alpha = _setup("abb")
    kg = precompute_kgrams(dm, k, monoid=md)
    assert len(kg._monoid_table) == len(alpha) ** k
    for gram in itertools.product(dm.alphabet, repeat=k):
        assert kg.lookup(gram) == _seq_compose(md, gram)
        assert 0 <= kg.lookup(gram) < md.size


def test_kgram_key_is_first_symbol_most_significant():
    _, dm, md, _ = _setup("abb")
    kg = precompute_kgrams(dm, 3, monoid=md)
    # key of (c0, c1, c2) = (c0 * sigma + c1) * sigma + c2 with a=0, b=1
    assert kg._gram_to_key(("b", "a", "a")) == 4
    assert kg._gram_to_key(("a", "a", "b")) == 1
    assert kg._monoid_table[4] == _seq_compose(md, "baa")


@pytest.mark.parametrize("name", ["abb", "even_a", "binary_div3", "abc_suffix"])
@pytest.mark.parametrize("k,tail", [(k, t) for k in (2, 3, 4) for t in range(k)])
def test_simulate_kgram_monoid_tails(name, k, tail):
    """The oracle's k-gram walk (whole k-grams by lookup, the tail character
    by character) decides as the closed form and the DFA walk do, for every
    tail length 0..k-1."""
    dfa, dm, md, alpha = _setup(name)
    kg = precompute_kgrams(dm, k, monoid=md)
    rng = np.random.default_rng(100 * k + tail)
    for _ in range(150):
        grams = int(rng.integers(0, 8))
        s = _rand_str(rng, alpha, grams * k + tail)
        want = TRUTH[name](s)
        assert simulate_kgram_monoid(kg, md, s) == want, s
        assert dfa_simulate(dfa, s) == want, s


def test_simulate_kgram_monoid_empty_and_unknown():
    dfa, dm, md, _ = _setup("abb")
    kg = precompute_kgrams(dm, 4, monoid=md)
    assert simulate_kgram_monoid(kg, md, "") is False
    _, dm2, md2, _ = _setup("even_a")
    kg2 = precompute_kgrams(dm2, 2, monoid=md2)
    assert simulate_kgram_monoid(kg2, md2, "") is True
    # a character outside the alphabet is the identity in the monoid walks
    # (src/monoid.py:185-190): it may only appear in the tail
    assert simulate_kgram_monoid(kg2, md2, "abab" + "z") == simulate_monoid(md2, "abab")


@pytest.mark.parametrize("k", [2, 3, 4])
def test_kgram_matrix_mode(k):
    dfa, dm, _, alpha = _setup("abb")
    kg = precompute_kgrams(dm, k, monoid=None)
    assert isinstance(kg, KGramTable) and len(kg._matrix_table) == len(alpha) ** k
    for gram in itertools.product(dm.alphabet, repeat=k):
        acc = dm.identity_matrix().astype(np.int32)
        for ch in gram:
            acc = dm.matrices[ch].astype(np.int32) @ acc      # newer symbol on the left
        assert np.array_equal(kg.lookup_matrix(gram), acc.astype(np.int8)), gram
    rng = np.random.default_rng(77 + k)
    for _ in range(60):
        s = _rand_str(rng, alpha, k * int(rng.integers(1, 6)))
        acc = dm.identity_matrix().astype(np.int32)
        for pos in range(0, len(s), k):
            acc = kg.lookup_matrix(tuple(s[pos:pos + k])).astype(np.int32) @ acc
        final = acc @ dm.start_vec.astype(np.int32)
        assert dm.check_accept(final.astype(np.int8)) == TRUTH["abb"](s), s


def test_kgram_oracle_catches_wrong_composition_order():
    """A k-gram table composed oldest-on-the-left (a plausible slip) is caught
    by the walk checks above: build one and show it disagrees."""
    _, dm, md, alpha = _setup("abb")
    k = 3
    kg = precompute_kgrams(dm, k, monoid=md)
    swapped = {}
    for key, gram in enumerate(itertools.product(dm.alphabet, repeat=k)):
        acc = md.identity_idx
        for ch in gram:
            acc = int(md.compose_table[acc, md.char_to_monoid[ch]])
        swapped[key] = acc
    bad = KGramTable(k, dm.alphabet, monoid_table=swapped)
    rng = np.random.default_rng(5)
    wrong = 0
    for _ in range(200):
        s = _rand_str(rng, alpha, 3 * int(rng.integers(1, 6)))
        assert simulate_kgram_monoid(kg, md, s) == TRUTH["abb"](s)
        wrong += simulate_kgram_monoid(bad, md, s) != TRUTH["abb"](s)
    assert wrong > 0


# ---------------------------------------------------------------------------
# precompute_batch_tables (reference tests/test_monoid.py:449-531)
# ---------------------------------------------------------------------------
def _replay(tables, s):
    sig_ext = tables["sigma_ext"]
    cur = tables["identity_idx"]
    for ch in s:
        cur = int(tables["char_compose"][cur * sig_ext + int(tables["raw_char_map"][ord(ch)])])
    return bool(tables["accept"][cur])


@pytest.mark.parametrize("name", ["abb", "even_a", "identifier", "binary_div3", "abcd_pairs"])
def test_batch_tables_reproduce_simulate_monoid(name):
    dfa, dm, md, alpha = _setup(name)
    t = precompute_batch_tables(md, dm)
    rng = np.random.default_rng(42)
    for _ in range(200):
        s = _rand_str(rng, alpha, int(rng.integers(0, 100)))
        want = simulate_monoid(md, s)
        assert want == TRUTH[name](s) == dfa_simulate(dfa, s), s
        assert _replay(t, s) == want, s
        # the fused byte table walks the same way: fused[m * 256 + byte]
        cur = t["identity_idx"]
        for ch in s:
            cur = int(t["fused_compose"][cur * 256 + ord(ch)])
        assert bool(t["accept"][cur]) == want, s


@pytest.mark.parametrize("name", ["abb", "even_a", "identifier"])
def test_batch_tables_identity_column_and_compose(name):
    _, dm, md, _ = _setup(name)
    t = precompute_batch_tables(md, dm)
    m, sig_ext = t["M"], t["sigma_ext"]
    z = t["raw_char_map"][ord("~")]             # a byte outside every alphabet
    assert z == sig_ext - 1 == len(dm.alphabet)
    assert all(t["char_compose"][i * sig_ext + z] == i for i in range(m))
    assert np.array_equal(t["monoid_compose"].reshape(m, m), md.compose_table.astype(np.uint8))
    assert t["char_compose"].shape == (m * sig_ext,) and t["char_compose"].dtype == np.uint8
    assert t["raw_char_map"].shape == (256,) and t["raw_char_map"].dtype == np.uint8
    assert t["accept"].shape == (m,) and t["accept"].dtype == np.uint8
    assert t["monoid_compose"].shape == (m * m,) and t["monoid_compose"].dtype == np.uint8
    assert t["fused_compose"].shape == (m * 256,)


@pytest.mark.parametrize("name", ["abb", "binary_div3", "abc_suffix"])
def test_batch_tables_kgram_compose(name):
    """kgram_compose[m * sigma^k + g] is the element after walking k-gram g
    (first symbol most significant) from element m; k = floor(log_sigma 256)."""
    _, dm, md, _ = _setup(name)
    t = precompute_batch_tables(md, dm)
    sigma, k = t["sigma"], t["kgram_k"]
    assert sigma ** k <= 256 < sigma ** (k + 1)
    assert t["sigma_k"] == sigma ** k
    kg = t["kgram_compose"]
    assert kg is not None and kg.shape == (t["M"] * sigma ** k,)
    for m in range(t["M"]):
        for g, gram in enumerate(itertools.product(range(sigma), repeat=k)):
            cur = m
            for c in gram:
                cur = int(t["char_compose"][cur * t["sigma_ext"] + c])
            assert kg[m * sigma ** k + g] == cur


def test_batch_tables_no_kgram_for_wide_alphabets():
    # sigma = 36 > 16: k = 1, no k-gram table (src/monoid.py:240-251)
    _, dm, md, _ = _setup("identifier")
    t = precompute_batch_tables(md, dm)
    assert t["kgram_k"] == 1 and t["kgram_compose"] is None


# ---------------------------------------------------------------------------
# DFA.simulate, DFAMatrices, NFAMatrices methods of the reference's API
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("entry", GOLD["regex_accept"], ids=lambda e: e["source"])
def test_dfa_simulate_method_golden(entry):
    strings = entry["strings"] if entry.get("strings") is not None else [c * n for c, n in entry["repeat"]]
    dfa = compile_regex(entry["pattern"])
    assert [dfa.simulate(s) for s in strings] == entry["expected"]


@pytest.mark.parametrize("name", sorted(PATTERNS))
def test_dfa_simulate_method_closed_forms(name):
    regex, alpha = PATTERNS[name]
    dfa = compile_regex(regex)
    rng = np.random.default_rng(3)
    for _ in range(300):
        s = _rand_str(rng, alpha + "z%", int(rng.integers(0, 20)))
        assert dfa.simulate(s) == (re.fullmatch(regex, s) is not None), s


def test_dfa_simulate_method_without_dead_state():
    dfa = compile_regex("(a|b)*abb")
    dfa.dead_state = None                        # a character with no edge rejects at once
    assert dfa.simulate("aabb") is True
    assert dfa.simulate("aazbb") is False


@pytest.mark.parametrize("name", ["abb", "even_a", "abcd_pairs"])
def test_matrix_sequence_and_check_accept(name):
    regex, alpha = PATTERNS[name]
    dfa = compile_regex(regex)
    dm = DFAMatrices(dfa)
    rng = np.random.default_rng(9)
    for _ in range(100):
        s = _rand_str(rng, alpha + "z", int(rng.integers(0, 16)))
        seq = dm.get_matrix_sequence(s)
        assert seq.shape == (len(s), dm.n_states, dm.n_states) and seq.dtype == np.int8
        vec = dm.start_vec.astype(np.int32)
        for i, ch in enumerate(s):
            if ch in dm.char_to_idx:
                assert np.array_equal(seq[i], dm.matrices[ch])
            else:                                    # identity (src/simulation.py:88-91)
                assert np.array_equal(seq[i], np.eye(dm.n_states, dtype=np.int8))
            vec = seq[i].astype(np.int32) @ vec
        # unknown characters are skipped: the monoid walk's semantics
        assert dm.check_accept(vec.astype(np.int8)) == TRUTH[name](s.replace("z", "")), s
    assert dm.get_matrix_sequence("").shape == (0, dm.n_states, dm.n_states)


def test_check_accept_ignores_padded_states():
    dm = DFAMatrices(compile_regex("(a|b)*abb"), pad_to=32)
    vec = np.zeros(dm.n_states, np.int8)
    vec[dm.n_states_raw:] = 1
    dm.accept_mask[dm.n_states_raw:] = 1             # even if a padded slot were marked
    assert dm.check_accept(vec) is False


@pytest.mark.parametrize("regex,alpha", [("(a|b)*a(a|b)", "ab"), ("(ab|a)*b+", "ab"),
                                         ("[a-c]*abc", "abc"), ("((ab|cd)(ef|gh))+", "abcdefgh")])
def test_nfa_check_accept(regex, alpha):
    nm = compile_nfa_matrices(regex)
    rng = np.random.default_rng(4)
    for _ in range(200):
        s = _rand_str(rng, alpha, int(rng.integers(0, 14)))
        vec = nm.start_vec.astype(np.int32)
        for ch in s:
            vec = np.minimum(nm.matrices[ch].astype(np.int32) @ vec, 1)
        want = re.fullmatch(regex, s) is not None
        assert nm.check_accept(vec) == want, s
        assert simulate_nfa(nm, s) == want, s
