"""OptimizedEngine: API contract on the CPU (selection, config_info), results
on the GPU against the oracle semantics of each configuration (the reference's
tests/test_optimized_engine.py:59-455 cover the same contract)."""

import re

import numpy as np
import pytest

from oracle.dfa import dfa_simulate, simulate_monoid
from oracle.nfa import simulate_nfa
from trex_amd.optimized_engine import CONFIGS, OptimizedEngine

REQUIRED = {"representation", "scan_backend", "alphabet_size", "kgram_k", "selection_reason"}


@pytest.mark.parametrize("config", CONFIGS)
def test_config_info_keys(config):
    eng = OptimizedEngine("(a|b)*abb", config=config)
    info = eng.config_info
    assert REQUIRED <= set(info)
    assert info["representation"] in ("dfa", "nfa")
    assert isinstance(info["selection_reason"], str) and info["selection_reason"]


def test_unknown_config_raises():
    with pytest.raises(ValueError):
        OptimizedEngine("a", config="bogus")


def test_auto_selection_families():
    assert OptimizedEngine("(a|b)*abb").config_info["scan_backend"] == "dfa16+gpu"
    wide = OptimizedEngine("(a|b)*a(a|b)(a|b)(a|b)(a|b)(a|b)")
    assert wide.config_info["scan_backend"] == "dfa+gpu" and wide.config_info["dfa_states"] > 16
    capped = OptimizedEngine("(a|b)*abb", dfa_state_cap=2)
    assert capped.config_info["representation"] == "nfa"


def test_monoid_cap_raises():
    with pytest.raises(RuntimeError):
        OptimizedEngine("(a|b)*a(a|b)(a|b)(a|b)", config="monoid", monoid_cap=4)


def test_monoid_batch_limit():
    eng = OptimizedEngine("(b*ab*ab*)*b*", config="monoid_batch+gpu")
    assert eng.config_info["monoid_size"] == 2


# ---------------------------------------------------------------------------
# GPU: every configuration against its reference semantics
# ---------------------------------------------------------------------------
def _strings(alpha, rng, n=300, max_len=30):
    return ["".join(alpha[i] for i in rng.integers(0, len(alpha), int(rng.integers(0, max_len))))
            for _ in range(n)]


@pytest.mark.gpu
@pytest.mark.parametrize("config", CONFIGS)
@pytest.mark.parametrize("pattern", ["(a|b)*abb", "(b*ab*ab*)*b*", "[a-c]*abc"])
def test_configs_match_their_semantics(gpu_lib, config, pattern):
    rng = np.random.default_rng(abs(hash((config, pattern))) % 2**32)
    eng = OptimizedEngine(pattern, config=config)
    strings = _strings("abcx", rng)
    got = eng.match_batch(strings)
    dfa = eng._dfa
    if config in ("monoid", "monoid+kgram", "monoid+gpu", "monoid_batch+gpu"):
        want = [simulate_monoid(eng._md, s) for s in strings]
    elif config == "prefix+gpu":
        # unknown characters are the identity: drop them, then the DFA decides
        want = [dfa_simulate(dfa, "".join(c for c in s if c in dfa.alphabet)) for s in strings]
    elif config == "nfa":
        want = [simulate_nfa(eng._nm, s) for s in strings]
    else:
        want = [dfa_simulate(dfa, s) for s in strings]
    assert got == want
    assert eng.match(strings[:10]) == want[:10]


@pytest.mark.gpu
def test_all_configs_agree_inside_alphabet(gpu_lib):
    rng = np.random.default_rng(12)
    strings = _strings("ab", rng, n=200)
    ref = [re.fullmatch("(a|b)*abb", s) is not None for s in strings]
    for config in CONFIGS:
        assert OptimizedEngine("(a|b)*abb", config=config).match_batch(strings) == ref, config


@pytest.mark.gpu
def test_timed_and_long_strings(gpu_lib):
    eng = OptimizedEngine("(b*ab*ab*)*b*")
    rng = np.random.default_rng(5)
    long_s = "".join("ab"[i] for i in rng.integers(0, 2, (1 << 20) + 4097))
    strings = [long_s, "aa", "a", ""]
    res, timing = eng.match_batch_timed(strings)
    assert res == [long_s.count("a") % 2 == 0, True, False, True]
    assert timing["kernel_ms"] >= 0 and timing["total_ms"] >= 0
    assert eng.match_batch([]) == []


def test_dfa_as_nfa_walks_as_the_table():
    """The NFA-64 form of a DFA table (for long strings of DFAs wider than 16
    states) reaches, from the start set, exactly the singleton of the table
    walk's state, or the empty set where the walk reaches the dead state."""
    from oracle.dfa import table_walk_one
    from oracle.nfa import set_walk_one
    from trex_amd.engine import dfa_as_nfa, nfa_mask_to_state
    from trex_amd.regex_to_dfa import compile_regex
    from trex_amd.tables import dfa_byte_table
    rng = np.random.default_rng(2)
    for pattern in ("(a|b)*a(a|b)(a|b)(a|b)(a|b)", "(a|b)*abb", "[a-c]*abc(a|c)*"):
        tab = dfa_byte_table(compile_regex(pattern))
        succ, b2s, start, accept = dfa_as_nfa(tab)
        assert (b2s >= 0).sum() == len(compile_regex(pattern).alphabet)
        for _ in range(200):
            s = bytes(rng.choice(np.frombuffer(b"abcx", np.uint8), int(rng.integers(0, 40))))
            state, acc = table_walk_one(tab.delta, tab.accept, tab.start, s)
            mask, nacc = set_walk_one(succ, b2s, start, accept, s)
            assert mask == 0 or mask & (mask - 1) == 0
            assert nfa_mask_to_state(np.array([mask], np.uint64), tab.dead)[0] == state
            assert nacc == acc
    wide = dfa_byte_table(compile_regex("(a|b)*a(a|b)(a|b)(a|b)(a|b)(a|b)(a|b)"))
    assert wide.n_states > 64 and dfa_as_nfa(wide) is None


@pytest.mark.gpu
@pytest.mark.parametrize("pattern,config", [("(a|b)*a(a|b)(a|b)", "nfa"),
                                            ("(a|b)*a(a|b)(a|b)(a|b)(a|b)", None),
                                            ("(b*ab*ab*)*b*", "nfa")])
def test_long_strings_split_for_nfa_and_wide_dfa(gpu_lib, pattern, config):
    """Strings over 1 MiB for an NFA and for a DFA of more than 16 states go
    through speculative chunks and fix-ups (the mixed queue with one pattern)
    and agree with the oracle's walks, final sets / states included."""
    from oracle import cpu as ocpu
    from trex_amd.engine import NfaEngine, pack_strings
    eng = OptimizedEngine(pattern, config=config)
    rng = np.random.default_rng(31)
    lengths = [(1 << 20) + 5, 3 << 20, 2_500_001, 17, 0, 70_000]
    strings = [bytes(rng.choice(np.frombuffer(b"ab", np.uint8), n)) for n in lengths]
    strings[3] = b"abx" + strings[3]          # a byte outside the alphabet
    data, offs = pack_strings(strings)
    final, acc = eng.match_arrays(data, offs)
    if config == "nfa":
        nm = eng._nm
        ref = NfaEngine.from_matrices(nm)
        assert ref.can_split()
        rf, ra = ocpu.nfa_walk(ref.succ, ref.byte_to_sym, ref.start, ref.accept, data, offs)
        if pattern == "(b*ab*ab*)*b*":
            assert [bool(a) for a in acc] == [s.count(b"a") % 2 == 0 and b"x" not in s for s in strings]
    else:
        tab = eng._table
        assert tab.n_states > 16 and eng._nfa_form() is not None
        rf, ra = ocpu.dfa_walk(tab.delta, tab.accept, tab.start, data, offs)
    assert np.array_equal(np.asarray(final, np.uint64), np.asarray(rf, np.uint64))
    assert np.array_equal(acc, ra)
    res, timing = eng.match_batch_timed(strings)
    assert res == [bool(a) for a in ra] and 0 < timing["kernel_ms"] <= timing["total_ms"]
