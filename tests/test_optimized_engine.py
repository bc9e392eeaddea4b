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
