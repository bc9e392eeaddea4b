"""PackedEngine: several regexes over one batch (the reference's
tests/test_packed_engine.py cover the same contract: per-pattern results equal
each pattern's own simulation). A character outside a pattern's alphabet is
the identity for that pattern (src/packed_engine.py:91-105, 383-389), so the
oracle is the DFA simulation of the string with those characters removed."""

import random

import pytest

from oracle.dfa import dfa_simulate
from trex_amd.packed_engine import PackedEngine, packed_byte_table
from trex_amd.regex_to_dfa import compile_regex

BINARY = ["(a|b)*abb", "(aa|b)*", "(ab)*", "(b*ab*ab*)*b*"]
MIXED = ["(a|b)*abb", "c(a|d)*", "[a-e]*ed", "(ab|cd|ef)*(x|y)(ab|cd)*z"]


def _strings(rng, alphabet, n, max_len):
    return ["".join(rng.choice(alphabet) for _ in range(rng.randint(0, max_len))) for _ in range(n)]


def _want(regexes, strings):
    out = []
    for r in regexes:
        dfa = compile_regex(r)
        out.append([dfa_simulate(dfa, "".join(c for c in s if c in dfa.alphabet)) for s in strings])
    return out


def _walk(table, s: str) -> bool:
    q = table.start
    for b in s.encode("latin-1"):
        q = int(table.delta[q, b])
    return bool(table.accept[q])


@pytest.mark.parametrize("regexes", [BINARY, MIXED])
def test_packed_tables_follow_the_semantics(regexes):
    """The byte tables alone (walked on the CPU) give the packed semantics for
    strings over the union of the alphabets."""
    rng = random.Random(3)
    union = sorted(set().union(*(compile_regex(r).alphabet for r in regexes)))
    strings = _strings(rng, "".join(union), 300, 40)
    want = _want(regexes, strings)
    for p, r in enumerate(regexes):
        t = packed_byte_table(compile_regex(r), union)
        assert t.n_states <= 16
        assert [_walk(t, s) for s in strings] == want[p]
        # a byte outside the union goes to the absorbing, non-accepting dead state
        assert int(t.delta[t.start, 0x01]) == t.dead and not t.accept[t.dead]
        assert (t.delta[t.dead] == t.dead).all()


def test_deleting_outside_bytes():
    pe = PackedEngine(["(a|b)*abb"], use_gpu=False)
    data, offsets = pe._prepare(["xaxbby", "", "zzz", "abb"])
    assert bytes(data) == b"abbabb"
    assert offsets.tolist() == [0, 3, 3, 3, 6]


def test_config_info():
    pe = PackedEngine(["(a|b)*abb", "(aa|b)*", "(ab)*"], use_gpu=False)
    info = pe.config_info
    assert info["n_patterns"] == 3 and len(info["state_counts"]) == 3
    assert info["NP"] > 0 and info["NP"] % 16 == 0 and info["NP"] >= sum(info["state_counts"])
    assert info["offsets"][0] == 0 and info["unified_alphabet_size"] == 2
    assert info["regexes"] == ["(a|b)*abb", "(aa|b)*", "(ab)*"]
    with pytest.raises(ValueError):
        PackedEngine(["a"], gpu_mode="bogus", use_gpu=False)


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("regexes,alphabet,n,max_len", [
    (BINARY[:2], "ab", 100, 64),
    (BINARY, "ab", 200, 128),
    (BINARY[:1], "ab", 50, 64),
    (MIXED, "abcdefxyz", 400, 90),
])
def test_gpu_packed_equals_oracle(gpu_lib, regexes, alphabet, n, max_len):
    rng = random.Random(42 + n)
    strings = _strings(rng, alphabet, n, max_len)
    pe = PackedEngine(regexes)
    assert pe.config_info["backend"] == "gpu-multi"
    assert pe.match_batch(strings) == _want(regexes, strings)


@pytest.mark.gpu
def test_gpu_packed_outside_characters_are_identity(gpu_lib):
    rng = random.Random(5)
    strings = _strings(rng, "ab#!Qz\x00\xff", 500, 60) + ["", "####", "abb#", "#abb"]
    pe = PackedEngine(BINARY)
    got = pe.match_batch(strings)
    assert got == _want(BINARY, strings)
    # variable lengths, long strings
    strings = _strings(rng, "ab", 20, 5000) + ["a" * 10000 + "bb"]
    assert pe.match_batch(strings) == _want(BINARY, strings)


@pytest.mark.gpu
def test_gpu_packed_without_a_common_window(gpu_lib):
    """Alphabets in two windows (letters and digits), and a DFA over 16
    states: one launch per pattern, same semantics."""
    regexes = ["(a|b)*abb", "[0-3]*12", "(a|b)*a(a|b)(a|b)(a|b)(a|b)"]
    rng = random.Random(8)
    strings = _strings(rng, "ab0123x", 300, 50)
    pe = PackedEngine(regexes)
    assert pe.config_info["backend"] == "gpu"
    assert pe.match_batch(strings) == _want(regexes, strings)


@pytest.mark.gpu
def test_gpu_packed_empty_and_timed(gpu_lib):
    pe = PackedEngine(BINARY[:2])
    assert pe.match_batch([]) == [[], []]
    res, timing = pe.match_batch_timed(["", "abb", "aab"])
    assert res == _want(BINARY[:2], ["", "abb", "aab"])
    assert {"kernel_ms", "total_ms", "n_strings", "n_patterns", "NP"} <= set(timing)
    assert timing["n_strings"] == 3 and timing["kernel_ms"] >= 0.0
    rng = random.Random(11)
    strings = _strings(rng, "ab", 2000, 256)
    assert pe.match_batch(strings) == _want(BINARY[:2], strings)
    from trex_amd.engine import pack_strings
    acc = pe.match_arrays(*pack_strings(strings))
    assert acc.shape == (2, 2000) and [[bool(x) for x in row] for row in acc] == _want(BINARY[:2], strings)
