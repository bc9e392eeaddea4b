"""GPU parity of the bit-parallel Boolean NFA-64 kernels (final state mask and
accept bit, bit-exact) against the CPU oracle."""

import ctypes
import json
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import cpu as ocpu
from oracle.nfa import set_walk_one, simulate_nfa
from trex_amd.engine import NfaEngine, nfa_byte_map, pack_strings
from trex_amd.nfa_matrices import compile_nfa_matrices
from trex_amd.workloads import random_nfa64, ragged_batch, uniform_batch

pytestmark = pytest.mark.gpu
GOLD = json.loads((Path(__file__).parent / "golden" / "reference_vectors.json").read_text())


def _check(eng: NfaEngine, data, offs):
    mask, acc, _, _ = eng.match_arrays(data, offs)
    rm, ra = ocpu.nfa_walk(eng.succ, eng.byte_to_sym, eng.start, eng.accept, data, offs)
    bad = np.nonzero((mask != rm) | (acc != ra))[0]
    assert bad.size == 0, f"{bad.size} mismatches at {bad[:5]}"
    return mask, acc


def test_reference_compat_three_as_golden(gpu_lib):
    """nfa_tc_engine_* with the reference's ">= 3 a's" NFA and its 10 expected bits."""
    from trex_amd import _lib
    e = GOLD["nfa_tables"][0]
    n, sig = e["n"], e["sigma"]
    T = np.zeros((sig, n, n), np.float32)
    for sym, edges in e["edges"].items():
        for src, dst in edges:
            T[int(sym), dst, src] = 1.0
    for s in range(4, n):
        T[:, s, s] = 1.0
    acc = np.zeros(n, np.float32)
    acc[e["accept"]] = 1.0
    start = np.zeros(n, np.float32)
    start[e["start"]] = 1.0
    _lib.call("nfa_tc_engine_init", T.ctypes.data_as(_lib.f32p), acc.ctypes.data_as(_lib.f32p),
              start.ctypes.data_as(_lib.f32p), n, sig, 64, 256)
    flat = np.array([c for s in e["symbol_strings"] for c in s], np.uint8)
    offs = np.concatenate([[0], np.cumsum([len(s) for s in e["symbol_strings"]])]).astype(np.int32)
    res = np.zeros(len(e["symbol_strings"]), np.int32)
    km, tm = ctypes.c_float(), ctypes.c_float()
    _lib.call("nfa_tc_engine_dispatch", flat.ctypes.data_as(_lib.u8p), offs.ctypes.data_as(_lib.intp),
              res.ctypes.data_as(_lib.intp), len(res), int(offs[-1]), ctypes.byref(km), ctypes.byref(tm))
    assert res.tolist() == e["expected"]
    _lib.lib().nfa_tc_engine_destroy()


@pytest.mark.parametrize("pattern", ["(a|b)*abb", "(ab)*", "a(b|c)*d", "[a-d]*a[a-d][a-d]", "(b*ab*ab*)*b*"])
def test_regex_nfas_vs_re_and_oracle(gpu_lib, pattern):
    nm = compile_nfa_matrices(pattern)
    eng = NfaEngine.from_matrices(nm)
    rng = np.random.default_rng(len(pattern))
    alpha = "".join(sorted(nm.alphabet)) + "#"
    strings = ["".join(alpha[i] for i in rng.integers(0, len(alpha), int(rng.integers(0, 40))))
               for _ in range(400)]
    data, offs = pack_strings(strings)
    mask, acc = _check(eng, data, offs)
    for v in ("window5", "window6"):               # 5-bit and 6/5-bit slice tables
        eng.set_variant(v)
        assert np.array_equal(_check(eng, data, offs)[0], mask)
    rx = re.compile(f"(?:{pattern})")
    assert [bool(x) for x in acc] == [rx.fullmatch(s) is not None for s in strings]
    assert [bool(x) for x in acc[:60]] == [simulate_nfa(nm, s) for s in strings[:60]]


@pytest.mark.parametrize("seed", range(3))
def test_random_nfa64_32_symbols_both_variants(gpu_lib, seed):
    rng = np.random.default_rng(500 + seed)
    nm = random_nfa64(rng)
    eng = NfaEngine.from_matrices(nm)
    assert eng.info()["variant"] == "window6"
    eng.set_variant("window")
    lengths = rng.integers(0, 600, 3000)
    lengths[:6] = [0, 1, 15, 16, 17, 33]
    data, offs = ragged_batch(rng, lengths, 0x40, 0x60)
    data[rng.integers(0, len(data), 30)] = rng.integers(0, 256, 30, dtype=np.uint8)
    first = _check(eng, data, offs)
    eng.set_variant("bytes")
    second = _check(eng, data, offs)
    assert np.array_equal(first[0], second[0])
    for v in ("window5", "window6"):
        eng.set_variant(v)
        assert np.array_equal(first[0], _check(eng, data, offs)[0])
    # spot-check the oracle against the set-of-states restatement
    for i in range(0, 3000, 500):
        b = bytes(data[offs[i]:offs[i + 1]])
        assert set_walk_one(eng.succ, eng.byte_to_sym, eng.start, eng.accept, b)[0] == int(first[0][i])


def test_skip_bytes_general_map(gpu_lib):
    """-2 (skip) bytes leave the state unchanged; symbols outside one window."""
    rng = np.random.default_rng(77)
    nm = random_nfa64(rng, n_symbols=40, alphabet="".join(chr(c) for c in range(0x30, 0x58)))
    b2s = nfa_byte_map(nm, unknown="skip")
    eng = NfaEngine(nm.successor_masks(), b2s, nm.start_mask(), nm.accept_bits())
    assert eng.info()["variant"] == "bytes"
    data, offs = ragged_batch(rng, rng.integers(0, 300, 1000), 0x20, 0x60)
    _check(eng, data, offs)


@pytest.mark.parametrize("variant", ["window", "window5", "window6"])
def test_uniform_batch_bit_exact(gpu_lib, variant):
    rng = np.random.default_rng(3)
    nm = random_nfa64(rng)
    eng = NfaEngine.from_matrices(nm)
    eng.set_variant(variant)
    data, offs = uniform_batch(rng, 8192, 1024, 0x40, 0x60)
    _check(eng, data, offs)


def test_window5_skip_and_reject_bytes_inside_the_window(gpu_lib):
    """Window bytes that skip (identity) or reject (empty set), and a
    sparse 40-state automaton, through the 5-bit slice tables."""
    rng = np.random.default_rng(91)
    nm = random_nfa64(rng, n_symbols=26, n_states=40, p=0.1, alphabet="".join(chr(c) for c in range(0x40, 0x5a)))
    b2s = np.full(256, -1, np.int16)
    b2s[0x40:0x5a] = np.arange(26)
    b2s[0x5a:0x5e] = -2                            # skip
    eng = NfaEngine(nm.successor_masks(), b2s, nm.start_mask(), nm.accept_bits())
    assert eng.info()["variant"] == "window6"
    lengths = rng.integers(0, 400, 2000)
    lengths[:5] = [0, 1, 15, 16, 31]
    data, offs = ragged_batch(rng, lengths, 0x40, 0x60)
    data[rng.integers(0, len(data), 20)] = 0x21    # outside the window: rejects
    _check(eng, data, offs)
    for v in ("window6", "window"):
        eng.set_variant(v)
        _check(eng, data, offs)
