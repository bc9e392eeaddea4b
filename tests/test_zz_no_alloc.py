"""The allocation-free look-back entry point, trex_dfa_match_chunked_static,
called through the C ABI: its final states and accept bits equal those of the
plan-based entry point (trex_plan_create + trex_dfa_match_plan) and of the C
oracle's byte walk, bit for bit, on two calls in a row (the second with a
smaller input, so stale contents of the shared static array would show), and
its size limits and range checks refuse what it cannot hold."""

import ctypes

import numpy as np
import pytest

from oracle import cpu as ocpu
from trex_amd.regex_to_dfa import compile_regex
from trex_amd.tables import dfa_byte_table
from trex_amd.workloads import ALPHABET_32, random_dfa_pattern

MAX_STRINGS = 65536
MAX_CHUNKS = 131072


def _static_call(eng, d_bytes, bytes_len, d_begin, d_end, n, chunk_len, d_f, d_a, d_err):
    from trex_amd import _lib
    return _lib.lib().trex_dfa_match_chunked_static(eng.handle, d_bytes, bytes_len, d_begin, d_end, n, chunk_len,
                                                    d_f, d_a, d_err, 0, None)


def _both_forms(tab, data, offs, chunk_len):
    """(static final, static accept, error word, plan final, plan accept)."""
    from trex_amd import _lib
    from trex_amd.device import DeviceBuffer
    from trex_amd.engine import ChunkPlan, DfaEngine
    eng = DfaEngine(tab)
    n = len(offs) - 1
    begin = np.ascontiguousarray(offs[:-1], np.int64)
    end = np.ascontiguousarray(offs[1:], np.int64)
    d_bytes = DeviceBuffer.from_array(data if len(data) else np.zeros(16, np.uint8))
    d_begin, d_end = DeviceBuffer.from_array(begin), DeviceBuffer.from_array(end)
    # outputs pre-filled with a value no result takes, so unwritten slots show
    d_f = DeviceBuffer.from_array(np.full(max(n, 1), 0xAB, np.uint8))
    d_a = DeviceBuffer.from_array(np.full(max(n, 1), 0xAB, np.uint8))
    d_err = DeviceBuffer.from_array(np.array([0xDEADBEEF], np.uint32))
    rc = _static_call(eng, d_bytes.ptr, len(data), d_begin.ptr, d_end.ptr, n, chunk_len, d_f.ptr, d_a.ptr, d_err.ptr)
    assert rc == 0, _lib.last_error()
    _lib.call("trex_device_sync")
    sf, sa = d_f.download(np.uint8, n), d_a.download(np.uint8, n)
    err = int(d_err.download(np.uint32, 1)[0])
    plan = ChunkPlan(begin, end, chunk_len)
    p_f, p_a = DeviceBuffer(max(n, 1)), DeviceBuffer(max(n, 1))
    eng.match_plan(plan, d_bytes.ptr, p_f.ptr, p_a.ptr)
    assert plan.error() == 0
    return sf, sa, err, p_f.download(np.uint8, n), p_a.download(np.uint8, n)


def _mod15():
    """Sum of the low nibbles mod 15 over the 0x40-0x5F window (a permutation
    automaton on its alphabet: every chunk keeps composing), other bytes to
    the dead state 15."""
    from trex_amd.tables import DfaByteTable
    delta = np.full((16, 256), 15, np.uint8)
    for b in range(0x40, 0x60):
        delta[:15, b] = (np.arange(15) + (b & 15)) % 15
    return DfaByteTable(delta, np.eye(16, dtype=np.uint8)[0], 0, 15, 16)


def _case(seed, n_short, long_lengths, chunk_len, alphabet=ALPHABET_32, pattern=None):
    rng = np.random.default_rng(seed)
    if pattern == "mod15":
        tab = _mod15()
    else:
        tab = dfa_byte_table(compile_regex(pattern) if pattern else random_dfa_pattern(rng, alphabet)[1])
    lengths = np.concatenate([rng.integers(0, 3 * chunk_len, n_short),
                              [0, 1, 15, 16, chunk_len - 1, chunk_len, chunk_len + 1], long_lengths]).astype(np.int64)
    rng.shuffle(lengths)
    offs = np.zeros(len(lengths) + 1, np.int64)
    np.cumsum(lengths, out=offs[1:])
    data = rng.choice(np.frombuffer(alphabet.encode(), np.uint8), int(offs[-1]))
    data[rng.integers(0, len(data), 8)] = 0x7E   # stray bytes (the dead state)
    return tab, data, offs


def _check(tab, data, offs, chunk_len):
    sf, sa, err, pf, pa = _both_forms(tab, data, offs, chunk_len)
    rf, ra = ocpu.dfa_walk(tab.delta, tab.accept, tab.start, data, offs)
    assert err == 0
    assert np.array_equal(sf, pf) and np.array_equal(sa, pa)
    assert np.array_equal(sf, rf) and np.array_equal(sa, ra)


@pytest.mark.gpu
@pytest.mark.parametrize("chunk_len", [256, 4096])
def test_static_equals_plan_two_calls(gpu_lib, chunk_len):
    """Window kernel (32-symbol alphabet): a larger call, then a smaller one."""
    tab, data, offs = _case(3 + chunk_len, 3000, [300_000, 1 << 20, 123_457], chunk_len)
    _check(tab, data, offs, chunk_len)
    _check(*_case(4 + chunk_len, 200, [70_001], chunk_len), chunk_len)
    # a permutation automaton over the same window: every chunk composes and looks back
    _check(*_case(5 + chunk_len, 1000, [400_000, 1 << 19], chunk_len, pattern="mod15"), chunk_len)
    _check(*_case(6 + chunk_len, 100, [90_000], chunk_len, pattern="mod15"), chunk_len)


@pytest.mark.gpu
def test_static_byte_kernel_two_calls(gpu_lib):
    """A pattern whose bytes lie in two 32-byte windows (the byte kernel), and a
    permutation automaton whose chunks keep composing (every chunk looks back)."""
    _check(*_case(11, 500, [250_000, 77_777], 1024, alphabet="aB", pattern="(a|B)*aBB"), 1024)
    _check(*_case(12, 50, [100_000], 1024, alphabet="ab", pattern="(b*ab*ab*)*b*"), 1024)


@pytest.mark.gpu
def test_static_limits_and_bad_ranges(gpu_lib):
    from trex_amd import _lib
    from trex_amd.device import DeviceBuffer
    from trex_amd.engine import DfaEngine
    tab = dfa_byte_table(compile_regex("(a|b)*abb"))
    eng = DfaEngine(tab)
    d = DeviceBuffer(4096)
    d_err = DeviceBuffer.from_array(np.array([7], np.uint32))
    # refused on the host: too many strings, too many chunks, bad chunk length
    assert _static_call(eng, d.ptr, 64, d.ptr, d.ptr, MAX_STRINGS + 1, 64, d.ptr, d.ptr, d_err.ptr) == -4
    assert _static_call(eng, d.ptr, MAX_CHUNKS * 64, d.ptr, d.ptr, 1, 64, d.ptr, d.ptr, d_err.ptr) == -4
    assert _static_call(eng, d.ptr, 64, d.ptr, d.ptr, 1, 100, d.ptr, d.ptr, d_err.ptr) == -3
    # no strings: the error word is cleared, nothing else is touched
    assert _static_call(eng, d.ptr, 64, d.ptr, d.ptr, 0, 64, d.ptr, d.ptr, d_err.ptr) == 0
    _lib.call("trex_device_sync")
    assert int(d_err.download(np.uint32, 1)[0]) == 0
    # a range beyond the buffer: error bit 4, outputs left unwritten
    begin = DeviceBuffer.from_array(np.array([0, 10], np.int64))
    end = DeviceBuffer.from_array(np.array([10, 5000], np.int64))
    d_f = DeviceBuffer.from_array(np.full(2, 0xAB, np.uint8))
    d_a = DeviceBuffer.from_array(np.full(2, 0xAB, np.uint8))
    assert _static_call(eng, d.ptr, 4096, begin.ptr, end.ptr, 2, 64, d_f.ptr, d_a.ptr, d_err.ptr) == 0
    _lib.call("trex_device_sync")
    assert int(d_err.download(np.uint32, 1)[0]) == 4
    assert d_f.download(np.uint8, 2).tolist() == [0xAB, 0xAB]
    # overlapping ranges that need more chunks than the bound: error bit 8
    begin = DeviceBuffer.from_array(np.zeros(4, np.int64))
    end = DeviceBuffer.from_array(np.full(4, 4096, np.int64))
    d_f = DeviceBuffer.from_array(np.full(4, 0xAB, np.uint8))
    d_a = DeviceBuffer.from_array(np.full(4, 0xAB, np.uint8))
    assert _static_call(eng, d.ptr, 4096, begin.ptr, end.ptr, 4, 64, d_f.ptr, d_a.ptr, d_err.ptr) == 0
    _lib.call("trex_device_sync")
    assert int(d_err.download(np.uint32, 1)[0]) == 8
    assert d_f.download(np.uint8, 4).tolist() == [0xAB] * 4
