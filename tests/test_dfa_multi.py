"""Multi-pattern DFA-16 pass (trex_dfa_multi_*): every (pattern, string)
result equals the C oracle's table walk and the pattern's own single-pattern
kernel, bit for bit, over ragged and empty strings, bytes outside the
patterns' windows, patterns with different windows, every layout, all four table
forms and every group size."""

import numpy as np
import pytest

ALPHABETS = {
    0x40: "".join(chr(c) for c in range(0x40, 0x60)),   # the 32-symbol benchmark alphabet
    0x60: "abcdefghijklmnopqrstuvwxyz{|}~",                # a second window
    0x20: "0123456789:;<=>?",                               # a third (digits and punctuation)
}


def _patterns(rng, count):
    from trex_amd.tables import dfa_byte_table
    from trex_amd.workloads import random_dfa_pattern
    bases = list(ALPHABETS)
    out = []
    for i in range(count):
        alphabet = ALPHABETS[bases[i % len(bases)]]
        # alphabets holding regex metacharacters are drawn as plain letters
        safe = "".join(c for c in alphabet if c not in "[]\\()|*+?.{}^$-")
        _, dfa = random_dfa_pattern(rng, safe)
        out.append(dfa_byte_table(dfa))
    return out


def _batch(rng, n, max_len, lo=0x20, hi=0x80, outside=0.002):
    lengths = rng.integers(0, max_len + 1, size=n)
    lengths[rng.random(n) < 0.05] = 0                  # empty strings
    offsets = np.zeros(n + 1, np.int64)
    np.cumsum(lengths, out=offsets[1:])
    data = rng.integers(lo, hi, size=int(offsets[-1]), dtype=np.uint8)
    stray = rng.random(len(data)) < outside            # bytes no pattern's window holds
    data[stray] = rng.integers(0x80, 0x100, size=int(stray.sum()), dtype=np.uint8)
    return data, offsets


def _oracle(tables, data, offsets):
    from oracle import cpu as ocpu
    fs, acs = [], []
    for t in tables:
        f, a = ocpu.dfa_walk(t.delta, t.accept, t.start, data, offsets)
        fs.append(f)
        acs.append(a)
    return np.stack(fs), np.stack(acs)


def test_multi_binding_declared():
    from trex_amd import _lib
    for name in ("trex_dfa_multi_create", "trex_dfa_multi_match_ranges", "trex_dfa_multi_match_interleaved",
                 "trex_dfa_multi_set_form", "trex_dfa_multi_set_group", "trex_dfa_multi_info",
                 "trex_dfa_multi_destroy"):
        assert name in _lib.exported_symbols()


@pytest.mark.gpu
@pytest.mark.parametrize("layout", ["csr", "staggered", "interleaved"])
@pytest.mark.parametrize("form", ["k2f", "k2w", "k2fp", "k2fc"])
def test_gpu_multi_equals_oracle(gpu_lib, layout, form):
    from trex_amd.engine import DfaEngine, DfaMultiEngine
    rng = np.random.default_rng(101 + len(layout) + len(form))
    tables = _patterns(rng, 11)
    engines = [DfaEngine(t) for t in tables]
    assert all(e.info()["window_base"] >= 0 for e in engines)
    assert len({e.info()["window_base"] for e in engines}) == 3
    data, offsets = _batch(rng, 3000, 300)
    want_f, want_a = _oracle(tables, data, offsets)
    multi = DfaMultiEngine(engines)
    multi.set_form(form)
    for group in (8, 4, 3, 1):            # 6+5, 4+4+3, 3+3+3+2 and 11 launches
        multi.set_group(group)
        f, a = multi.match_arrays(data, offsets, layout=layout)
        assert np.array_equal(f, want_f), (group, np.argwhere(f != want_f)[:5])
        assert np.array_equal(a, want_a), group
    # and equal to each pattern's own kernel
    for i, e in enumerate(engines):
        f1, a1, _, _ = e.match_arrays(data, offsets)
        assert np.array_equal(f1, want_f[i]) and np.array_equal(a1, want_a[i])
    multi.close()
    for e in engines:
        e.close()


@pytest.mark.gpu
@pytest.mark.parametrize("n_patterns", [1, 2, 5, 7, 8])
def test_gpu_multi_every_launch_width(gpu_lib, n_patterns):
    """P patterns in one launch for every P the kernels hold, unaligned CSR
    offsets (head bytes) and long strings."""
    from trex_amd.engine import DfaEngine, DfaMultiEngine
    rng = np.random.default_rng(200 + n_patterns)
    tables = _patterns(rng, n_patterns)
    data, offsets = _batch(rng, 700, 5000)
    offsets = offsets + 3                                   # no string starts 16-byte aligned
    data = np.concatenate([np.zeros(3, np.uint8), data])
    want_f, want_a = _oracle(tables, data, offsets)
    small = sum(t.n_states <= 4 for t in tables)
    for form in ("k2f", "k2w", "k2fp", "k2fc"):
        multi = DfaMultiEngine([DfaEngine(t) for t in tables])
        # a new handle: compact tables, all patterns in one launch, when two or
        # more patterns have at most 4 states; else paired in launches of four
        start = multi.info()
        assert start["form"] == ("k2fc" if small >= 2 else "k2fp")
        assert start["group"] == (8 if small >= 2 else 4)
        multi.set_form(form)
        assert multi.info() == {"n_patterns": n_patterns, "form": form, "group": start["group"]}
        multi.set_group(8)                                  # all in one launch
        f, a = multi.match_arrays(data, offsets, layout="csr")
        assert np.array_equal(f, want_f) and np.array_equal(a, want_a), form
        multi.close()


def _sized_patterns(rng, small, large):
    """`small` tables of at most 4 states and `large` of more, over the three
    windows (random regexes redrawn until they fall in the wanted class)."""
    from trex_amd.tables import dfa_byte_table
    from trex_amd.workloads import random_dfa_pattern
    bases = list(ALPHABETS)
    out, want = [], [True] * small + [False] * large
    i = 0
    while want:
        alphabet = ALPHABETS[bases[i % len(bases)]]
        safe = "".join(c for c in alphabet if c not in "[]\\()|*+?.{}^$-")
        t = dfa_byte_table(random_dfa_pattern(rng, safe)[1])
        i += 1
        if (t.n_states <= 4) in want:
            want.remove(t.n_states <= 4)
            out.append(t)
    order = rng.permutation(len(out))            # classes interleaved in pattern order
    return [out[k] for k in order]


@pytest.mark.gpu
@pytest.mark.parametrize("small,large", [(8, 0), (0, 8), (5, 3), (3, 9), (11, 2)])
def test_gpu_multi_compact_mixed_widths(gpu_lib, small, large):
    """The compact form with every mix of one-byte (at most 4 states) and
    full patterns per launch: launch order puts full patterns first and the
    outputs go back to the patterns' own rows; every layout and group size,
    head bytes, long strings and bytes outside the windows."""
    from trex_amd.engine import DfaEngine, DfaMultiEngine
    rng = np.random.default_rng(300 + 16 * small + large)
    tables = _sized_patterns(rng, small, large)
    assert sum(t.n_states <= 4 for t in tables) == small
    data, offsets = _batch(rng, 900, 3000)
    offsets = offsets + 5
    data = np.concatenate([np.zeros(5, np.uint8), data])
    want_f, want_a = _oracle(tables, data, offsets)
    multi = DfaMultiEngine([DfaEngine(t) for t in tables])
    multi.set_form("k2fc")
    for layout in ("csr", "interleaved"):
        for group in (8, 5, 3, 1):
            multi.set_group(group)
            f, a = multi.match_arrays(data, offsets, layout=layout)
            assert np.array_equal(f, want_f), (layout, group, np.argwhere(f != want_f)[:5])
            assert np.array_equal(a, want_a), (layout, group)
    assert multi.info()["form"] == "k2fc"
    multi.close()


@pytest.mark.gpu
def test_gpu_multi_uniform_interleaved_batch(gpu_lib):
    """The benchmark's shape at reduced size: 8 patterns over the 32-symbol
    window, 8,192 strings x 4,096 bytes, interleaved layout."""
    from trex_amd.device import DeviceBuffer
    from trex_amd.engine import DeviceBatch, DfaEngine, DfaMultiEngine
    from trex_amd.tables import dfa_byte_table
    from trex_amd.workloads import ALPHABET_32, random_dfa_pattern
    rng = np.random.default_rng(7)
    tables = [dfa_byte_table(random_dfa_pattern(rng, ALPHABET_32)[1]) for _ in range(8)]
    n, L = 8192, 4096
    data = rng.integers(0x40, 0x60, size=n * L, dtype=np.uint8)
    offsets = np.arange(n + 1, dtype=np.int64) * L
    engines = [DfaEngine(t) for t in tables]
    batch = DeviceBatch(data, offsets, layout="interleaved")
    d_f, d_a = DeviceBuffer(8 * n), DeviceBuffer(8 * n)
    multi = DfaMultiEngine(engines)
    multi.match_batch_device(batch, d_f.ptr, d_a.ptr)
    f = d_f.download(np.uint8).reshape(8, n)
    a = d_a.download(np.uint8).reshape(8, n)
    s_f, s_a = DeviceBuffer(n), DeviceBuffer(n)
    for i, e in enumerate(engines):
        e.match_batch_device(batch, s_f.ptr, s_a.ptr)
        assert np.array_equal(f[i], s_f.download(np.uint8)) and np.array_equal(a[i], s_a.download(np.uint8))
    want_f, want_a = _oracle(tables, data[:512 * L], offsets[:513])
    assert np.array_equal(f[:, :512], want_f) and np.array_equal(a[:, :512], want_a)
    batch.free()


@pytest.mark.gpu
def test_gpu_multi_refuses_tables_without_a_window(gpu_lib):
    from trex_amd import _lib
    from trex_amd.engine import DfaEngine, DfaMultiEngine
    from trex_amd.regex_to_dfa import compile_regex
    from trex_amd.tables import dfa_byte_table
    ok = DfaEngine(dfa_byte_table(compile_regex("(a|b)*abb")))
    two_windows = DfaEngine(dfa_byte_table(compile_regex("(a|B)*aBB")))      # 'B' and 'a' in different windows
    assert two_windows.info()["window_base"] == -1
    wide = DfaEngine(dfa_byte_table(compile_regex("(a|b)*a(a|b)(a|b)(a|b)(a|b)(a|b)")))   # > 16 states
    for bad in (two_windows, wide):
        with pytest.raises(_lib.TrexError, match="pattern 1"):
            DfaMultiEngine([ok, bad])
    with pytest.raises(ValueError):
        DfaMultiEngine([])
    multi = DfaMultiEngine([ok])
    with pytest.raises(_lib.TrexError):
        multi.set_group(9)
    with pytest.raises(KeyError):
        multi.set_form("k1")
    # an empty batch launches nothing and returns empty results
    f, a = multi.match_arrays(np.zeros(0, np.uint8), np.zeros(1, np.int64))
    assert f.shape == (1, 0) and a.shape == (1, 0)
