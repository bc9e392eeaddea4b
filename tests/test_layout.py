"""HBM batch layouts (host side, no GPU): every string's bytes can be read back
from the device image through the ranges the kernels are given."""

from __future__ import annotations

import numpy as np
import pytest

from trex_amd.engine import IL_GROUP, STAGGER_GAP, STAGGER_GROUP, interleaved_image, staggered_ranges


def _unpack_interleaved(image, begin, end):
    """The kernels' reading rule: vector k of string i at begin[i] + 16 * IL_GROUP * k."""
    out = []
    for b, e in zip(begin.tolist(), end.tolist()):
        n = e - b
        rows = [image[b + 16 * IL_GROUP * k:b + 16 * IL_GROUP * k + 16] for k in range(-(-n // 16))]
        out.append(np.concatenate(rows)[:n] if rows else np.zeros(0, np.uint8))
    return out


@pytest.mark.parametrize("n,uniform", [(0, False), (1, False), (63, False), (64, True), (130, False),
                                       (128, True), (192, False)])
def test_interleaved_image_round_trip(n, uniform):
    rng = np.random.default_rng(n)
    lengths = np.full(n, 64) if uniform else rng.integers(0, 100, size=n)
    if n > 8 and not uniform:
        lengths[:6] = [0, 1, 15, 16, 17, 0]
    offsets = np.zeros(n + 1, np.int64)
    np.cumsum(lengths, out=offsets[1:])
    data = rng.integers(0, 256, size=int(offsets[-1])).astype(np.uint8)
    image, begin, end = interleaved_image(data, offsets)
    assert np.all(begin % 16 == 0) and np.array_equal(end - begin, lengths)
    got = _unpack_interleaved(image, begin, end)
    for i in range(n):
        assert np.array_equal(got[i], data[offsets[i]:offsets[i + 1]]), i
    # the lanes of one group read one contiguous row: begin steps by 16 inside a group
    for g in range(0, n, IL_GROUP):
        grp = begin[g:g + IL_GROUP]
        assert np.array_equal(np.diff(grp), np.full(len(grp) - 1, 16))


def test_interleaved_uniform_path_equals_general_path():
    """The one-copy path for equal lengths gives the image of the per-group path."""
    rng = np.random.default_rng(3)
    n, L = 3 * IL_GROUP, 48
    offsets = np.arange(n + 1, dtype=np.int64) * L
    data = rng.integers(0, 256, size=n * L).astype(np.uint8)
    fast, b1, e1 = interleaved_image(data, offsets)
    # same strings with one byte less in the last one: general path
    offsets2 = offsets.copy()
    offsets2[-1] -= 1
    slow, b2, e2 = interleaved_image(data[:-1], offsets2)
    assert np.array_equal(b1, b2)
    last = b1[-1] + 16 * IL_GROUP * (L // 16 - 1) + 15    # the byte dropped
    mask = np.ones(len(fast), bool)
    mask[last] = False
    assert np.array_equal(fast[:len(slow)][mask[:len(slow)]], slow[mask[:len(slow)]])


def test_staggered_ranges_shift_per_group():
    offsets = np.arange(0, 200 * 10 + 1, 10, dtype=np.int64)
    begin, end, size = staggered_ranges(offsets)
    groups = np.arange(200) // STAGGER_GROUP
    assert np.array_equal(begin - offsets[:-1], groups * STAGGER_GAP)
    assert np.array_equal(end - begin, np.diff(offsets))
    assert size == 2000 + (-(-200 // STAGGER_GROUP)) * STAGGER_GAP


@pytest.mark.parametrize("items", [
    ["", "a", "\xff\x00b", "xyz" * 100],
    [b"", b"x", bytearray(b"yz"), b"\x00\xff"],
    ["ab", b"cd", memoryview(b"ef"), bytearray(b"")],
    [],
    ("t", "u", ""),
])
def test_pack_strings_fast_paths_equal_the_per_item_packing(items):
    """Joined-once packing (all str, all bytes) equals encoding item by item."""
    from trex_amd.engine import pack_strings
    data, offsets = pack_strings(items)
    parts = [s.encode("latin-1") if isinstance(s, str) else bytes(s) for s in items]
    assert bytes(data) == b"".join(parts)
    assert offsets.tolist() == np.concatenate([[0], np.cumsum([len(p) for p in parts], dtype=np.int64)]).tolist()
    assert data.dtype == np.uint8 and offsets.dtype == np.int64


def test_pack_strings_rejects_characters_beyond_latin1():
    from trex_amd.engine import pack_strings
    with pytest.raises(UnicodeEncodeError):
        pack_strings(["ok", "Ā"])


def test_plain_dfa_workload_shape():
    """The plain DFA workload is seeded and deterministic: a regex over abcd
    whose table has at most 16 states, 1,024 strings of 1,024 bytes over the
    alphabet."""
    from trex_amd.tables import dfa_byte_table
    from trex_amd.workloads import plain_dfa_workload
    p1, d1, data1, off1 = plain_dfa_workload()
    p2, _, data2, off2 = plain_dfa_workload()
    assert p1 == p2 and np.array_equal(data1, data2) and np.array_equal(off1, off2)
    assert set(p1) <= set("abcd()|*")
    assert dfa_byte_table(d1).n_states <= 16
    assert len(off1) == 1025 and np.all(np.diff(off1) == 1024)
    assert set(np.unique(data1).tolist()) == {97, 98, 99, 100}
