"""GPU parity of the DFA byte-walk kernels against the CPU oracle.

Bit-exact: final state and accept bit of every string, through the C ABI.
"""

import json
from pathlib import Path

import numpy as np
import pytest

from oracle import cpu as ocpu
from oracle.dfa import dfa_simulate, table_walk
from trex_amd.engine import DfaEngine, pack_strings, upload_batch
from trex_amd.regex_to_dfa import compile_regex
from trex_amd.tables import DfaByteTable, dfa_byte_table
from trex_amd.workloads import ALPHABET_32, ALPHABET_ABCD, random_dfa_pattern, ragged_batch, uniform_batch

pytestmark = pytest.mark.gpu
GOLD = json.loads((Path(__file__).parent / "golden" / "reference_vectors.json").read_text())


def _strings(entry):
    if entry.get("strings") is not None:
        return entry["strings"]
    return [ch * n for ch, n in entry["repeat"]]


def _check(eng: DfaEngine, tab: DfaByteTable, data, offsets):
    f, a, kms, tms = eng.match_arrays(data, offsets)
    rf, ra = ocpu.dfa_walk(tab.delta, tab.accept, tab.start, data, offsets)
    bad = np.nonzero((f != rf) | (a != ra))[0]
    assert bad.size == 0, f"{bad.size} mismatches, first at {bad[:5]}: gpu {f[bad[:5]]} cpu {rf[bad[:5]]}"
    return f, a


@pytest.mark.parametrize("entry", GOLD["regex_accept"], ids=lambda e: e["source"])
def test_golden_regex_vectors(gpu_lib, entry):
    dfa = compile_regex(entry["pattern"])
    tab = dfa_byte_table(dfa)
    eng = DfaEngine(tab)
    strings = _strings(entry)
    f, a, _, _ = eng.match_strings(strings)
    assert [bool(x) for x in a] == entry["expected"]
    rf, _ = table_walk(tab.delta, tab.accept, tab.start, *pack_strings(strings))
    assert np.array_equal(f, rf)


def test_reference_compat_prefix_engine_golden(gpu_lib):
    """prefix_engine_* with the reference's (a|b)*a(a|b) tmap (identity for unknown bytes)."""
    import ctypes
    from trex_amd import _lib
    e = GOLD["tmap_tables"][0]
    n = e["N"]
    tmap = np.tile(np.arange(n, dtype=np.uint8), 256)
    for ch, row in e["transitions"].items():
        tmap[ord(ch) * n: ord(ch) * n + n] = row
    acc = np.array(e["accept"], np.uint8)
    _lib.call("prefix_engine_init", tmap.ctypes.data_as(_lib.u8p), acc.ctypes.data_as(_lib.u8p),
              e["start"], n, 32, 1 << 20, 1 << 16)
    data, offs = pack_strings(e["strings"])
    offs32 = offs.astype(np.int32)
    res = np.zeros(len(e["strings"]), np.int32)
    km, tm = ctypes.c_float(), ctypes.c_float()
    _lib.call("prefix_engine_dispatch", data.ctypes.data_as(_lib.u8p), offs32.ctypes.data_as(_lib.intp),
              res.ctypes.data_as(_lib.intp), len(e["strings"]), int(offs[-1]), ctypes.byref(km), ctypes.byref(tm))
    assert res.tolist() == e["expected"]
    # capacity is checked instead of overrun
    rc = _lib.lib().prefix_engine_dispatch(data.ctypes.data_as(_lib.u8p), offs32.ctypes.data_as(_lib.intp),
                                           res.ctypes.data_as(_lib.intp), 1 << 17, int(offs[-1]),
                                           ctypes.byref(km), ctypes.byref(tm))
    assert rc == -4
    _lib.lib().prefix_engine_destroy()


def test_reference_compat_monoid_engine_golden(gpu_lib):
    import ctypes
    from trex_amd import _lib
    e = GOLD["monoid_tables"][0]
    cc = np.array(e["char_compose"], np.uint8)
    cmap = np.full(256, e["unmapped"], np.uint8)
    for ch, i in e["char_map"].items():
        cmap[ord(ch)] = i
    acc = np.array(e["accept"], np.uint8)
    mc = np.array([0, 1, 1, 0], np.uint8)
    _lib.call("monoid_batch_engine_init", e["M"], e["sigma_ext"], e["identity"], cc.ctypes.data_as(_lib.u8p),
              cmap.ctypes.data_as(_lib.u8p), acc.ctypes.data_as(_lib.u8p), mc.ctypes.data_as(_lib.u8p),
              1 << 20, 1 << 16)
    data, offs = pack_strings(e["strings"])
    offs32 = offs.astype(np.int32)
    res = np.zeros(len(e["strings"]), np.int32)
    km, tm = ctypes.c_float(), ctypes.c_float()
    _lib.call("monoid_batch_engine_dispatch", data.ctypes.data_as(_lib.u8p), offs32.ctypes.data_as(_lib.intp),
              res.ctypes.data_as(_lib.intp), len(e["strings"]), int(offs[-1]), ctypes.byref(km), ctypes.byref(tm))
    assert res.tolist() == e["expected"]
    _lib.lib().monoid_batch_engine_destroy()


@pytest.mark.parametrize("seed", range(6))
def test_random_patterns_ragged_all_variants(gpu_lib, seed):
    rng = np.random.default_rng(1000 + seed)
    alpha = ALPHABET_ABCD if seed % 2 == 0 else ALPHABET_32
    _, dfa = random_dfa_pattern(rng, alpha)
    tab = dfa_byte_table(dfa)
    eng = DfaEngine(tab)
    lengths = rng.integers(0, 400, size=3000)
    lengths[:5] = [0, 1, 15, 16, 17]
    lo, hi = ord(min(alpha)), ord(max(alpha)) + 1
    data, offs = ragged_batch(rng, lengths, lo, hi)
    # sprinkle bytes outside the alphabet (and outside the 32-byte window)
    hits = rng.integers(0, len(data), size=50)
    data[hits] = rng.integers(0, 256, size=50, dtype=np.uint8)
    variants = ["k1_generic", "k1_16"]
    if eng.info()["window_base"] >= 0:
        variants += ["k2_window", "k2_window_f", "k2_window_fr"]
    results = []
    for v in variants:
        eng.set_variant(v)
        for pf in (8, 16):
            eng.set_prefetch(pf)
            results.append(_check(eng, tab, data, offs))
    for f, a in results[1:]:
        assert np.array_equal(f, results[0][0]) and np.array_equal(a, results[0][1])


def test_window_variant_out_of_window_bytes(gpu_lib):
    """A byte outside the 32-byte window anywhere in a string forces the dead state."""
    rng = np.random.default_rng(5)
    _, dfa = random_dfa_pattern(rng, ALPHABET_32)
    tab = dfa_byte_table(dfa)
    eng = DfaEngine(tab)
    assert eng.info()["variant"] == "k2_window_f"
    data, offs = uniform_batch(rng, 512, 300, 0x40, 0x60)
    for i in range(512):
        pos = int(offs[i] + (i * 7) % 300)
        data[pos] = [0x00, 0x3F, 0x60, 0x7F, 0xC0, 0xFF, 0x20, 0x41][i % 8]
    for v in ["k2_window", "k2_window_f", "k2_window_fr"]:
        eng.set_variant(v)
        f, a = _check(eng, tab, data, offs)
        assert np.all(f[np.arange(512) % 8 != 7] == tab.dead)


def test_interleaved_uniform_config2_shape(gpu_lib):
    """The interleaved layout on whole groups of equal-length strings (the
    one-copy packing path), k2f and k2fr, against the oracle."""
    from trex_amd.device import DeviceBuffer
    from trex_amd.engine import DeviceBatch
    rng = np.random.default_rng(37)
    _, dfa = random_dfa_pattern(rng, ALPHABET_32)
    tab = dfa_byte_table(dfa)
    eng = DfaEngine(tab)
    data, offs = uniform_batch(rng, 64 * 40, 4096, 0x40, 0x60)
    batch = DeviceBatch(data, offs, layout="interleaved")
    n = len(offs) - 1
    d_f, d_a = DeviceBuffer(n), DeviceBuffer(n)
    rf, ra = ocpu.dfa_walk(tab.delta, tab.accept, tab.start, data, offs)
    for v in ["k2_window_f", "k2_window_fr"]:
        eng.set_variant(v)
        eng.match_batch_device(batch, d_f.ptr, d_a.ptr)
        assert np.array_equal(d_f.download(np.uint8, n), rf) and np.array_equal(d_a.download(np.uint8, n), ra), v


@pytest.mark.parametrize("layout", ["csr", "staggered", "interleaved"])
def test_device_batch_layouts_bit_exact(gpu_lib, layout):
    """Ragged batch (empty strings, lengths off the 16-byte grid, a last group
    of fewer than 64 strings) in every HBM layout, every variant, against the
    oracle."""
    from trex_amd.device import DeviceBuffer
    from trex_amd.engine import DeviceBatch
    rng = np.random.default_rng(31)
    _, dfa = random_dfa_pattern(rng, ALPHABET_32)
    tab = dfa_byte_table(dfa)
    eng = DfaEngine(tab)
    lengths = rng.integers(0, 3000, size=1000)
    lengths[::97] = 4096
    lengths[1:9] = [0, 1, 15, 16, 17, 31, 33, 0]
    data, offs = ragged_batch(rng, lengths, 0x40, 0x60)
    data[rng.integers(0, len(data), 20)] = 0x21
    batch = DeviceBatch(data, offs, layout=layout)
    n = len(offs) - 1
    d_f, d_a = DeviceBuffer(n), DeviceBuffer(n)
    rf, ra = ocpu.dfa_walk(tab.delta, tab.accept, tab.start, data, offs)
    for v in ["k1_generic", "k1_16", "k2_window", "k2_window_f", "k2_window_fr"]:
        eng.set_variant(v)
        for pf in (8, 16):
            eng.set_prefetch(pf)
            eng.match_batch_device(batch, d_f.ptr, d_a.ptr)
            assert np.array_equal(d_f.download(np.uint8, n), rf), (v, pf)
            assert np.array_equal(d_a.download(np.uint8, n), ra), (v, pf)


def test_unaligned_offsets_and_empty(gpu_lib):
    rng = np.random.default_rng(9)
    _, dfa = random_dfa_pattern(rng, ALPHABET_32)
    tab = dfa_byte_table(dfa)
    eng = DfaEngine(tab)
    lengths = np.array([0, 0, 3, 0, 65, 1, 0, 130, 77, 0], np.int64)
    data, offs = ragged_batch(rng, lengths, 0x40, 0x60)
    # offsets that start mid-buffer (first string not at 0)
    data2 = np.concatenate([np.full(5, 0x41, np.uint8), data])
    _check(eng, tab, data2, offs + 5)
    f, a, _, _ = eng.match_arrays(np.zeros(0, np.uint8), np.zeros(4, np.int64))
    assert f.tolist() == [tab.start] * 3 and a.tolist() == [tab.accept[tab.start]] * 3
    f, a, _, _ = eng.match_arrays(np.zeros(0, np.uint8), np.zeros(1, np.int64))
    assert len(f) == 0


def test_wide_dfa_generic_kernel(gpu_lib):
    """More than 16 states: the general byte walk (stride > 16)."""
    dfa = compile_regex("(a|b)*a(a|b)(a|b)(a|b)(a|b)(a|b)")   # 2^6 = 64 states
    tab = dfa_byte_table(dfa)
    assert tab.n_states > 16
    eng = DfaEngine(tab)
    assert eng.info()["variant"] == "k1_generic"
    rng = np.random.default_rng(3)
    data, offs = ragged_batch(rng, rng.integers(0, 200, 2000), ord("a"), ord("c"))
    f, a = _check(eng, tab, data, offs)
    strings = [bytes(data[offs[i]:offs[i + 1]]).decode() for i in range(50)]
    assert [bool(x) for x in a[:50]] == [dfa_simulate(dfa, s) for s in strings]


def test_device_resident_dispatch_matches_host(gpu_lib):
    from trex_amd.device import DeviceBuffer, Stream
    rng = np.random.default_rng(77)
    _, dfa = random_dfa_pattern(rng, ALPHABET_32)
    tab = dfa_byte_table(dfa)
    eng = DfaEngine(tab)
    data, offs = uniform_batch(rng, 4096, 1024, 0x40, 0x60)
    d_bytes, d_offs = upload_batch(data, offs)
    n = len(offs) - 1
    d_f, d_a = DeviceBuffer(n), DeviceBuffer(n)
    s = Stream()
    eng.match_device(d_bytes.ptr, d_offs.ptr, n, d_f.ptr, d_a.ptr, s)
    s.synchronize()
    f = d_f.download(np.uint8, n)
    a = d_a.download(np.uint8, n)
    rf, ra = ocpu.dfa_walk(tab.delta, tab.accept, tab.start, data, offs)
    assert np.array_equal(f, rf) and np.array_equal(a, ra)


def test_config2_shape_slice_bit_exact(gpu_lib):
    """Eight seeded 32-symbol patterns over 16,384 x 4,096 B (1/16 of the 1 GiB batch)."""
    rng = np.random.default_rng(2026)
    data, offs = uniform_batch(rng, 16384, 4096, 0x40, 0x60)
    d_bytes, d_offs = upload_batch(data, offs)
    from trex_amd.device import DeviceBuffer
    n = len(offs) - 1
    d_f, d_a = DeviceBuffer(n), DeviceBuffer(n)
    for p in range(8):
        _, dfa = random_dfa_pattern(rng, ALPHABET_32)
        tab = dfa_byte_table(dfa)
        eng = DfaEngine(tab)
        eng.match_device(d_bytes.ptr, d_offs.ptr, n, d_f.ptr, d_a.ptr)
        f = d_f.download(np.uint8, n)
        a = d_a.download(np.uint8, n)
        rf, ra = ocpu.dfa_walk(tab.delta, tab.accept, tab.start, data, offs)
        assert np.array_equal(f, rf) and np.array_equal(a, ra), f"pattern {p}"


def test_plain_dfa_workload(gpu_lib):
    """The reference CPU path's plain DFA workload (seeded regex over abcd,
    <= 16 states, 1,024 x 1,024 B): final state and accept bit of every string
    equal the C oracle's, the accept bits equal DFA.simulate, and the final
    states equal the dict walk's."""
    from oracle.dfa import dfa_final_state
    from trex_amd.workloads import plain_dfa_workload
    pattern, dfa, data, offsets = plain_dfa_workload()
    tab = dfa_byte_table(dfa)
    assert tab.n_states <= 16 and len(offsets) == 1025 and int(offsets[-1]) == 1 << 20
    eng = DfaEngine(tab)
    f, a = _check(eng, tab, data, offsets)
    strings = [bytes(data[offsets[i]:offsets[i + 1]]).decode("latin-1") for i in range(1024)]
    assert [bool(x) for x in a] == [dfa.simulate(s) for s in strings] == [dfa_simulate(dfa, s) for s in strings]
    assert [int(x) for x in f[:128]] == [dfa_final_state(dfa, s, tab.dead) for s in strings[:128]]
    # uniform text over abcd leaves every string of this workload in the dead
    # state; strings that walk the pattern's live states reach all of them
    from trex_amd.workloads import live_walk_batch
    data2, offs2 = live_walk_batch(tab, np.random.default_rng(5), 1024, 64, "abcd")
    f2, a2 = _check(eng, tab, data2, offs2)
    assert len(np.unique(f2)) == tab.n_states and a2.any()
    strings2 = [bytes(data2[offs2[i]:offs2[i + 1]]).decode("latin-1") for i in range(1024)]
    assert [bool(x) for x in a2] == [dfa_simulate(dfa, s) for s in strings2]


@pytest.mark.gpu
@pytest.mark.parametrize("regex,alpha", [("(a|b)*abb", "ab"), ("(b*ab*ab*)*b*", "ab"), ("(a|b|c|d)*a(b|c)", "abcd"),
                                         ("(a|c|(b|d)(b|d))*", "abcd"), ("(a|b)*a(a|b)(a|b)", "ab")])
def test_reference_compat_monoid_engine_batch_tables(gpu_lib, regex, alpha):
    """precompute_batch_tables' output (the tables the reference's bridge hands
    to monoid_batch_engine_init, src/monoid.py:195-287) through the reference
    entry points, both dispatch forms, against simulate_monoid: short strings,
    strings with characters outside the alphabet (the identity), and strings
    long enough for the prefix dispatch to split."""
    import ctypes
    from oracle.dfa import simulate_monoid
    from trex_amd import _lib
    from trex_amd.monoid import compute_monoid, precompute_batch_tables
    from trex_amd.simulation import DFAMatrices
    dm = DFAMatrices(compile_regex(regex))
    md = compute_monoid(dm)
    t = precompute_batch_tables(md, dm)
    rng = np.random.default_rng(len(regex))
    chars = np.array(list(alpha + "xZ"))
    p = np.r_[np.full(len(alpha), 0.98 / len(alpha)), 0.01, 0.01]
    lengths = np.r_[[0, 1, 2, 3], rng.integers(0, 300, 400), [70_000, 200_001]]
    strings = ["".join(rng.choice(chars, int(n), p=p)) for n in lengths]
    want = [int(simulate_monoid(md, s)) for s in strings]
    assert 0 < sum(want) < len(want)
    _lib.call("monoid_batch_engine_init", t["M"], t["sigma_ext"], t["identity_idx"],
              t["char_compose"].ctypes.data_as(_lib.u8p), t["raw_char_map"].ctypes.data_as(_lib.u8p),
              t["accept"].ctypes.data_as(_lib.u8p), t["monoid_compose"].ctypes.data_as(_lib.u8p), 1 << 22, 1 << 12)
    if t["kgram_compose"] is not None:
        _lib.call("monoid_batch_engine_set_kgram", t["fused_compose"].ctypes.data_as(_lib.u8p),
                  t["kgram_compose"].ctypes.data_as(_lib.u8p), t["kgram_k"], t["sigma_k"], t["sigma"])
    data, offs = pack_strings(strings)
    offs32 = offs.astype(np.int32)
    km, tm = ctypes.c_float(), ctypes.c_float()
    try:
        for entry in ("monoid_batch_engine_dispatch", "monoid_batch_engine_dispatch_prefix"):
            res = np.full(len(strings), -1, np.int32)
            _lib.call(entry, data.ctypes.data_as(_lib.u8p), offs32.ctypes.data_as(_lib.intp),
                      res.ctypes.data_as(_lib.intp), len(strings), int(offs[-1]), ctypes.byref(km), ctypes.byref(tm))
            assert res.tolist() == want, entry
    finally:
        _lib.lib().monoid_batch_engine_destroy()


def _mod15_table():
    """sum of (byte & 15) mod 15 over the window 0x40..0x5f (a permutation on
    every window byte: never synchronises), with a dead state for the rest."""
    delta = np.full((16, 256), 15, np.uint8)
    for b in range(0x40, 0x60):
        delta[:15, b] = (np.arange(15) + (b & 15)) % 15
    accept = np.zeros(16, np.uint8)
    accept[0] = 1
    return DfaByteTable(delta, accept, 0, 15, 16)


def _rotation_tables():
    rng = np.random.default_rng(77)
    tabs = {f"random{i}": dfa_byte_table(random_dfa_pattern(rng, ALPHABET_32)[1]) for i in range(3)}
    tabs["suffix_ABC"] = dfa_byte_table(compile_regex("(A|B|C|D)*ABC"))   # dead state only off the alphabet
    tabs["mod15"] = _mod15_table()
    return tabs


@pytest.mark.parametrize("name", ["random0", "random1", "random2", "suffix_ABC", "mod15"])
def test_rotated_read_order_matches_oracle(gpu_lib, name):
    """The per-wave rotated read order (phase A composed from the rotation
    point, phase B walked, joined) gives the oracle's results with rotation
    off, on and automatic: 2,048 packed 4 KiB strings (every rotation class),
    then ragged strings at unaligned offsets with stray bytes in heads,
    bodies and tails, over synchronising automata, one whose dead state no
    window byte reaches, and a permutation that never synchronises."""
    tab = _rotation_tables()[name]
    eng = DfaEngine(tab)
    assert eng.info()["variant"] == "k2_window_f"
    rng = np.random.default_rng(len(name))
    alpha = np.frombuffer(b"ABCD", np.uint8) if name == "suffix_ABC" else np.arange(0x40, 0x60, dtype=np.uint8)
    packed = alpha[rng.integers(0, len(alpha), 2048 * 4096)]
    packed_offs = np.arange(2049, dtype=np.int64) * 4096
    lengths = rng.integers(0, 5000, 1500)
    lengths[:12] = [0, 1, 15, 16, 17, 255, 256, 257, 4095, 4096, 4097, 40000]
    offs = np.zeros(len(lengths) + 1, np.int64)
    np.cumsum(lengths, out=offs[1:])
    ragged = np.concatenate([alpha[rng.integers(0, len(alpha), 7)], alpha[rng.integers(0, len(alpha), int(offs[-1]))]])
    offs += 7
    hits = rng.integers(0, len(ragged), 400)
    ragged[hits] = rng.integers(0, 256, 400, dtype=np.uint8)
    packed_stray = packed.copy()
    hits = rng.integers(0, len(packed_stray), 300)
    packed_stray[hits] = np.where(rng.random(300) < 0.5, 0x3f, 0x60).astype(np.uint8)
    for data, o in ((packed, packed_offs), (packed_stray, packed_offs), (ragged, offs)):
        results = []
        for mode in (0, 1, -1):
            eng.set_rotation(mode)
            results.append(_check(eng, tab, data, o))
        for f, a in results[1:]:
            assert np.array_equal(f, results[0][0]) and np.array_equal(a, results[0][1])
    eng.set_rotation(1)
    assert eng.rotation()[0]
    eng.set_rotation(0)
    assert not eng.rotation()[0]
    eng.set_variant("k2_window")   # the rotated order is a k2f form only
    eng.set_rotation(1)
    _check(eng, tab, ragged, offs)


def test_rotation_auto_follows_synchronisation(gpu_lib):
    """Automatic rotation is on exactly for tables that synchronise, and the
    library's synchronisation test agrees with the host's."""
    from trex_amd import _lib
    from trex_amd.engine import synchronizes
    tabs = _rotation_tables()
    rng = np.random.default_rng(5)
    for i in range(12):
        tabs[f"extra{i}"] = dfa_byte_table(random_dfa_pattern(rng, ALPHABET_32)[1])
    seen = set()
    for name, tab in tabs.items():
        eng = DfaEngine(tab)
        if eng.info()["window_base"] < 0:
            continue
        active, sync = eng.rotation()
        want = synchronizes(tab.delta, tab.n_states, tab.dead)
        assert sync == want, name
        assert active == sync, name
        seen.add(sync)
        eng.set_rotation(0)
        assert eng.rotation() == (False, sync)
        with pytest.raises(_lib.TrexError):
            eng.set_rotation(2)
    assert seen == {True, False}
