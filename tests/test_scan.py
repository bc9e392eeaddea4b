"""Look-back scan: the oracle pinned on the CPU, then the chunked GPU kernel
against the oracle (bit-exact final state and accept bit)."""

import numpy as np
import pytest

from oracle import cpu as ocpu
from oracle.dfa import table_walk_one
from oracle.scan import lookback_final_state, plan_chunks, wave_protocol
from trex_amd.regex_to_dfa import compile_regex
from trex_amd.tables import dfa_byte_table
from trex_amd.workloads import ALPHABET_32, random_dfa_pattern, ragged_batch

EVEN_A = "(b*ab*ab*)*b*"


# ---------------------------------------------------------------------------
# oracle pins (CPU)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("tile", [1, 3, 16, 256])
def test_lookback_oracle_equals_sequential_walk(tile):
    """Associativity: any tiling and any publication order gives the walk's state."""
    rng = np.random.default_rng(tile)
    _, dfa = random_dfa_pattern(rng, "abcd")
    tab = dfa_byte_table(dfa)
    for _ in range(20):
        n = int(rng.integers(1, 700))
        data = rng.integers(ord("a"), ord("e"), n, dtype=np.uint8)
        want = table_walk_one(tab.delta, tab.accept, tab.start, bytes(data))[0]
        n_tiles = -(-n // tile)
        for order in (None, list(range(n_tiles)), list(rng.permutation(n_tiles))):
            assert lookback_final_state(tab.delta, tab.start, data, 0, n, tile, order) == want


def test_lookback_oracle_even_a_closed_form():
    """Even number of a's (cuda/monoid_scan.cu:544-578): parity of the count."""
    tab = dfa_byte_table(compile_regex(EVEN_A))
    rng = np.random.default_rng(4)
    for n in (100, 255, 256, 257, 1000, 5000):
        data = rng.choice(np.frombuffer(b"ab", np.uint8), n)
        s = lookback_final_state(tab.delta, tab.start, data, 0, n, 256)
        assert bool(tab.accept[s]) == (int(np.count_nonzero(data == ord("a"))) % 2 == 0)


def test_lookback_oracle_mutation_caught():
    """Composing aggregates in the wrong order (a planted mistake) is caught."""
    tab = dfa_byte_table(compile_regex("(a|b)*abb"))
    rng = np.random.default_rng(8)

    def wrong(delta, start, data, tile):
        n = len(data)
        tiles = [(k, min(n, k + tile)) for k in range(0, n, tile)]
        f = np.arange(delta.shape[0])
        for a, b in tiles:
            g = np.arange(delta.shape[0])
            for x in data[a:b]:
                g = delta[g, int(x)]
            f = f[g]                     # applies the new tile first: wrong
        return int(f[start])

    bad = 0
    for _ in range(50):
        data = rng.choice(np.frombuffer(b"ab", np.uint8), int(rng.integers(10, 60)))
        want = table_walk_one(tab.delta, tab.accept, tab.start, bytes(data))[0]
        assert lookback_final_state(tab.delta, tab.start, data, 0, len(data), 4) == want
        bad += wrong(tab.delta, tab.start, data, 4) != want
    assert bad > 0


# ---------------------------------------------------------------------------
# the kernel's wave-level protocol, modelled on the CPU
# ---------------------------------------------------------------------------
def _model_case(seed, alpha=b"ab", pattern=EVEN_A):
    rng = np.random.default_rng(seed)
    tab = dfa_byte_table(compile_regex(pattern))
    lengths = rng.integers(0, 700, int(rng.integers(3, 12)))
    strings = [bytes(rng.choice(np.frombuffer(alpha, np.uint8), int(n))) for n in lengths]
    want = [table_walk_one(tab.delta, tab.accept, tab.start, s)[0] for s in strings]
    return rng, tab, strings, want


@pytest.mark.parametrize("seed", range(12))
@pytest.mark.parametrize("late", [False, True], ids=["publish-first", "publish-last"])
def test_wave_protocol_reads_only_earlier_waves(seed, late):
    """With in-wave chains resolved in registers, every look-back read is of
    a chunk in an earlier wave, every wave finishes whatever the interleaving
    and whichever order a wave gives its stores and its look-back, and every
    string's final state equals the sequential walk."""
    rng, tab, strings, want = _model_case(seed)
    for wave, chunk in ((4, 16), (8, 32), (64, 16)):
        out, reads, stalled = wave_protocol(tab.delta, tab.start, strings, chunk, rng, True, late, wave)
        assert not stalled
        assert out == want
        assert all(j // wave < c // wave for c, j, _ in reads)


def test_previous_protocol_can_wait_on_its_own_wave():
    """The protocol before in-wave chains linked a lane only to the lane just
    before it, so a segment head could look back to a chunk of its own wave
    (a string's last, shorter chunk sorted next to its predecessor). That read
    depends on the wave's own store coming first: with the store placed after
    the look-back loop, the wave never finishes. This is the failure class of
    a look-back that did not end; the kernel no longer has such reads."""
    own_reads = stalls = 0
    for seed in range(12):
        rng, tab, strings, want = _model_case(seed)
        out, reads, stalled = wave_protocol(tab.delta, tab.start, strings, 16, rng, False, False, 8)
        assert not stalled and out == want            # publishing first, it completes
        own_reads += sum(j // 8 == c // 8 for c, j, _ in reads)
        _, _, stalled = wave_protocol(tab.delta, tab.start, strings, 16, rng, False, True, 8)
        stalls += stalled
    assert own_reads > 0 and stalls > 0


def test_plan_chunks_order():
    chunks = plan_chunks([40, 0, 33, 16], 16)
    # full chunks first (first chunks ahead), then the shorter tails
    assert [(s, lo, hi) for s, lo, hi, _ in chunks[:5]] == [(0, 0, 16), (2, 0, 16), (3, 0, 16), (0, 16, 32),
                                                             (2, 16, 32)]
    for k, (sid, lo, hi, pred) in enumerate(chunks):
        assert pred < k
        if lo > 0:
            assert chunks[pred][0] == sid and chunks[pred][2] == lo


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------
def _run_plan(tab, data, offsets, chunk_len, layout="csr"):
    from trex_amd.device import DeviceBuffer
    from trex_amd.engine import ChunkPlan, DeviceBatch, DfaEngine
    eng = DfaEngine(tab)
    batch = DeviceBatch(data, offsets, layout=layout)
    plan = ChunkPlan.for_batch(batch, chunk_len)
    n = len(offsets) - 1
    d_f, d_a = DeviceBuffer(max(n, 1)), DeviceBuffer(max(n, 1))
    eng.match_plan(plan, batch.d_bytes.ptr, d_f.ptr, d_a.ptr)
    assert plan.error() == 0
    return d_f.download(np.uint8, n), d_a.download(np.uint8, n), plan.n_chunks


@pytest.mark.gpu
@pytest.mark.parametrize("pattern,chunk_len", [(EVEN_A, 64), (EVEN_A, 1024), ("(a|b)*abb", 256),
                                               ("random32", 128), ("random32", 4096)])
def test_plan_matches_oracle(gpu_lib, pattern, chunk_len):
    rng = np.random.default_rng(chunk_len)
    if pattern == "random32":
        _, dfa = random_dfa_pattern(rng, ALPHABET_32)
        lo, hi = 0x40, 0x60
    else:
        dfa = compile_regex(pattern)
        lo, hi = ord("a"), ord("c")
    tab = dfa_byte_table(dfa)
    lengths = np.concatenate([rng.integers(0, 300, 400), rng.integers(1000, 20000, 60),
                              [0, 1, 15, 16, 17, chunk_len - 1, chunk_len, chunk_len + 1, 200_000]])
    rng.shuffle(lengths)
    data, offs = ragged_batch(rng, lengths, lo, hi)
    data[rng.integers(0, len(data), 10)] = 0x7e
    f, a, n_chunks = _run_plan(tab, data, offs, chunk_len, layout="staggered" if chunk_len > 1000 else "csr")
    assert n_chunks > len(lengths)
    rf, ra = ocpu.dfa_walk(tab.delta, tab.accept, tab.start, data, offs)
    bad = np.nonzero((f != rf) | (a != ra))[0]
    assert bad.size == 0, f"{bad.size} mismatches at {bad[:5]}, lengths {lengths[bad[:5]]}"


@pytest.mark.gpu
def test_plan_even_a_long_strings_parity(gpu_lib):
    """Long strings (1 MiB) across many chunks: accept = even number of a's."""
    tab = dfa_byte_table(compile_regex(EVEN_A))
    rng = np.random.default_rng(99)
    lengths = np.array([1 << 20, (1 << 20) + 3, 777_777, 5], np.int64)
    data, offs = ragged_batch(rng, lengths, ord("a"), ord("c"))
    f, a, n_chunks = _run_plan(tab, data, offs, 4096)
    for i in range(len(lengths)):
        seg = data[offs[i]:offs[i + 1]]
        assert bool(a[i]) == (int(np.count_nonzero(seg == ord("a"))) % 2 == 0)


def test_lookback_chunk_len():
    from trex_amd.engine import LOOKBACK_TARGET_CHUNKS, lookback_chunk_len
    assert lookback_chunk_len(1 << 30) == 16384
    assert lookback_chunk_len(0) == 1024 and lookback_chunk_len(1 << 20) == 1024
    assert lookback_chunk_len(1 << 40) == 1 << 16
    for total in (3 << 27, 5 << 28, 1 << 31):
        c = lookback_chunk_len(total)
        assert c & (c - 1) == 0 and 1024 <= c <= 1 << 16
        assert total / c / LOOKBACK_TARGET_CHUNKS < 2 ** 0.5 * 1.0001 or c == 1 << 16


@pytest.mark.gpu
def test_plan_constant_and_composed_chunks_side_by_side(gpu_lib):
    """Chunks whose function is constant (a byte sending every state to the
    dead state) publish their final state at once, next to chunks of the same
    string that still look back (a permutation automaton): both kinds share
    waves, and every result equals the sequential walk."""
    tab = dfa_byte_table(compile_regex(EVEN_A))
    rng = np.random.default_rng(5)
    lengths = np.array([300_000, 123_457, 64, 0, 90_001], np.int64)
    data, offs = ragged_batch(rng, lengths, ord("a"), ord("c"))
    # a '~' (dead for this automaton) in about a third of the 256-byte chunks
    starts = np.arange(0, len(data), 256)
    starts = starts[rng.random(len(starts)) < 0.3]
    data[np.minimum(starts + rng.integers(0, 256, size=len(starts)), len(data) - 1)] = 0x7e
    f, a, n_chunks = _run_plan(tab, data, offs, 256)
    rf, ra = ocpu.dfa_walk(tab.delta, tab.accept, tab.start, data, offs)
    assert np.array_equal(f, rf) and np.array_equal(a, ra)
    assert n_chunks > 2000


@pytest.mark.gpu
def test_plan_adaptive_chunks_small_table(gpu_lib):
    """A table of fewer than 16 states (the padding slots start at state 0 so
    that a synchronised function is seen constant) over long strings, with the
    chunk length picked from the batch size."""
    from trex_amd.device import DeviceBuffer
    from trex_amd.engine import ChunkPlan, DeviceBatch, DfaEngine, lookback_chunk_len
    rng = np.random.default_rng(17)
    _, dfa = random_dfa_pattern(rng, ALPHABET_32)
    tab = dfa_byte_table(dfa)
    assert tab.n_states < 16
    lengths = np.array([40 << 20, (24 << 20) + 5, 3, 1 << 20], np.int64)
    data, offs = ragged_batch(rng, lengths, 0x40, 0x60)
    eng = DfaEngine(tab)
    batch = DeviceBatch(data, offs, layout="csr")
    plan = ChunkPlan.for_batch(batch)
    assert plan.chunk_len == lookback_chunk_len(int(lengths.sum()))
    d_f, d_a = DeviceBuffer(len(lengths)), DeviceBuffer(len(lengths))
    eng.match_plan(plan, batch.d_bytes.ptr, d_f.ptr, d_a.ptr)
    assert plan.error() == 0
    rf, ra = ocpu.dfa_walk(tab.delta, tab.accept, tab.start, data, offs)
    assert np.array_equal(d_f.download(np.uint8, len(lengths)), rf)
    assert np.array_equal(d_a.download(np.uint8, len(lengths)), ra)


@pytest.mark.gpu
@pytest.mark.parametrize("chunk_len", [64, 1024, 8192])
def test_plan_byte_kernel_two_windows(gpu_lib, chunk_len):
    """A pattern whose bytes lie in two 32-byte windows runs the byte kernel
    (dfa16_chunks_bytes): same results as the sequential walk."""
    from trex_amd.engine import DfaEngine
    tab = dfa_byte_table(compile_regex("(a|B)*aBB"))
    assert DfaEngine(tab).info()["window_base"] == -1
    rng = np.random.default_rng(chunk_len + 1)
    lengths = np.concatenate([rng.integers(0, 200, 300), [0, 1, 16, chunk_len, chunk_len + 1, 250_000, 77_777]])
    offs = np.zeros(len(lengths) + 1, np.int64)
    np.cumsum(lengths, out=offs[1:])
    data = np.frombuffer(b"aBB~", np.uint8)[rng.choice(4, size=int(offs[-1]), p=[0.45, 0.45, 0.099, 0.001])]
    f, a, n_chunks = _run_plan(tab, np.ascontiguousarray(data), offs, chunk_len)
    rf, ra = ocpu.dfa_walk(tab.delta, tab.accept, tab.start, data, offs)
    assert np.array_equal(f, rf) and np.array_equal(a, ra)


def _run_plan_image(tab, data, offsets, chunk_len, chains=None, gap=None):
    """Match through the plan's chunk-interleaved image; also returns the
    plain (CSR) results of the same plan."""
    from trex_amd.device import DeviceBuffer
    from trex_amd.engine import ChunkPlan, DeviceBatch, DfaEngine
    eng = DfaEngine(tab)
    batch = DeviceBatch(data, offsets, layout="csr")
    plan = ChunkPlan.for_batch(batch, chunk_len)
    if chains is not None:
        plan.set_chains(chains)
    if gap is not None:
        plan.set_image_gap(gap)
    image = plan.build_image(data)
    n_tiles = -(-plan.n_chunks // 64)
    g = 256 if gap is None else gap
    assert plan.image_size() >= len(data) + n_tiles * g and (plan.image_size() - n_tiles * g) % 1024 == 0
    d_img = DeviceBuffer(image.nbytes)
    d_img.upload(image)
    n = len(offsets) - 1
    d_f, d_a = DeviceBuffer(max(n, 1)), DeviceBuffer(max(n, 1))
    from trex_amd import _lib
    with pytest.raises(_lib.TrexError, match="image buffer"):   # refused on the host, nothing launched
        eng.match_plan_image(plan, d_img.ptr, plan.image_size() - 16, d_f.ptr, d_a.ptr)
    eng.match_plan_image(plan, d_img.ptr, d_img.nbytes, d_f.ptr, d_a.ptr)
    assert plan.error() == 0
    fi, ai = d_f.download(np.uint8, n), d_a.download(np.uint8, n)
    eng.match_plan(plan, batch.d_bytes.ptr, d_f.ptr, d_a.ptr)
    assert plan.error() == 0
    return fi, ai, d_f.download(np.uint8, n), d_a.download(np.uint8, n)


@pytest.mark.gpu
@pytest.mark.parametrize("chains,gap", [(1, None), (2, None), (1, 0), (1, 4112)])
@pytest.mark.parametrize("pattern,chunk_len", [("random32", 1024), ("random32", 8192), (EVEN_A, 64),
                                               (EVEN_A, 4096), ("(a|B)*aBB", 512)])
def test_plan_image_matches_oracle(gpu_lib, pattern, chunk_len, chains, gap):
    """The chunk-interleaved image gives the sequential walk's results (and
    the CSR launch's) over ragged lengths, partial vectors, empty strings,
    synchronising and permutation automata and the byte kernel, with one or
    two chains per lane and with the default, no and an odd tile gap."""
    rng = np.random.default_rng(chunk_len + len(pattern))
    if pattern == "random32":
        _, dfa = random_dfa_pattern(rng, ALPHABET_32)
        alphabet = bytes(range(0x40, 0x60))
    else:
        dfa = compile_regex(pattern)
        alphabet = b"ab" if pattern == EVEN_A else b"aBB"
    tab = dfa_byte_table(dfa)
    lengths = np.concatenate([rng.integers(0, 300, 300), rng.integers(1000, 40000, 40),
                              [0, 1, 15, 16, 17, chunk_len - 1, chunk_len, chunk_len + 1, 3 * chunk_len + 5,
                               300_001]])
    rng.shuffle(lengths)
    offs = np.zeros(len(lengths) + 1, np.int64)
    np.cumsum(lengths, out=offs[1:])
    data = np.frombuffer(alphabet, np.uint8)[rng.integers(0, len(alphabet), int(offs[-1]))].copy()
    data[rng.integers(0, len(data), 8)] = 0x7e
    fi, ai, fc, ac = _run_plan_image(tab, data, offs, chunk_len, chains, gap)
    rf, ra = ocpu.dfa_walk(tab.delta, tab.accept, tab.start, data, offs)
    assert np.array_equal(fi, rf) and np.array_equal(ai, ra)
    assert np.array_equal(fc, rf) and np.array_equal(ac, ra)


@pytest.mark.gpu
@pytest.mark.parametrize("gap", [0, 256, 4112])
def test_plan_image_size_equal_lengths(gpu_lib, gap):
    """Strings of one length that is a multiple of the chunk length: the image
    is exactly the bytes plus one gap per tile (no padding)."""
    from trex_amd import _lib
    from trex_amd.engine import ChunkPlan
    L, C = 1 << 20, 8192
    begin = np.arange(4, dtype=np.int64) * L
    plan = ChunkPlan(begin, begin + L, C)
    if gap != 256:
        assert plan.image_size() == (4 << 20) + 8 * 256   # the default gap
        plan.set_image_gap(gap)
    assert plan.n_chunks == 512 and plan.image_size() == (4 << 20) + 8 * gap
    # the layout itself: chunks of equal length run first chunks first, then
    # the others in string order; vector i of tile t's k-th chunk sits at
    # t * (C * 64 + gap) + (64 i + k) * 16, and the gaps stay zero
    data = np.random.default_rng(0).integers(0, 256, 4 * L, dtype=np.uint8)
    image = plan.build_image(data)
    tiles = image[:plan.image_size()].reshape(8, C * 64 + gap)
    assert not tiles[:, C * 64:].any()
    starts = [s * L for s in range(4)] + [s * L + j * C for s in range(4) for j in range(1, L // C)]
    rows = tiles[:, :C * 64].reshape(8, C // 16, 64, 16).transpose(0, 2, 1, 3).reshape(512, C)
    want = np.stack([data[a:a + C] for a in starts])
    assert np.array_equal(rows, want)
    with pytest.raises(_lib.TrexError):
        plan.build_image(data[:-1])          # a chunk would lie outside the source
    for bad in (-16, 8, 65552):
        with pytest.raises(_lib.TrexError, match="image gap"):
            plan.set_image_gap(bad)


@pytest.mark.gpu
@pytest.mark.parametrize("pattern", [EVEN_A, "random32"])
def test_plan_delayed_publication(gpu_lib, pattern):
    """Tiles publish in reverse order within groups of 8 (the plan's hold
    hook), so successors wait for and walk back over unpublished
    predecessors: results still equal the sequential walk, both layouts."""
    from trex_amd.device import DeviceBuffer
    from trex_amd.engine import ChunkPlan, DeviceBatch, DfaEngine
    rng = np.random.default_rng(77)
    if pattern == "random32":
        _, dfa = random_dfa_pattern(rng, ALPHABET_32)
        lo, hi = 0x40, 0x60
    else:
        dfa = compile_regex(pattern)
        lo, hi = ord("a"), ord("c")
    tab = dfa_byte_table(dfa)
    lengths = np.array([400_000, 250_003, 64, 0, 99_999, 1, 150_000], np.int64)
    data, offs = ragged_batch(rng, lengths, lo, hi)
    eng = DfaEngine(tab)
    batch = DeviceBatch(data, offs, layout="csr")
    plan = ChunkPlan.for_batch(batch, 256)
    assert plan.n_chunks > 64 * 16
    plan.set_hold(40)
    n = len(lengths)
    d_f, d_a = DeviceBuffer(n), DeviceBuffer(n)
    rf, ra = ocpu.dfa_walk(tab.delta, tab.accept, tab.start, data, offs)
    eng.match_plan(plan, batch.d_bytes.ptr, d_f.ptr, d_a.ptr)
    assert plan.error() == 0
    assert np.array_equal(d_f.download(np.uint8, n), rf) and np.array_equal(d_a.download(np.uint8, n), ra)
    image = plan.build_image(data)
    d_img = DeviceBuffer(image.nbytes)
    d_img.upload(image)
    eng.match_plan_image(plan, d_img.ptr, d_img.nbytes, d_f.ptr, d_a.ptr)
    assert plan.error() == 0
    assert np.array_equal(d_f.download(np.uint8, n), rf) and np.array_equal(d_a.download(np.uint8, n), ra)


@pytest.mark.gpu
def test_plan_wait_expiry_reported(gpu_lib):
    """A poll limit, or a wall-clock bound, far below the hold makes waits
    expire: the launch still ends, and the error word says so (failure forced
    from the host)."""
    from trex_amd import _lib
    from trex_amd.device import DeviceBuffer
    from trex_amd.engine import ChunkPlan, DeviceBatch, DfaEngine
    tab = dfa_byte_table(compile_regex(EVEN_A))
    rng = np.random.default_rng(3)
    lengths = np.array([200_000, 100_000], np.int64)
    data, offs = ragged_batch(rng, lengths, ord("a"), ord("c"))
    eng = DfaEngine(tab)
    batch = DeviceBatch(data, offs, layout="csr")
    plan = ChunkPlan.for_batch(batch, 256)
    plan.set_hold(200)
    d_f, d_a = DeviceBuffer(2), DeviceBuffer(2)
    eng.match_plan(plan, batch.d_bytes.ptr, d_f.ptr, d_a.ptr, max_spins=4)
    assert plan.error() & 1
    # the wall-clock bound alone (no poll limit): 2 us against a hold of
    # hundreds of microseconds per tile
    _lib.call("trex_set_wait_bound_us", 2)
    try:
        eng.match_plan(plan, batch.d_bytes.ptr, d_f.ptr, d_a.ptr, max_spins=0)
        assert plan.error() & 1
    finally:
        _lib.call("trex_set_wait_bound_us", 0)
    # the default bound (100 ms) lets the same held launch finish cleanly
    eng.match_plan(plan, batch.d_bytes.ptr, d_f.ptr, d_a.ptr)
    assert plan.error() == 0
    plan.set_hold(0)
    eng.match_plan(plan, batch.d_bytes.ptr, d_f.ptr, d_a.ptr)
    assert plan.error() == 0
    rf, ra = ocpu.dfa_walk(tab.delta, tab.accept, tab.start, data, offs)
    assert np.array_equal(d_a.download(np.uint8, 2), ra)
    with pytest.raises(_lib.TrexError):
        plan.set_hold(2001)


def test_synchronizes_and_chunk_target():
    """Synchronisation over the live bytes: permutation automata (even number
    of a's, a sum mod 15 over a window) do not synchronise although a byte
    outside their alphabet kills every state; suffix and counting automata
    do (a pair with the dead state merges once the other state dies)."""
    from trex_amd.engine import (LOOKBACK_TARGET_CHUNKS, LOOKBACK_TARGET_CHUNKS_COMPOSING, lookback_chunk_len,
                                 lookback_target_chunks, synchronizes)
    from trex_amd.tables import DfaByteTable
    delta = np.full((16, 256), 15, np.uint8)
    for b in range(0x40, 0x60):
        delta[:15, b] = (np.arange(15) + (b & 15)) % 15
    mod15 = DfaByteTable(delta, np.eye(16, dtype=np.uint8)[0], 0, 15, 16)
    assert not synchronizes(mod15.delta, 16, 15)
    want = {EVEN_A: False, "(a|b)*abb": True, "(aa|b)*": True, "(ab)*": True, "(a|b)*a(a|b)(a|b)": True}
    for rx, sync in want.items():
        t = dfa_byte_table(compile_regex(rx))
        assert synchronizes(t.delta, t.n_states, t.dead) == sync, rx
    assert lookback_target_chunks(mod15) == LOOKBACK_TARGET_CHUNKS_COMPOSING
    assert lookback_target_chunks(dfa_byte_table(compile_regex("(a|b)*abb"))) == LOOKBACK_TARGET_CHUNKS
    assert lookback_chunk_len(1 << 30, LOOKBACK_TARGET_CHUNKS_COMPOSING) == 8192
    # brute force on a small case: a synchronising word over the live bytes
    # exists exactly when the check says so (the dead state counts only when
    # a live byte reaches it)
    import itertools
    for rx, sync in want.items():
        t = dfa_byte_table(compile_regex(rx))
        alive = [s for s in range(t.n_states) if s != t.dead]
        states = alive + ([t.dead] if any(int(t.delta[s, c]) == t.dead for s in alive for c in b"ab") else [])
        found = False
        for L in range(1, 7):
            for w in itertools.product(b"ab", repeat=L):
                img = set()
                for s in states:
                    for c in w:
                        s = int(t.delta[s, c])
                    img.add(s)
                found |= len(img) == 1
        assert found == sync, rx
