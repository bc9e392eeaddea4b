"""Mixed DFA-16 / NFA-64 batches in one persistent launch.

CPU: invariants of the work-queue plan (the host planner runs without a
device). GPU: every string's final state / state set and accept bit are
bit-exact against the C oracle, with long strings split into look-back DFA
chunks and speculative NFA chunks, including NFAs whose speculative walks
never meet their checkpoints, bytes outside the symbol window, empty strings,
and replays of a captured HIP graph.
"""

from __future__ import annotations

import numpy as np
import pytest

from trex_amd.mix import BLOCK, DFA, NFA, MixEngine, MixParams

SMALL = MixParams(dfa_split=4096, dfa_chunk=1024, nfa_split=2048, nfa_chunk=1024, ck=64)


def _batch(rng, n, lo, hi, n_patterns):
    lengths = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), size=n))).astype(np.int64)
    lengths[rng.random(n) < 0.05] = 0
    offsets = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lengths, out=offsets[1:])
    pattern_of = rng.integers(0, n_patterns, size=n).astype(np.int32)
    return offsets, pattern_of


def check_plan(m: MixEngine, begin, end, pattern_of, kinds, params: MixParams, block: int = BLOCK):
    L = m.layout()
    nb = len(L["bfirst"])
    roles = L["brole"]
    # queue sections: NFA_SPEC, DFA_CHUNK, NFA_FIXUP, then whole strings
    section = np.array([{3: 0, 1: 1, 4: 2, 0: 3, 2: 3}[r] for r in roles])
    assert np.all(np.diff(section) >= 0)
    item_batch = np.full(len(L["ibeg"]), -1)
    for b in range(nb):
        f, c, p, r = L["bfirst"][b], L["bcount"][b], L["bpattern"][b], roles[b]
        assert 1 <= c <= block
        assert (kinds[p] == DFA) == (r in (0, 1))
        assert np.all(item_batch[f:f + c] == -1)
        item_batch[f:f + c] = b
    assert np.all(item_batch >= 0)
    # whole batches longest first (weighted)
    whole = [b for b in range(nb) if roles[b] in (0, 2)]
    w = [-(L["iend"][L["bfirst"][b]] - L["ibeg"][L["bfirst"][b]]) * (10 if roles[b] == 2 else 1) for b in whole]
    assert w == sorted(w)
    # every string has exactly one output item, and its bytes are covered exactly once
    outputs = np.zeros(len(begin), dtype=np.int64)
    covered = {i: [] for i in range(len(begin))}
    role_of = roles[item_batch]
    for it in range(len(L["ibeg"])):
        r = role_of[it]
        sid = L["ib"][it]
        if sid >= 0:
            outputs[sid] += 1
        if r in (0, 2):
            covered[sid].append((L["ibeg"][it], L["iend"][it]))
        if r == 1 and not L["iflags"][it] & 1:
            pred = L["ia"][it]
            assert 0 <= pred < len(L["ibeg"]) and L["iend"][pred] == L["ibeg"][it]
            assert item_batch[pred] <= item_batch[it]
        if r == 4:
            c0, c1 = L["ibeg"][it], L["iend"][it]
            assert c1 > c0 and np.all(role_of[c0:c1] == 3)
            assert L["iflags"][c0] & 1 and not np.any(L["iflags"][c0 + 1:c1] & 1)
            assert np.all(L["iend"][c0:c1 - 1] == L["ibeg"][c0 + 1:c1])
            covered[sid].append((L["ibeg"][c0], L["iend"][c1 - 1]))
            assert np.all(L["ia"][c0:c1] == L["ia"][it])
            assert np.all(item_batch[c0:c1] < item_batch[it])
    assert np.all(outputs == 1)
    # DFA chunks: the first chunk starts the string, the last one ends it
    for it in np.flatnonzero(role_of == 1):
        if L["ib"][it] >= 0:
            s = L["ib"][it]
            assert L["iend"][it] == end[s]
            j, n_chunks = it, 1
            while not L["iflags"][j] & 1:
                j = L["ia"][j]
                n_chunks += 1
            assert L["ibeg"][j] == begin[s]
            covered[s].append((begin[s], end[s]))
            assert n_chunks == -(-(end[s] - begin[s]) // params.dfa_chunk)
    whole_strings = set(L["ib"][np.isin(role_of, (0, 2))].tolist())
    for i in range(len(begin)):
        assert covered[i] == [(begin[i], end[i])], i
        split = params.dfa_split if kinds[pattern_of[i]] == DFA else params.nfa_split
        assert (i in whole_strings) == (end[i] - begin[i] <= split)


@pytest.mark.parametrize("form,block", [("nibble", BLOCK), ("slice", 2 * BLOCK)])
def test_plan_invariants(form, block):
    rng = np.random.default_rng(0)
    kinds = [DFA, NFA, DFA, NFA, NFA]
    offsets, pattern_of = _batch(rng, 3000, 1, 200_000, len(kinds))
    m = MixEngine(kinds=kinds)
    m.set_form(form)                      # batches are cut to the form's workgroup size
    info = m.plan(offsets[:-1], offsets[1:], pattern_of, SMALL)
    assert info["chunks"] > 0 and info["fixups"] > 0
    check_plan(m, offsets[:-1], offsets[1:], pattern_of, kinds, SMALL, block)
    assert max(m.layout()["bcount"]) == block


def test_plan_edge_cases():
    m = MixEngine(kinds=[DFA, NFA])
    m.plan(np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0, np.int32), SMALL)
    assert m.info()["items"] == 0
    begin = np.array([0, 0, 10], np.int64)
    end = np.array([0, 4097, 10 + 2049], np.int64)
    m.plan(begin, end, np.array([0, 0, 1], np.int32), SMALL)
    check_plan(m, begin, end, np.array([0, 0, 1]), [DFA, NFA], SMALL)
    from trex_amd._lib import TrexError
    with pytest.raises(TrexError):
        m.plan(begin, end, np.array([0, 2, 1], np.int32), SMALL)
    with pytest.raises(TrexError):
        m.plan(begin, end, np.array([0, 0, 1], np.int32), MixParams(nfa_chunk=1000, ck=64))
    with pytest.raises(TrexError):
        m.plan(end, begin, np.array([0, 0, 1], np.int32), SMALL)


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------

def _patterns(rng, n_dfa, n_nfa, permutation_nfa=0):
    from trex_amd.engine import DfaEngine, NfaEngine
    from trex_amd.nfa_matrices import NFAMatrices
    from trex_amd.tables import dfa_byte_table
    from trex_amd.workloads import ALPHABET_32, random_dfa_pattern, random_nfa64
    pats, refs = [], []
    for _ in range(n_dfa):
        _, dfa = random_dfa_pattern(rng, ALPHABET_32)
        t = dfa_byte_table(dfa)
        pats.append(DfaEngine(t))
        refs.append(("dfa", t))
    for k in range(n_nfa):
        if k < permutation_nfa:
            # every symbol permutes the states: sets never merge, so a walk from
            # all states never meets the true one
            stack = np.zeros((32, 64, 64), np.int8)
            for c in range(32):
                stack[c, rng.permutation(64), np.arange(64)] = 1
            start = np.zeros(64, np.int8)
            start[[0, 7]] = 1
            accept = np.zeros(64, np.int8)
            accept[rng.choice(64, 4, replace=False)] = 1
            nm = NFAMatrices.from_matrices(stack, start, accept, list(ALPHABET_32))
        else:
            nm = random_nfa64(rng)
        e = NfaEngine.from_matrices(nm)
        pats.append(e)
        refs.append(("nfa", e))
    return pats, refs


def _oracle(refs, data, offsets, pattern_of):
    from oracle import cpu as ocpu
    n = len(offsets) - 1
    mask = np.zeros(n, np.uint64)
    acc = np.zeros(n, np.uint8)
    for p, (kind, r) in enumerate(refs):
        idx = np.flatnonzero(pattern_of == p)
        if len(idx) == 0:
            continue
        sub_off = np.zeros(len(idx) + 1, np.int64)
        np.cumsum(offsets[idx + 1] - offsets[idx], out=sub_off[1:])
        sub = np.concatenate([data[offsets[i]:offsets[i + 1]] for i in idx]) if sub_off[-1] else np.zeros(0, np.uint8)
        if kind == "dfa":
            f, a = ocpu.dfa_walk(r.delta, r.accept, r.start, sub, sub_off)
        else:
            f, a = ocpu.nfa_walk(r.succ, r.byte_to_sym, r.start, r.accept, sub, sub_off)
        mask[idx] = f
        acc[idx] = a
    return mask, acc


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["nibble", "slice"])
@pytest.mark.parametrize("params", [SMALL, MixParams(dfa_split=1 << 20, dfa_chunk=1 << 16, nfa_split=1 << 20,
                                                     nfa_chunk=1 << 16, ck=256)], ids=["small", "large"])
def test_gpu_mix_matches_oracle(gpu_lib, params, form):
    rng = np.random.default_rng(1)
    pats, refs = _patterns(rng, 5, 5, permutation_nfa=1)
    offsets, pattern_of = _batch(rng, 2500, 1, 300_000, len(pats))
    data = rng.integers(0x40, 0x60, size=int(offsets[-1]), dtype=np.uint8)
    # a few bytes outside the window: DFA strings go dead, NFA sets empty
    bad = rng.choice(len(data), size=20, replace=False)
    data[bad] = rng.integers(0, 0x40, size=20, dtype=np.uint8)
    m = MixEngine(pats)
    m.set_form(form)
    mask, acc = m.match_arrays(data, offsets, pattern_of, params)
    ref_mask, ref_acc = _oracle(refs, data, offsets, pattern_of)
    bad_idx = np.flatnonzero((mask != ref_mask) | (acc != ref_acc))
    assert len(bad_idx) == 0, (bad_idx[:10], pattern_of[bad_idx[:10]], np.diff(offsets)[bad_idx[:10]])


TUNED = MixParams(dfa_split=64 << 10, dfa_chunk=64 << 10, nfa_split=32 << 10, nfa_chunk=32 << 10, ck=64)


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["nibble", "slice"])
def test_gpu_mix_long_strings_tuned_split(gpu_lib, form):
    """Strings of 2-4 MiB at the tuned split and chunk lengths (DFA 64 KiB
    chunks chained by the look-back, NFA 32 KiB speculative chunks with
    fix-ups, a permutation NFA whose chunks never meet a checkpoint, and a
    DFA that never synchronises), beside short strings of every pattern:
    bit-exact against the oracle."""
    from trex_amd.engine import DfaEngine
    from trex_amd.regex_to_dfa import compile_regex
    from trex_amd.tables import dfa_byte_table
    rng = np.random.default_rng(41)
    pats, refs = _patterns(rng, 2, 3, permutation_nfa=1)
    # even number of '@' (0x40) over the window: a permutation automaton
    even = dfa_byte_table(compile_regex("([A-_]*@[A-_]*@)*[A-_]*"))
    pats.append(DfaEngine(even))
    refs.append(("dfa", even))
    n_p = len(pats)
    long_len = rng.integers(2 << 20, 4 << 20, size=2 * n_p)
    short_len = rng.integers(0, 70_000, size=300)
    lengths = np.concatenate([long_len, short_len]).astype(np.int64)
    pattern_of = np.concatenate([np.repeat(np.arange(n_p), 2), rng.integers(0, n_p, size=300)]).astype(np.int32)
    perm = rng.permutation(len(lengths))
    lengths, pattern_of = lengths[perm], pattern_of[perm]
    offsets = np.zeros(len(lengths) + 1, np.int64)
    np.cumsum(lengths, out=offsets[1:])
    data = rng.integers(0x40, 0x60, size=int(offsets[-1]), dtype=np.uint8)
    bad = rng.choice(len(data), size=12, replace=False)
    data[bad] = rng.integers(0, 0x40, size=12, dtype=np.uint8)
    m = MixEngine(pats)
    m.set_form(form)
    mask, acc = m.match_arrays(data, offsets, pattern_of, TUNED)
    info = m.info()
    assert info["chunks"] > 2 * n_p * 32 and info["fixups"] >= 6
    ref_mask, ref_acc = _oracle(refs, data, offsets, pattern_of)
    bad_idx = np.flatnonzero((mask != ref_mask) | (acc != ref_acc))
    assert len(bad_idx) == 0, (bad_idx[:10], pattern_of[bad_idx[:10]], np.diff(offsets)[bad_idx[:10]])


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["nibble", "slice"])
def test_gpu_mix_graph_replay(gpu_lib, form):
    from trex_amd.device import DeviceBuffer, Stream
    rng = np.random.default_rng(2)
    pats, refs = _patterns(rng, 3, 3)
    offsets, pattern_of = _batch(rng, 800, 1, 100_000, len(pats))
    data = rng.integers(0x40, 0x60, size=int(offsets[-1]), dtype=np.uint8)
    ref_mask, ref_acc = _oracle(refs, data, offsets, pattern_of)
    n = len(offsets) - 1
    m = MixEngine(pats)
    m.set_form(form)
    m.plan(offsets[:-1], offsets[1:], pattern_of, SMALL)
    stream = Stream()
    d_bytes = DeviceBuffer.from_array(data, stream)
    d_mask, d_acc = DeviceBuffer(8 * n), DeviceBuffer(n)
    m.capture(d_bytes.ptr, d_mask.ptr, d_acc.ptr, stream)
    for k in range(3):
        d_mask.upload(np.full(n, 0xAB, np.uint64), stream)   # eager work between replays
        d_acc.upload(np.full(n, 7, np.uint8), stream)
        m.replay(stream)
        assert m.error(stream) == 0
        assert np.array_equal(d_mask.download(np.uint64, n, stream), ref_mask), k
        assert np.array_equal(d_acc.download(np.uint8, n, stream), ref_acc), k


@pytest.mark.gpu
def test_gpu_mix_rejects_unsupported_pattern(gpu_lib):
    from trex_amd._lib import TrexError
    from trex_amd.engine import DfaEngine
    from trex_amd.regex_to_dfa import compile_regex
    from trex_amd.tables import dfa_byte_table
    e = DfaEngine(dfa_byte_table(compile_regex("(a|A)*b")))   # live bytes 0x41 and 0x61: two windows
    with pytest.raises(TrexError):
        MixEngine([e])


@pytest.mark.gpu
def test_gpu_mix_run_refuses_a_plan_for_the_other_form(gpu_lib):
    from trex_amd._lib import TrexError
    from trex_amd.device import DeviceBuffer
    rng = np.random.default_rng(4)
    pats, _ = _patterns(rng, 1, 1)
    offsets, pattern_of = _batch(rng, 50, 1, 5000, len(pats))
    data = rng.integers(0x40, 0x60, size=int(offsets[-1]), dtype=np.uint8)
    m = MixEngine(pats)
    m.plan(offsets[:-1], offsets[1:], pattern_of, SMALL)
    m.set_form("slice")
    d_bytes = DeviceBuffer.from_array(data)
    n = len(offsets) - 1
    d_mask, d_acc = DeviceBuffer(8 * n), DeviceBuffer(n)
    with pytest.raises(TrexError, match="plan again"):
        m.run(d_bytes.ptr, d_mask.ptr, d_acc.ptr)


# ---------------------------------------------------------------------------
# synthetic input generator (device) and its host restatement (oracle)
# ---------------------------------------------------------------------------

def _splitmix64(x):
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return x ^ (x >> np.uint64(31))


def test_generator_oracle_pins():
    from oracle import cpu as ocpu
    # published SplitMix64 outputs for state 0: the first two values of the sequence
    assert int(_splitmix64(np.array([0], np.uint64))[0]) == 0xE220A8397B1DCDAF
    assert int(_splitmix64(np.array([0x9E3779B97F4A7C15], np.uint64))[0]) == 0x6E789E6AA1B965F4
    i = np.arange(3, 9000, dtype=np.uint64)
    with np.errstate(over="ignore"):
        h = _splitmix64(np.uint64(123) ^ ((i >> np.uint64(3)) * np.uint64(0x9E3779B97F4A7C15)))
    ref = (0x40 + ((h >> (np.uint64(8) * (i & np.uint64(7)))) & np.uint64(31))).astype(np.uint8)
    assert np.array_equal(ocpu.gen_symbols(3, len(i), 123), ref)
    assert np.array_equal(ocpu.gen_symbols(0, 16, 5, lo=0, span=256)[7:9], ocpu.gen_symbols(7, 2, 5, lo=0, span=256))


@pytest.mark.gpu
def test_gpu_generator_matches_oracle(gpu_lib):
    from oracle import cpu as ocpu
    from trex_amd import _lib
    from trex_amd.device import DeviceBuffer, Stream
    n = 1 << 20
    stream = Stream()
    d = DeviceBuffer(n)
    _lib.call("trex_gen_symbols", d.ptr, n, 99, 0x40, 32, stream.handle)
    got = d.download(np.uint8, n, stream)
    assert np.array_equal(got, ocpu.gen_symbols(0, n, 99))
    with pytest.raises(_lib.TrexError):
        _lib.call("trex_gen_symbols", d.ptr, 100, 99, 0x40, 32, stream.handle)


@pytest.mark.gpu
def test_gpu_tune_mix_chooses_a_form_and_the_choice_matches(gpu_lib, tmp_path, monkeypatch):
    """The autotuner times both NFA table forms (two workgroup shapes); the
    chosen plan gives the oracle's results, and a second call is a cache hit."""
    monkeypatch.setenv("TREX_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    from trex_amd.autotune import tune_mix
    from trex_amd.device import DeviceBuffer, Stream
    rng = np.random.default_rng(6)
    pats, refs = _patterns(rng, 2, 2)
    offsets, pattern_of = _batch(rng, 600, 1, 60_000, len(pats))
    data = rng.integers(0x40, 0x60, size=int(offsets[-1]), dtype=np.uint8)
    stream = Stream()
    d_bytes = DeviceBuffer.from_array(data, stream)
    grid = {"dfa_split": (4096,), "dfa_chunk": (1024,), "nfa_split": (2048,), "nfa_chunk": (1024,), "ck": (64,),
            "form": ("nibble", "slice")}
    params, info = tune_mix(pats, offsets[:-1], offsets[1:], pattern_of, d_bytes.ptr, stream, sample=300, grid=grid)
    assert info["source"] == "tuned" and info["candidates"] == 2 and params.form in ("nibble", "slice")
    m = MixEngine(pats)
    mask, acc = m.match_arrays(data, offsets, pattern_of, params)
    ref_mask, ref_acc = _oracle(refs, data, offsets, pattern_of)
    assert np.array_equal(mask, ref_mask) and np.array_equal(acc, ref_acc)
    again, info2 = tune_mix(pats, offsets[:-1], offsets[1:], pattern_of, d_bytes.ptr, stream, sample=300, grid=grid)
    assert info2["source"] == "cache" and again == params
