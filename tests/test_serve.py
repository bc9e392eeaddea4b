"""Continuous-batching server: the HBM arena allocator (CPU) and, on the GPU,
that every request's result is the same served in a mix of concurrent
requests as computed alone (bit-exact for DFA / NFA results, and also for the
probabilistic log-likelihoods, whose per-sequence arithmetic does not depend on
the other sequences of a launch), and that every served result agrees with the
oracle: DFA / NFA bit-exact against the C walk, log-likelihoods within 1e-3
relative of the FP64 forward pass."""

from __future__ import annotations

import numpy as np
import pytest

from oracle_checks import assert_loglik_matches_oracle, assert_walk_matches_oracle


def test_arena_first_fit_and_merge():
    from trex_amd.serve import ALIGN, Arena

    class _Fake(Arena):
        def __init__(self, n):   # no device buffer
            self.size = n
            self.free = [(0, n)]
            self.used = 0

    a = _Fake(10 * ALIGN)
    x = a.alloc(1)
    y = a.alloc(3 * ALIGN)
    z = a.alloc(2 * ALIGN + 1)
    assert (x, y, z) == (0, ALIGN, 4 * ALIGN)
    assert a.alloc(4 * ALIGN) == -1            # 3 * ALIGN left
    a.release(y, 3 * ALIGN)
    assert a.alloc(4 * ALIGN) == -1            # two separate holes
    a.release(z, 2 * ALIGN + 1)
    assert a.free == [(ALIGN, 9 * ALIGN)]      # merged with the tail
    assert a.alloc(9 * ALIGN) == ALIGN and a.used == 10 * ALIGN


def _requests(rng, n_patterns, kinds, count, max_strings, median, sigma, cap):
    from trex_amd.serve import Request
    reqs = []
    for _ in range(count):
        p = int(rng.integers(0, n_patterns))
        n = int(rng.integers(1, max_strings + 1))
        lengths = np.minimum(np.floor(np.exp(rng.normal(np.log(median), sigma, size=n))), cap).astype(np.int64)
        offsets = np.zeros(n + 1, np.int64)
        np.cumsum(lengths, out=offsets[1:])
        hi = 32 if kinds[p] == "prob" else 0x60
        lo = 0 if kinds[p] == "prob" else 0x40
        data = rng.integers(lo, hi, size=int(offsets[-1]), dtype=np.uint8)
        reqs.append(Request(p, data, offsets))
    return reqs


@pytest.mark.gpu
@pytest.mark.parametrize("mix_form", ["nibble", "slice"])
def test_gpu_served_equals_alone(gpu_lib, mix_form):
    from trex_amd.engine import DfaEngine, NfaEngine, ProbNfaEngine
    from trex_amd.serve import Server
    from trex_amd.tables import dfa_byte_table
    from trex_amd.workloads import ALPHABET_32, dirichlet_pnfa, random_dfa_pattern, random_nfa64

    rng = np.random.default_rng(21)
    engines = [DfaEngine(dfa_byte_table(random_dfa_pattern(rng, ALPHABET_32)[1])) for _ in range(3)]
    engines += [NfaEngine.from_matrices(random_nfa64(rng)) for _ in range(2)]
    engines += [ProbNfaEngine(*dirichlet_pnfa(rng, 32))]
    kinds = ["mix"] * 5 + ["prob"]
    reqs = _requests(rng, len(engines), kinds, 60, 64, 512, 2.0, 300_000)
    server = Server(engines, hbm_budget=256 << 20, small_bytes=64 << 10, prob_chunk_min=1024, prob_chunk_len=128,
                    prob_long_chunk=8192, prob_long_chunk_len=512, mix_form=mix_form)
    with server:
        for r in reqs:
            server.submit(r)
        for r in reqs:
            assert r.wait(60), "request not served"
    assert server.ticks["small"] >= 1 and server.ticks["bulk"] >= 1
    assert any(r.lane == "prob_small" and np.any(np.diff(r.offsets) > server.prob_chunk_min) for r in reqs), \
        "no small probabilistic request exercises the parallel-in-time path"
    assert any(r.lane == "prob_small" and np.any(np.diff(r.offsets) > server.prob_long_chunk) for r in reqs), \
        "no small probabilistic request takes the longer chunks"
    for r in reqs:
        e = engines[r.pattern]
        if isinstance(e, ProbNfaEngine):
            # alone: the server's routing rule per sequence (each sequence's result
            # depends only on itself and the fixed chunk length)
            lengths = np.diff(r.offsets)
            chunked = (lengths > server.prob_chunk_min) if r.lane == "prob_small" else np.zeros(len(lengths), bool)
            longer = chunked & (lengths > server.prob_long_chunk)
            ll = np.empty(len(lengths))
            lm = np.empty(len(lengths))
            for sel, kw in ((chunked & ~longer, {"chunk_len": server.prob_chunk_len}),
                            (longer, {"chunk_len": server.prob_long_chunk_len}), (~chunked, {})):
                idx = np.flatnonzero(sel)
                if len(idx):
                    parts = [r.data[r.offsets[i]:r.offsets[i + 1]] for i in idx]
                    offs = np.concatenate([[0], np.cumsum([len(x) for x in parts])]).astype(np.int64)
                    a, b = e.loglik_ragged(np.concatenate(parts) if offs[-1] else np.zeros(0, np.uint8), offs, **kw)
                    ll[idx], lm[idx] = a, b
            assert np.array_equal(r.result[0], ll) and np.array_equal(r.result[1], lm)
            assert_loglik_matches_oracle(e, r.data, r.offsets, r.result[0])
        else:
            final, acc, _, _ = e.match_arrays(r.data, r.offsets)
            assert np.array_equal(r.result[0].astype(final.dtype), final), r.pattern
            assert np.array_equal(r.result[1], acc)
            assert_walk_matches_oracle(e, r.data, r.offsets, r.result[0], r.result[1])
        assert r.arrival <= r.admitted <= r.launched <= r.done


@pytest.mark.gpu
def test_gpu_expired_waits_fail_the_tick_not_the_server(gpu_lib):
    """A spin bound of one makes the fix-up of a long NFA string give up
    waiting for its speculative chunks (the kernel's own bounded wait: it sets
    the error word and every wave still exits). The server reads the error
    word behind the launch and completes that request with an error instead
    of a wrong result; requests whose launches need no wait are served."""
    from trex_amd.engine import NfaEngine
    from trex_amd.serve import Request, Server
    from trex_amd.workloads import random_nfa64

    rng = np.random.default_rng(29)
    engines = [NfaEngine.from_matrices(random_nfa64(rng))]
    long_req = Request(0, rng.integers(0x40, 0x60, size=4 << 20, dtype=np.uint8), np.array([0, 4 << 20]))
    small = []
    for _ in range(4):   # short strings: walked whole, no waits
        lengths = rng.integers(0, 1500, size=20)
        offsets = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
        small.append(Request(0, rng.integers(0x40, 0x60, size=int(offsets[-1]), dtype=np.uint8), offsets))
    server = Server(engines, hbm_budget=64 << 20, small_bytes=64 << 10, mix_max_spins=1)
    with server:
        for r in [long_req] + small:
            server.submit(r)
        for r in [long_req] + small:
            assert r.wait(60), "request not completed"
    assert long_req.result is None and "error word" in (long_req.error or "")
    for r in small:
        assert r.error is None
        final, acc, _, _ = engines[0].match_arrays(r.data, r.offsets)
        assert np.array_equal(r.result[0], final) and np.array_equal(r.result[1], acc)
        assert_walk_matches_oracle(engines[0], r.data, r.offsets, r.result[0], r.result[1])


@pytest.mark.gpu
def test_gpu_kernel_copy_pinned_and_device(gpu_lib):
    """trex_memcpy_kernel: pinned host -> device and device -> device copies of
    sizes around the 16-byte vector and the grid limit are exact; unaligned or
    pageable sources are rejected before any launch."""
    from trex_amd import _lib
    from trex_amd.device import DeviceBuffer, PinnedBuffer, Stream
    rng = np.random.default_rng(5)
    stream = Stream()
    n_max = 1024 * 256 * 16 + 77          # beyond one grid's worth of vectors, with a tail
    src = PinnedBuffer(n_max)
    src.array[:] = rng.integers(0, 256, size=n_max, dtype=np.uint8)
    d = DeviceBuffer(n_max + 32)
    d2 = DeviceBuffer(n_max + 32)
    for n in (1, 15, 16, 17, 4096 + 3, n_max):
        _lib.call("trex_memcpy_kernel", d.ptr, src.ptr, n, stream.handle)
        _lib.call("trex_memcpy_kernel", d2.ptr, d.ptr, n, stream.handle)
        got = d2.download(np.uint8, n, stream)
        assert np.array_equal(got, src.array[:n]), n
    from trex_amd._lib import TrexError
    with pytest.raises(TrexError):
        _lib.call("trex_memcpy_kernel", d.ptr + 1, src.ptr, 64, stream.handle)
    pageable = np.zeros(4096, np.uint8)
    if pageable.ctypes.data % 16 == 0:
        with pytest.raises(TrexError):
            _lib.call("trex_memcpy_kernel", d.ptr, pageable.ctypes.data, 4096, stream.handle)


@pytest.mark.gpu
def test_gpu_restart_from_table_file(gpu_lib, tmp_path):
    """A server restarted from its table file serves the same results as the
    engines it was saved from (no pattern compiled at the restart)."""
    from trex_amd.engine import DfaEngine, NfaEngine, ProbNfaEngine
    from trex_amd.serve import Server
    from trex_amd.tables import dfa_byte_table
    from trex_amd.workloads import ALPHABET_32, dirichlet_pnfa, random_dfa_pattern, random_nfa64

    rng = np.random.default_rng(23)
    engines = [DfaEngine(dfa_byte_table(random_dfa_pattern(rng, ALPHABET_32)[1])),
               NfaEngine.from_matrices(random_nfa64(rng)), ProbNfaEngine(*dirichlet_pnfa(rng, 32))]
    first = Server(engines, hbm_budget=64 << 20, small_bytes=64 << 10)
    first.save_tables(tmp_path / "serving.trxt")
    # sequences <= 1000 symbols: every probabilistic one takes the sequential path
    reqs = _requests(rng, len(engines), ["mix", "mix", "prob"], 12, 32, 256, 1.0, 1000)
    with Server.from_tables(tmp_path / "serving.trxt", hbm_budget=64 << 20, small_bytes=64 << 10) as restarted:
        assert [type(e) for e in restarted.engines] == [type(e) for e in engines]
        for r in reqs:
            restarted.submit(r)
        for r in reqs:
            assert r.wait(60) and r.error is None
    for r in reqs:
        e = engines[r.pattern]
        if isinstance(e, ProbNfaEngine):
            ll, lm = e.loglik_ragged(r.data, r.offsets)
            assert np.array_equal(r.result[0], ll) and np.array_equal(r.result[1], lm)
            assert_loglik_matches_oracle(e, r.data, r.offsets, r.result[0])
        else:
            final, acc, _, _ = e.match_arrays(r.data, r.offsets)
            assert np.array_equal(r.result[0].astype(final.dtype), final)
            assert np.array_equal(r.result[1], acc)
            assert_walk_matches_oracle(e, r.data, r.offsets, r.result[0], r.result[1])
