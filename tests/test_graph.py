"""Launches recorded as HIP graphs (trex_graph_*) and replayed: the headline
DFA launches of several patterns over an interleaved batch, and a look-back
plan over long strings. Every replay, with other eager work and host
synchronisations between the replays, equals the eager launches and the C
oracle."""

import numpy as np
import pytest

from oracle import cpu as ocpu
from trex_amd.tables import dfa_byte_table
from trex_amd.workloads import ALPHABET_32, random_dfa_pattern, ragged_batch, uniform_batch

pytestmark = pytest.mark.gpu


def test_graph_replays_dfa_launches_and_plan(gpu_lib):
    from trex_amd.device import DeviceBuffer, Graph, Stream
    from trex_amd.engine import ChunkPlan, DeviceBatch, DfaEngine
    rng = np.random.default_rng(21)
    tabs = [dfa_byte_table(random_dfa_pattern(rng, ALPHABET_32)[1]) for _ in range(3)]
    engs = [DfaEngine(t) for t in tabs]
    data, offs = uniform_batch(rng, 64 * 20, 2048, 0x40, 0x60)
    batch = DeviceBatch(data, offs, layout="interleaved")
    n = len(offs) - 1
    lengths = np.array([300_000, 5, 0, 123_457, 64_000], np.int64)
    ldata, loffs = ragged_batch(rng, lengths, 0x40, 0x60)
    lbatch = DeviceBatch(ldata, loffs, layout="csr")
    plan = ChunkPlan.for_batch(lbatch, 4096)
    nl = len(lengths)
    stream = Stream()
    outs = [(DeviceBuffer(n), DeviceBuffer(n)) for _ in engs]
    p_f, p_a = DeviceBuffer(nl), DeviceBuffer(nl)

    def launches():
        for e, (f, a) in zip(engs, outs):
            e.match_batch_device(batch, f.ptr, a.ptr, stream)
        engs[0].match_plan(plan, lbatch.d_bytes.ptr, p_f.ptr, p_a.ptr, stream)

    def results():
        got = [(f.download(np.uint8, n, stream), a.download(np.uint8, n, stream)) for f, a in outs]
        return got + [(p_f.download(np.uint8, nl, stream), p_a.download(np.uint8, nl, stream))]

    def clear():   # synchronous uploads of a value no result takes
        stream.synchronize()
        for f, a in outs:
            f.upload(np.full(n, 0xAB, np.uint8))
            a.upload(np.full(n, 0xAB, np.uint8))
        p_f.upload(np.full(nl, 0xAB, np.uint8))
        p_a.upload(np.full(nl, 0xAB, np.uint8))

    launches()
    stream.synchronize()
    eager = results()
    want = [ocpu.dfa_walk(t.delta, t.accept, t.start, data, offs) for t in tabs]
    want.append(ocpu.dfa_walk(tabs[0].delta, tabs[0].accept, tabs[0].start, ldata, loffs))
    for (gf, ga), (rf, ra) in zip(eager, want):
        assert np.array_equal(gf, rf) and np.array_equal(ga, ra)

    with Graph.capture(stream) as g:
        launches()
    other = DeviceBuffer(n), DeviceBuffer(n)
    for rep in range(4):
        clear()
        g.launch(stream)
        # eager work of the program between replays
        engs[1].match_batch_device(batch, other[0].ptr, other[1].ptr, stream)
        stream.synchronize()
        got = results()
        for k, ((gf, ga), (ef, ea)) in enumerate(zip(got, eager)):
            assert np.array_equal(gf, ef) and np.array_equal(ga, ea), (rep, k)
        assert plan.error() == 0
    g.close()


def test_graph_capture_failure_is_dropped(gpu_lib):
    """A capture that fails part way is ended and dropped; the stream is left
    usable for eager work."""
    from trex_amd.device import DeviceBuffer, Graph, Stream
    from trex_amd.engine import DeviceBatch, DfaEngine
    rng = np.random.default_rng(3)
    tab = dfa_byte_table(random_dfa_pattern(rng, ALPHABET_32)[1])
    eng = DfaEngine(tab)
    data, offs = uniform_batch(rng, 256, 300, 0x40, 0x60)
    stream = Stream()
    with pytest.raises(Exception):
        with Graph.capture(stream):
            raise RuntimeError("stop")
    batch = DeviceBatch(data, offs, layout="csr")
    f, a = DeviceBuffer(256), DeviceBuffer(256)
    eng.match_batch_device(batch, f.ptr, a.ptr, stream)
    stream.synchronize()
    rf, ra = ocpu.dfa_walk(tab.delta, tab.accept, tab.start, data, offs)
    assert np.array_equal(f.download(np.uint8, 256), rf) and np.array_equal(a.download(np.uint8, 256), ra)
