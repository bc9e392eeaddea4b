"""Mixed DFA-16 / Boolean NFA-64 batches in one persistent launch
(``trex_mix_*``).

A :class:`MixEngine` holds many resident patterns. ``plan`` turns a batch
(string i = bytes[begin[i]:end[i]) against pattern pattern_of[i]) into a
length-sorted work queue; ``run`` is one launch sequence on a stream (it can
be captured once into a HIP graph and replayed). Long strings are cut into
chunks: DFA chunks are chained by the decoupled look-back, NFA chunks are
walked speculatively from the set of all states and fixed up exactly.

The reference has one global engine per library and one synchronous launch
per pattern and batch (cuda/monoid_batch.cu:596-671).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from trex_amd import _lib
from trex_amd.device import DeviceBuffer, Stream

DFA, NFA = 0, 1
ROLES = {0: "dfa_whole", 1: "dfa_chunk", 2: "nfa_whole", 3: "nfa_spec", 4: "nfa_fixup"}
BLOCK = 512   # lane items per work batch (one workgroup)


@dataclass
class MixParams:
    """Split lengths, chunk lengths and checkpoint spacing of a plan, and the
    NFA table form ("nibble" / "slice"; None keeps the engine's form)."""
    dfa_split: int = 256 << 10
    dfa_chunk: int = 64 << 10
    nfa_split: int = 32 << 10
    nfa_chunk: int = 32 << 10
    ck: int = 64
    form: str | None = None

    def astuple(self):
        return (int(self.dfa_split), int(self.dfa_chunk), int(self.nfa_split), int(self.nfa_chunk), int(self.ck))


MIX_FORMS = {"nibble": 0, "slice": 1}


class MixEngine:
    def __init__(self, patterns=None, kinds=None):
        """``patterns``: DfaEngine / NfaEngine objects (kept alive here). With
        ``kinds`` only (a list of DFA / NFA), the engine plans without a device."""
        from trex_amd.engine import DfaEngine, NfaEngine
        if patterns is not None:
            kinds = []
            for p in patterns:
                if isinstance(p, DfaEngine):
                    kinds.append(DFA)
                elif isinstance(p, NfaEngine):
                    kinds.append(NFA)
                else:
                    raise TypeError(f"not a DfaEngine / NfaEngine: {type(p).__name__}")
        if not kinds:
            raise ValueError("need at least one pattern")
        self.patterns = list(patterns) if patterns is not None else None
        self.kinds = np.ascontiguousarray(kinds, dtype=np.int32)
        handles = None
        if self.patterns is not None:
            handles = (ctypes.c_void_p * len(self.patterns))(*[p.handle for p in self.patterns])
        self._handles = handles
        h = ctypes.c_void_p()
        _lib.call("trex_mix_create", self.kinds.ctypes.data, handles, len(self.kinds), ctypes.byref(h))
        self.handle = h.value
        self.n_strings = 0
        self.params = None

    def plan(self, begin, end, pattern_of, params: MixParams | None = None) -> dict:
        params = params or MixParams()
        self._begin = np.ascontiguousarray(begin, dtype=np.int64)
        self._end = np.ascontiguousarray(end, dtype=np.int64)
        self._pat = np.ascontiguousarray(pattern_of, dtype=np.int32)
        n = len(self._begin)
        if len(self._end) != n or len(self._pat) != n:
            raise ValueError("begin, end and pattern_of must have the same length")
        if params.form is not None:
            self.set_form(params.form)
        _lib.call("trex_mix_plan", self.handle, self._begin.ctypes.data, self._end.ctypes.data, self._pat.ctypes.data,
                  n, *params.astuple())
        self.n_strings = n
        self.params = params
        return self.info()

    def info(self) -> dict:
        v = [ctypes.c_int64() for _ in range(4)]
        _lib.call("trex_mix_info", self.handle, *[ctypes.byref(x) for x in v])
        return dict(zip(("items", "batches", "chunks", "fixups"), (x.value for x in v)))

    def layout(self) -> dict:
        inf = self.info()
        ni, nb = inf["items"], inf["batches"]
        out = {"ibeg": np.empty(ni, np.int64), "iend": np.empty(ni, np.int64), "ia": np.empty(ni, np.int32),
               "ib": np.empty(ni, np.int32), "iflags": np.empty(ni, np.uint32), "bfirst": np.empty(nb, np.int32),
               "bcount": np.empty(nb, np.int32), "bpattern": np.empty(nb, np.int32), "brole": np.empty(nb, np.int32)}
        _lib.call("trex_mix_layout", self.handle, *[a.ctypes.data for a in out.values()])
        return out

    def run(self, d_bytes: int, d_mask: int, d_accept: int, stream: Stream | None = None, max_spins: int = 0) -> None:
        _lib.call("trex_mix_run", self.handle, d_bytes, d_mask, d_accept, int(max_spins),
                  stream.handle if stream else None)

    def set_grid(self, workgroups: int) -> None:
        """Persistent workgroups of later runs (0: two per compute unit, one
        with the "slice" form)."""
        _lib.call("trex_mix_set_grid", self.handle, int(workgroups))

    def set_form(self, form: str) -> None:
        """NFA table form: "nibble" (default, 16 LDS reads per NFA byte) or
        "slice" (12 reads, 1024-thread workgroups). Call before :meth:`plan`."""
        _lib.call("trex_mix_set_form", self.handle, MIX_FORMS[form])

    def status_to(self, dst: int, stream: Stream | None = None) -> None:
        """Enqueue a copy of the last run's queue words (ticket, error word, converged,
        flag; 16 bytes) to dst (device or pinned host memory) behind the run."""
        _lib.call("trex_mix_status", self.handle, dst, stream.handle if stream else None)

    def error(self, stream: Stream | None = None) -> int:
        e = ctypes.c_uint32()
        _lib.call("trex_mix_error", self.handle, ctypes.byref(e), stream.handle if stream else None)
        return e.value

    def converged(self, stream: Stream | None = None) -> int:
        """Speculative NFA chunks of the last run whose fix-up met a checkpoint."""
        c = ctypes.c_uint32()
        _lib.call("trex_mix_stats", self.handle, ctypes.byref(c), stream.handle if stream else None)
        return c.value

    def capture(self, d_bytes: int, d_mask: int, d_accept: int, stream: Stream, max_spins: int = 0) -> None:
        _lib.call("trex_mix_capture", self.handle, d_bytes, d_mask, d_accept, int(max_spins), stream.handle)

    def replay(self, stream: Stream) -> None:
        _lib.call("trex_mix_replay", self.handle, stream.handle)

    def match_arrays(self, data: np.ndarray, offsets: np.ndarray, pattern_of: np.ndarray,
                     params: MixParams | None = None, timed: bool = False):
        """Host CSR batch -> (final uint64[n], accept uint8[n]); final is the DFA
        state or the NFA state set of each string. With ``timed``, also the
        launch sequence's time in ms (HIP events on its stream)."""
        from trex_amd.device import Event
        data = np.ascontiguousarray(data, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        n = len(offsets) - 1
        self.plan(offsets[:-1], offsets[1:], pattern_of, params)
        stream = Stream()
        d_bytes = DeviceBuffer.from_array(data if len(data) else np.zeros(16, np.uint8), stream)
        d_mask, d_acc = DeviceBuffer(8 * max(n, 1)), DeviceBuffer(max(n, 1))
        e0, e1 = Event(), Event()
        e0.record(stream)
        self.run(d_bytes.ptr, d_mask.ptr, d_acc.ptr, stream)
        e1.record(stream)
        if self.error(stream):
            raise _lib.TrexError(-7, "trex_mix_run", "a bounded wait expired")
        mask = d_mask.download(np.uint64, n, stream)
        acc = d_acc.download(np.uint8, n, stream)
        for b in (d_bytes, d_mask, d_acc):
            b.free()
        if timed:
            return mask, acc, e0.elapsed_ms(e1)
        return mask, acc

    def close(self) -> None:
        if self.handle:
            _lib.lib().trex_mix_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
