"""Continuous-batching request loop over resident patterns.

A :class:`Server` keeps compiled patterns resident on the device (DFA-16,
Boolean NFA-64 and probabilistic NFA-64 engines) and serves match requests
(one pattern, a CSR batch of strings each) as they arrive:

* admission by HBM budget: a request's bytes are copied into a device arena
  (on a copy stream, overlapping running work) as soon as the arena has room;
  requests that have not found room wait in FIFO order per lane;
* four lanes, each with its own stream: DFA/NFA requests of at most
  ``small_bytes`` bytes, larger DFA/NFA requests, and the same two classes of
  probabilistic requests. A lane launches a *tick* -- every admitted request
  of the lane not yet launched -- as soon as one of its two tick slots is free,
  so a new tick queues behind the running one and requests that arrive while
  others are on the GPU join the next tick;
* a DFA/NFA tick is one persistent mixed-queue launch (:mod:`trex_amd.mix`);
  a probabilistic tick is one ragged launch per request. The bulk lane's
  persistent grid is one workgroup per compute unit, so a small-request tick
  can run beside it;
* results are copied back into pinned host memory on the lane stream behind
  the kernels; a tick is complete when its event has fired, and then every
  request of the tick gets its result and completion time, and its arena space
  is released.

The reference has no serving path: one synchronous launch per request and
pattern through a process-global engine (cuda/monoid_batch.cu:596-671).
"""

from __future__ import annotations

import collections
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from trex_amd import _lib
from trex_amd.device import DeviceBuffer, Event, PinnedBuffer, Stream
from trex_amd.engine import DfaEngine, NfaEngine, ProbNfaEngine
from trex_amd.mix import MixEngine, MixParams

ALIGN = 256


@dataclass(eq=False)
class Request:
    """One match request: ``data[offsets[i]:offsets[i+1]]`` is string i."""
    pattern: int
    data: np.ndarray
    offsets: np.ndarray
    arrival: float = 0.0
    admitted: float = 0.0
    launched: float = 0.0
    done: float = 0.0
    lane: str = ""
    result: tuple | None = None
    error: str | None = None
    _arena: int = -1
    _event: threading.Event = field(default_factory=threading.Event)

    @property
    def nbytes(self) -> int:
        return int(self.offsets[-1] - self.offsets[0])

    @property
    def latency_ms(self) -> float:
        return (self.done - self.arrival) * 1e3

    def wait(self, timeout: float | None = None) -> bool:
        return self._event.wait(timeout)


class Arena:
    """First-fit allocator over one device buffer (offsets aligned to 256 B)."""

    def __init__(self, nbytes: int):
        self.buf = DeviceBuffer(nbytes)
        self.size = int(nbytes)
        self.free = [(0, self.size)]   # sorted (offset, size)
        self.used = 0

    def alloc(self, n: int) -> int:
        n = max(ALIGN, -(-int(n) // ALIGN) * ALIGN)
        for i, (off, sz) in enumerate(self.free):
            if sz >= n:
                if sz == n:
                    self.free.pop(i)
                else:
                    self.free[i] = (off + n, sz - n)
                self.used += n
                return off
        return -1

    def release(self, off: int, n: int) -> None:
        n = max(ALIGN, -(-int(n) // ALIGN) * ALIGN)
        self.used -= n
        self.free.append((off, n))
        self.free.sort()
        merged = []
        for o, s in self.free:
            if merged and merged[-1][0] + merged[-1][1] == o:
                merged[-1] = (merged[-1][0], merged[-1][1] + s)
            else:
                merged.append((o, s))
        self.free = merged


class _Slot:
    """Device and pinned buffers of one in-flight tick."""

    def __init__(self, engines_dfa_nfa):
        self.mix = MixEngine(engines_dfa_nfa) if engines_dfa_nfa else None
        self.cap = 0
        self.d_mask = self.d_acc = self.d_begin = self.d_end = self.d_lm = None
        self.h_mask = self.h_acc = None
        self.h_f64 = None
        self.event = Event()
        self.requests: list[Request] = []
        self.busy = False

    def reserve(self, n: int) -> None:
        if n <= self.cap:
            return
        cap = max(n, 2 * self.cap, 1024)
        for name in ("d_mask", "d_acc", "d_begin", "d_end", "d_lm"):
            b = getattr(self, name)
            if b is not None:
                b.free()
        self.d_mask, self.d_acc = DeviceBuffer(8 * cap), DeviceBuffer(cap)
        self.d_begin, self.d_end, self.d_lm = DeviceBuffer(8 * cap), DeviceBuffer(8 * cap), DeviceBuffer(8 * cap)
        self.h_mask, self.h_acc = PinnedBuffer(8 * cap), PinnedBuffer(cap)
        self.h_f64 = PinnedBuffer(16 * cap)
        self.cap = cap


class Server:
    def __init__(self, engines: list, hbm_budget: int = 8 << 30, small_bytes: int = 1 << 20,
                 small_params: MixParams | None = None, bulk_params: MixParams | None = None,
                 bulk_grid: int | None = None):
        self.engines = list(engines)
        self.kind = []
        mix_index = {}
        mix_engines = []
        for i, e in enumerate(self.engines):
            if isinstance(e, (DfaEngine, NfaEngine)):
                self.kind.append("mix")
                mix_index[i] = len(mix_engines)
                mix_engines.append(e)
            elif isinstance(e, ProbNfaEngine):
                self.kind.append("prob")
            else:
                raise TypeError(f"pattern {i}: unsupported engine {type(e).__name__}")
        self.mix_index = mix_index
        self.small_bytes = int(small_bytes)
        self.arena = Arena(hbm_budget)
        self.copy_stream = Stream()
        self.lanes = {name: {"stream": Stream(), "pending": collections.deque(), "slots": [], "queue": collections.deque()}
                      for name in ("small", "bulk", "prob_small", "prob_bulk")}
        for name, lane in self.lanes.items():
            is_mix = not name.startswith("prob")
            lane["slots"] = [_Slot(mix_engines if is_mix else None) for _ in range(2)]
            lane["params"] = (small_params or MixParams(dfa_split=64 << 10, dfa_chunk=16 << 10, nfa_split=8 << 10,
                                                        nfa_chunk=8 << 10, ck=64)) if name == "small" else \
                (bulk_params or MixParams())
        grid = bulk_grid if bulk_grid is not None else _lib.device_info()["cus"]
        for slot in self.lanes["bulk"]["slots"]:
            slot.mix.set_grid(grid)
        for slot in self.lanes["small"]["slots"]:
            slot.mix.set_grid(_lib.device_info()["cus"])
        self._arrivals: collections.deque[Request] = collections.deque()
        self._lock = threading.Lock()
        self._wake = threading.Event()
        self._stop = False
        self._thread: threading.Thread | None = None
        self.completed: list[Request] = []
        self.ticks = collections.Counter()

    # -- client side ---------------------------------------------------------------
    def submit(self, req: Request) -> Request:
        if not 0 <= req.pattern < len(self.engines):
            raise ValueError(f"pattern {req.pattern} out of range")
        req.offsets = np.ascontiguousarray(req.offsets, dtype=np.int64)
        req.data = np.ascontiguousarray(req.data, dtype=np.uint8)
        if req.offsets[0] != 0:
            raise ValueError("offsets must start at 0")
        if req.nbytes > self.arena.size:
            raise ValueError("request larger than the HBM budget")
        req.arrival = time.perf_counter()
        with self._lock:
            self._arrivals.append(req)
        self._wake.set()
        return req

    def start(self) -> "Server":
        self._stop = False
        self._thread = threading.Thread(target=self._loop, name="trex-serve", daemon=True)
        self._thread.start()
        return self

    def stop(self) -> None:
        self._stop = True
        self._wake.set()
        if self._thread:
            self._thread.join()
            self._thread = None

    def __enter__(self):
        return self.start()

    def __exit__(self, *exc):
        self.stop()

    # -- dispatcher ----------------------------------------------------------------
    def _lane_of(self, req: Request) -> str:
        small = req.nbytes <= self.small_bytes
        if self.kind[req.pattern] == "mix":
            return "small" if small else "bulk"
        return "prob_small" if small else "prob_bulk"

    def _loop(self) -> None:
        while True:
            with self._lock:
                while self._arrivals:
                    r = self._arrivals.popleft()
                    r.lane = self._lane_of(r)
                    self.lanes[r.lane]["pending"].append(r)
            progressed = self._admit()
            progressed |= self._complete()
            progressed |= self._launch()
            if self._stop and not any(l["pending"] or l["queue"] or any(s.busy for s in l["slots"])
                                      for l in self.lanes.values()):
                return
            if not progressed:
                self._wake.wait(0.0002)
                self._wake.clear()

    def _admit(self) -> bool:
        """Copies pending requests into the arena while it has room (FIFO per lane)."""
        progressed = False
        for name in ("small", "prob_small", "bulk", "prob_bulk"):
            lane = self.lanes[name]
            while lane["pending"]:
                r = lane["pending"][0]
                off = self.arena.alloc(max(r.nbytes, 16))
                if off < 0:
                    break
                lane["pending"].popleft()
                r._arena = off
                if r.nbytes:
                    _lib.call("trex_memcpy_h2d", self.arena.buf.ptr + off, r.data.ctypes.data, r.nbytes,
                              self.copy_stream.handle)
                r.admitted = time.perf_counter()
                lane["queue"].append(r)
                progressed = True
        return progressed

    def _launch(self) -> bool:
        progressed = False
        for name, lane in self.lanes.items():
            if not lane["queue"]:
                continue
            slot = next((s for s in lane["slots"] if not s.busy), None)
            if slot is None:
                continue
            reqs = list(lane["queue"])
            lane["queue"].clear()
            copied = Event()
            copied.record(self.copy_stream)
            stream = lane["stream"]
            stream.wait(copied)
            if name.startswith("prob"):
                self._launch_prob(slot, reqs, stream)
            else:
                self._launch_mix(slot, reqs, stream, lane["params"])
            slot.event.record(stream)
            slot.requests = reqs
            slot.busy = True
            now = time.perf_counter()
            for r in reqs:
                r.launched = now
            self.ticks[name] += 1
            progressed = True
        return progressed

    def _launch_mix(self, slot: _Slot, reqs: list[Request], stream: Stream, params: MixParams) -> None:
        base = self.arena.buf.ptr
        begin = np.concatenate([r._arena + r.offsets[:-1] for r in reqs])
        end = np.concatenate([r._arena + r.offsets[1:] for r in reqs])
        pat = np.concatenate([np.full(len(r.offsets) - 1, self.mix_index[r.pattern], np.int32) for r in reqs])
        n = len(begin)
        slot.reserve(n)
        slot.mix.plan(begin, end, pat, params)
        if n:
            slot.mix.run(base, slot.d_mask.ptr, slot.d_acc.ptr, stream)
            _lib.call("trex_memcpy_d2h", slot.h_mask.ptr, slot.d_mask.ptr, 8 * n, stream.handle)
            _lib.call("trex_memcpy_d2h", slot.h_acc.ptr, slot.d_acc.ptr, n, stream.handle)

    def _launch_prob(self, slot: _Slot, reqs: list[Request], stream: Stream) -> None:
        base = self.arena.buf.ptr
        n = sum(len(r.offsets) - 1 for r in reqs)
        slot.reserve(n)
        begin = np.empty(n, np.int64)
        end = np.empty(n, np.int64)
        k = 0
        slot.order = []
        for r in reqs:
            m = len(r.offsets) - 1
            lengths = r.offsets[1:] - r.offsets[:-1]
            order = np.argsort(-lengths, kind="stable")      # similar lengths share a group of 64
            begin[k:k + m] = r._arena + r.offsets[:-1][order]
            end[k:k + m] = r._arena + r.offsets[1:][order]
            slot.order.append((k, m, order))
            k += m
        if n == 0:
            return
        pin = slot.h_f64.array.view(np.int64)
        pin[:n] = begin
        pin[n:2 * n] = end
        _lib.call("trex_memcpy_h2d", slot.d_begin.ptr, slot.h_f64.ptr, 8 * n, stream.handle)
        _lib.call("trex_memcpy_h2d", slot.d_end.ptr, slot.h_f64.ptr + 8 * n, 8 * n, stream.handle)
        for r, (k0, m, _) in zip(reqs, slot.order):
            if m:
                self.engines[r.pattern].loglik_ranges_device(base, slot.d_begin.ptr + 8 * k0, slot.d_end.ptr + 8 * k0,
                                                             m, slot.d_mask.ptr + 8 * k0, slot.d_lm.ptr + 8 * k0,
                                                             stream)
        # LL (d_mask as float64) and log-mass, copied back behind the kernels
        _lib.call("trex_memcpy_d2h", slot.h_f64.ptr, slot.d_mask.ptr, 8 * n, stream.handle)
        _lib.call("trex_memcpy_d2h", slot.h_f64.ptr + 8 * n, slot.d_lm.ptr, 8 * n, stream.handle)

    def _complete(self) -> bool:
        progressed = False
        for name, lane in self.lanes.items():
            for slot in lane["slots"]:
                if not slot.busy or not slot.event.query():
                    continue
                now = time.perf_counter()
                if name.startswith("prob"):
                    n = sum(len(r.offsets) - 1 for r in slot.requests)
                    f64 = slot.h_f64.array.view(np.float64)
                    for r, (k0, m, order) in zip(slot.requests, slot.order):
                        ll = np.empty(m, np.float64)
                        lm = np.empty(m, np.float64)
                        ll[order] = f64[k0:k0 + m]
                        lm[order] = f64[n + k0:n + k0 + m]
                        r.result = (ll, lm)
                else:
                    k = 0
                    mask = slot.h_mask.array.view(np.uint64)
                    acc = slot.h_acc.array
                    for r in slot.requests:
                        m = len(r.offsets) - 1
                        r.result = (mask[k:k + m].copy(), acc[k:k + m].copy())
                        k += m
                for r in slot.requests:
                    r.done = now
                    self.arena.release(r._arena, max(r.nbytes, 16))
                    r._event.set()
                    self.completed.append(r)
                slot.requests = []
                slot.busy = False
                progressed = True
        return progressed
