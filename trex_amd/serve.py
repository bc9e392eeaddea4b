"""Continuous-batching request loop over resident patterns.

A :class:`Server` keeps compiled patterns resident on the device (DFA-16,
Boolean NFA-64 and probabilistic NFA-64 engines) and serves match requests
(one pattern, a CSR batch of strings each) as they arrive:

* admission by HBM budget: a request's bytes are copied into a device arena
  (on a copy stream, overlapping running work) as soon as the arena has room;
  requests that have not found room wait in FIFO order per lane. Small
  requests have an arena of their own, so a backlog of large ones never
  blocks their admission; they are staged in pinned memory and copied by a
  kernel on their lane's stream just ahead of their tick, so they neither
  queue behind the copies of large ones on the copy engines nor wait on
  another stream;
* four lanes, each with its own stream: DFA/NFA requests of at most
  ``small_bytes`` bytes, larger DFA/NFA requests, and the same two classes of
  probabilistic requests. A lane launches a *tick* -- every admitted request
  of the lane not yet launched -- as soon as one of its two tick slots is free,
  so a new tick queues behind the running one and requests that arrive while
  others are on the GPU join the next tick;
* a DFA/NFA tick is one persistent mixed-queue launch (:mod:`trex_amd.mix`);
  the bulk lane's persistent grid is one workgroup on half the compute units
  and it keeps one tick in flight (a second kernel on the same stream would
  be dispatched onto the remaining room as the first one's workgroups are
  placed), so a small-request tick always finds room beside it. A
  probabilistic request is a tick of its own (one ragged call, at most
  ``prob_grid`` workgroups per kernel launch, set on every probabilistic
  engine the server holds); the bulk probabilistic lane has ``prob_slots``
  slots and the small one ``prob_small_slots``, with a stream each, so a few
  long sequential launches overlap one another while the compute units they
  hold stay bounded, and small requests do not queue behind one another. In
  small probabilistic requests, sequences longer than ``prob_chunk_min``
  symbols take the parallel-in-time path (chunks multiplied concurrently, then
  chained), whose latency does not grow with the sequence length: chunks of
  ``prob_chunk_len`` symbols, or ``prob_long_chunk_len`` for sequences longer
  than ``prob_long_chunk`` (fewer chunks to chain);
* ticks complete in whatever order their events fire, so a slow tick never
  holds back the slots of faster ones;
* no call that synchronises the whole device (``hipFree``, a synchronous
  copy) runs while serving: tick buffers are reserved up front and grow
  geometrically, and mixed-queue plans are copied asynchronously;
* tick plans and sequence ranges go to the device by copy kernels, and the
  kernels write their results straight into pinned host memory, so no tick
  waits behind the copy engines; a tick is complete when its event has fired,
  and then every request of the tick gets its result and completion time, and
  its arena space is released.

The server uses nine streams (a copy stream, two DFA/NFA lanes, two slots
in the bulk probabilistic lane and four in the small one); streams that
share a hardware queue run their work in order, so a process that serves
should start with ``GPU_MAX_HW_QUEUES`` of at least 8 (the HIP runtime reads
it once, before the first GPU call; the default stream needs one too).

The reference has no serving path: one synchronous launch per request and
pattern through a process-global engine (cuda/monoid_batch.cu:596-671).
"""

from __future__ import annotations

import collections
import os
import threading
import time
import warnings
from dataclasses import dataclass, field

import numpy as np

from trex_amd import _lib
from trex_amd.device import DeviceBuffer, Event, PinnedBuffer, Stream
from trex_amd.engine import DfaEngine, NfaEngine, ProbNfaEngine
from trex_amd.mix import MixEngine, MixParams

ALIGN = 256


@dataclass(eq=False)
class Request:
    """One match request: ``data[offsets[i]:offsets[i+1]]`` is string i. Once
    ``wait()`` returns True, ``result`` holds the answer, or ``error`` says why
    there is none (``result`` is then None)."""
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
    _arena_obj: "Arena | None" = None
    _copied: "Event | None" = None
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

    def __init__(self, engines_dfa_nfa, stream: Stream):
        self.stream = stream
        self.mix = MixEngine(engines_dfa_nfa) if engines_dfa_nfa else None
        self.h_status = PinnedBuffer(16) if self.mix is not None else None   # queue words of the last run
        self.cap = 0
        self.d_begin = None
        self.h_mask = self.h_acc = None
        self.h_f64 = None
        self.event = Event()
        self.requests: list[Request] = []
        self.busy = False

    def reserve(self, n: int) -> None:
        if n <= self.cap:
            return
        cap = max(n, 2 * self.cap, 1024)
        if self.d_begin is not None:
            self.d_begin.free()
        if self.mix is not None:   # DFA/NFA results land in pinned memory
            self.h_mask, self.h_acc = PinnedBuffer(8 * cap), PinnedBuffer(cap)
        else:                      # probabilistic ranges (begin, end) and results (LL, log-mass)
            self.d_begin = DeviceBuffer(16 * cap)
            self.h_f64 = PinnedBuffer(16 * cap)
        self.cap = cap


class Server:
    def __init__(self, engines: list, hbm_budget: int = 8 << 30, small_bytes: int = 1 << 20,
                 small_params: MixParams | None = None, bulk_params: MixParams | None = None,
                 bulk_grid: int | None = None, small_budget: int | None = None, prob_slots: int = 2,
                 bulk_slots: int = 1, prob_grid: int = 32, reserve_strings: int = 1 << 18,
                 prob_chunk_min: int = 1024, prob_chunk_len: int = 256, mix_max_spins: int = 0,
                 mix_form: str = "nibble", prob_small_slots: int = 4, prob_long_chunk: int = 1 << 14,
                 prob_long_chunk_len: int = 1024):
        hwq = int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4)   # HIP's default is 4
        if hwq < 8:
            warnings.warn(f"GPU_MAX_HW_QUEUES is {hwq}: the server's nine streams share hardware queues, so "
                          "small requests wait behind long launches; start the process with 8 or more",
                          RuntimeWarning, stacklevel=2)
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
        self.prob_chunk_min = int(prob_chunk_min)
        # chunk lengths of the parallel-in-time path depend only on the
        # sequence's own length, so a result does not depend on its neighbours:
        # prob_chunk_len up to prob_long_chunk symbols, prob_long_chunk_len
        # beyond (fewer chunks to chain, so a 1 MiB sequence stays within a
        # few milliseconds)
        self.prob_chunk_len = int(prob_chunk_len)
        self.prob_long_chunk = int(prob_long_chunk)
        self.prob_long_chunk_len = int(prob_long_chunk_len)
        self.mix_max_spins = int(mix_max_spins)     # bound of every wait in a mixed launch (0: the library's)
        self.arena = Arena(hbm_budget)
        self.small_arena = Arena(small_budget if small_budget is not None else max(64 * self.small_bytes, 256 << 20))
        self.small_stage = PinnedBuffer(self.small_arena.size)   # host mirror of the small arena
        self.copy_stream = Stream()   # copies of large requests (copy engines)
        cus = _lib.device_info()["cus"]
        self.lanes = {name: {"pending": collections.deque(), "slots": [], "queue": collections.deque()}
                      for name in ("small", "bulk", "prob_small", "prob_bulk")}
        for name, lane in self.lanes.items():
            if name.startswith("prob"):
                n_slots = prob_small_slots if name == "prob_small" else prob_slots
                lane["slots"] = [_Slot(None, Stream()) for _ in range(n_slots)]
            else:
                stream = Stream()
                lane["slots"] = [_Slot(mix_engines, stream) for _ in range(bulk_slots if name == "bulk" else 2)]
            lane["params"] = (small_params or MixParams(dfa_split=4 << 10, dfa_chunk=4 << 10, nfa_split=2 << 10,
                                                        nfa_chunk=2 << 10, ck=64)) if name == "small" else \
                (bulk_params or MixParams())
            if lane["params"].form is None:   # the lanes' NFA table form (and workgroup shape)
                lane["params"] = MixParams(**{**lane["params"].__dict__, "form": mix_form})
        for slot in self.lanes["bulk"]["slots"]:
            slot.mix.set_grid(bulk_grid if bulk_grid is not None else max(1, cus // 2))
        for slot in self.lanes["small"]["slots"]:
            slot.mix.set_grid(cus)
        for lane in self.lanes.values():
            for slot in lane["slots"]:
                slot.reserve(reserve_strings)
        for e in self.engines:
            if isinstance(e, ProbNfaEngine):
                e.set_grid(prob_grid)
        self._lock = threading.Lock()   # arenas, lane queues and pending lists
        for lane in self.lanes.values():
            lane["cv"] = threading.Condition(self._lock)
        self._stop = False
        self._threads: list[threading.Thread] = []
        self.completed: list[Request] = []
        self.ticks = collections.Counter()

    # -- resident tables ---------------------------------------------------------------
    def save_tables(self, path) -> None:
        """Write every pattern's tables (pattern index = record index) to one
        table file, so a restarted server loads them instead of compiling."""
        from trex_amd.table_cache import save_tables
        save_tables(path, self.engines)

    @classmethod
    def from_tables(cls, path, **kwargs) -> "Server":
        """A server over the patterns of a table file (written by
        :meth:`save_tables` or :func:`trex_amd.table_cache.save_tables`)."""
        from trex_amd.table_cache import engines_of, load_tables
        return cls(engines_of(load_tables(path)), **kwargs)

    # -- client side ---------------------------------------------------------------
    def submit(self, req: Request) -> Request:
        """Admits the request at once when the arena has room (its bytes are
        copied on the copy stream), else queues it in FIFO order for its lane."""
        if not 0 <= req.pattern < len(self.engines):
            raise ValueError(f"pattern {req.pattern} out of range")
        req.offsets = np.ascontiguousarray(req.offsets, dtype=np.int64)
        req.data = np.ascontiguousarray(req.data, dtype=np.uint8)
        if req.offsets[0] != 0:
            raise ValueError("offsets must start at 0")
        req.lane = self._lane_of(req)
        if req.nbytes > self._arena_of(req.lane).size:
            raise ValueError("request larger than the HBM budget")
        req.arrival = time.perf_counter()
        with self._lock:
            lane = self.lanes[req.lane]
            if lane["pending"] or not self._try_admit(req):
                lane["pending"].append(req)
            else:
                lane["queue"].append(req)
                lane["cv"].notify()
        return req

    def start(self) -> "Server":
        self._stop = False
        self._threads = [threading.Thread(target=self._worker, args=(name,), name=f"trex-serve-{name}", daemon=True)
                         for name in self.lanes]
        for t in self._threads:
            t.start()
        return self

    def stop(self) -> None:
        """Serves everything submitted so far, then stops the lane threads."""
        with self._lock:
            self._stop = True
            for lane in self.lanes.values():
                lane["cv"].notify_all()
        for t in self._threads:
            t.join()
        self._threads = []

    def __enter__(self):
        return self.start()

    def __exit__(self, *exc):
        self.stop()

    # -- admission -------------------------------------------------------------------
    def _lane_of(self, req: Request) -> str:
        small = req.nbytes <= self.small_bytes
        if self.kind[req.pattern] == "mix":
            return "small" if small else "bulk"
        return "prob_small" if small else "prob_bulk"

    def _arena_of(self, lane: str) -> "Arena":
        return self.small_arena if lane.endswith("small") else self.arena

    def _try_admit(self, r: Request) -> bool:
        """Lock held. Reserves arena space and starts the copy."""
        arena = self._arena_of(r.lane)
        off = arena.alloc(max(r.nbytes, 16))
        if off < 0:
            return False
        r._arena, r._arena_obj = off, arena
        if arena is self.small_arena:
            # staged on the host now; copied by a kernel on the lane's stream just
            # ahead of its tick (no wait across streams)
            if r.nbytes:
                self.small_stage.array[off:off + r.nbytes] = r.data[:r.nbytes]
            r._copied = None
        else:
            if r.nbytes:
                _lib.call("trex_memcpy_h2d", arena.buf.ptr + off, r.data.ctypes.data, r.nbytes,
                          self.copy_stream.handle)
            r._copied = Event()
            r._copied.record(self.copy_stream)
        r.admitted = time.perf_counter()
        return True

    def _release(self, reqs: list[Request]) -> None:
        """Lock held. Frees the requests' arena space and admits waiting requests."""
        for r in reqs:
            r._arena_obj.release(r._arena, max(r.nbytes, 16))
        for lane in self.lanes.values():
            while lane["pending"] and self._try_admit(lane["pending"][0]):
                lane["queue"].append(lane["pending"].popleft())
                lane["cv"].notify()

    # -- lane workers ------------------------------------------------------------------
    def _worker(self, name: str) -> None:
        """One thread per lane: launch ticks while a slot is free, complete them
        as their events fire; every wait releases the GIL. A request joins a tick
        once its bytes are on the device: the host polls the copy's event, so
        no stream holds a wait for another (a wait blocks the hardware queue it
        sits in, which other streams may share)."""
        lane = self.lanes[name]
        prob = name.startswith("prob")
        inflight: collections.deque[_Slot] = collections.deque()
        while True:
            reqs, slot = [], None
            with self._lock:
                if not inflight:
                    while not lane["queue"] and not (self._stop and not lane["pending"]):
                        lane["cv"].wait(0.05)
                    if not lane["queue"] and self._stop and not lane["pending"]:
                        return
                slot = next((s for s in lane["slots"] if not s.busy), None)
                if lane["queue"] and slot is not None:
                    q = lane["queue"]
                    for r in list(q):
                        if r._copied is None or r._copied.query():
                            reqs.append(r)
                            if prob:   # one ragged launch per tick
                                break
                    for r in reqs:
                        q.remove(r)
            if reqs:
                stream = slot.stream
                try:
                    for r in reqs:
                        if r._copied is None and r.nbytes:   # small: staged bytes, copied on this stream
                            _lib.call("trex_memcpy_kernel", r._arena_obj.buf.ptr + r._arena,
                                      self.small_stage.ptr + r._arena, r.nbytes, stream.handle)
                    if prob:
                        self._launch_prob(slot, reqs, stream, self.prob_chunk_min if name == "prob_small" else 0)
                    else:
                        self._launch_mix(slot, reqs, stream, lane["params"])
                    slot.event.record(stream)
                except Exception as e:   # the tick's requests fail; the lane goes on
                    self._fail(slot, reqs, f"launch failed: {e}")
                    continue
                slot.requests = reqs
                slot.busy = True
                now = time.perf_counter()
                for r in reqs:
                    r.launched = now
                inflight.append(slot)
                self.ticks[name] += 1
                continue
            if inflight and len(inflight) == 1 and not lane["queue"]:
                inflight[0].event.synchronize()   # one tick running, nothing else to watch
            # complete whichever ticks have finished, in any order: a slow tick
            # does not hold back the slots of faster ones behind it
            done = next((sl for sl in inflight if sl.event.query()), None)
            if done is not None:
                inflight.remove(done)
                try:
                    self._complete(name, done)
                except Exception as e:
                    self._fail(done, done.requests, f"completion failed: {e}")
                continue
            with self._lock:   # a tick or a copy is still running: look again shortly
                lane["cv"].wait(0.0002)

    def _launch_mix(self, slot: _Slot, reqs: list[Request], stream: Stream, params: MixParams) -> None:
        base = reqs[0]._arena_obj.buf.ptr
        begin = np.concatenate([r._arena + r.offsets[:-1] for r in reqs])
        end = np.concatenate([r._arena + r.offsets[1:] for r in reqs])
        pat = np.concatenate([np.full(len(r.offsets) - 1, self.mix_index[r.pattern], np.int32) for r in reqs])
        n = len(begin)
        slot.reserve(n)
        slot.mix.plan(begin, end, pat, params)
        slot.h_status.array[:] = 0
        if n:   # results written by the kernel into pinned host memory
            slot.mix.run(base, slot.h_mask.ptr, slot.h_acc.ptr, stream, self.mix_max_spins)
            slot.mix.status_to(slot.h_status.ptr, stream)

    def _launch_prob(self, slot: _Slot, reqs: list[Request], stream: Stream, chunk_min: int = 0) -> None:
        """One ragged launch per request; with chunk_min > 0, the request's sequences
        longer than chunk_min symbols (a prefix of the longest-first order) take the
        parallel-in-time path, whose latency does not grow with their length."""
        base = reqs[0]._arena_obj.buf.ptr
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
        # ranges: staged in the pinned tick buffer, copied by a kernel into
        # d_begin (begin, then end); LL and log-mass come back into the same
        # pinned buffer, written by the kernels after the copy has run
        pin = slot.h_f64.array.view(np.int64)
        pin[:n] = begin
        pin[n:2 * n] = end
        _lib.call("trex_memcpy_kernel", slot.d_begin.ptr, slot.h_f64.ptr, 16 * n, stream.handle)
        d_end = slot.d_begin.ptr + 8 * n
        for r, (k0, m, order) in zip(reqs, slot.order):
            if not m:
                continue
            eng = self.engines[r.pattern]
            n_long = 0
            if chunk_min > 0:
                lengths = end[k0:k0 + m] - begin[k0:k0 + m]
                n_long = int(np.count_nonzero(lengths > chunk_min))
            if n_long:   # longest first: those over prob_long_chunk take the longer chunks
                lengths = end[k0:k0 + n_long] - begin[k0:k0 + n_long]
                n_longer = int(np.count_nonzero(lengths > self.prob_long_chunk))
                for a, b, clen in ((0, n_longer, self.prob_long_chunk_len), (n_longer, n_long, self.prob_chunk_len)):
                    if b > a:
                        eng.loglik_chunked_device(base, begin[k0 + a:k0 + b], end[k0 + a:k0 + b],
                                                  slot.h_f64.ptr + 8 * (k0 + a), slot.h_f64.ptr + 8 * (n + k0 + a),
                                                  clen, stream)
            if n_long < m:
                k1 = k0 + n_long
                eng.loglik_ranges_device(base, slot.d_begin.ptr + 8 * k1, d_end + 8 * k1, m - n_long,
                                         slot.h_f64.ptr + 8 * k1, slot.h_f64.ptr + 8 * (n + k1), stream)

    def _fail(self, slot: _Slot, reqs: list[Request], message: str) -> None:
        """Completes `reqs` with `error` set and no result, once the slot's stream
        has drained (its kernels may still read the requests' arena space)."""
        try:
            slot.stream.synchronize()
        except Exception:
            pass
        now = time.perf_counter()
        slot.requests = []
        with self._lock:
            self._release(reqs)
            slot.busy = False
            self.completed.extend(reqs)
        for r in reqs:
            r.result = None
            r.error = message
            r.launched = r.launched or now
            r.done = now
            r._event.set()

    def _complete(self, name: str, slot: _Slot) -> None:
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
            err = int(slot.h_status.array.view(np.uint32)[1])
            if err:   # a bounded wait expired: the tick's results are not to be trusted
                self._fail(slot, slot.requests, f"mixed launch reported error word {err:#x} (a bounded wait expired)")
                return
            k = 0
            mask = slot.h_mask.array.view(np.uint64)
            acc = slot.h_acc.array
            for r in slot.requests:
                m = len(r.offsets) - 1
                r.result = (mask[k:k + m].copy(), acc[k:k + m].copy())
                k += m
        reqs = slot.requests
        slot.requests = []
        with self._lock:
            self._release(reqs)
            slot.busy = False
            self.completed.extend(reqs)
        for r in reqs:
            r.done = now
            r._event.set()
