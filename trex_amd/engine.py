"""Pattern engines over the C ABI.

:class:`DfaEngine` keeps one compiled DFA byte table resident on the device
and matches CSR string batches with the HIP kernels. Results are the final
state and the accept bit of every string.

Batches are given either as Python strings / bytes (packed by
:func:`pack_strings`, latin-1 for ``str`` as in the reference bridges,
``src/gpu_bridge_monoid_batch.py:89-101``) or directly as a ``uint8`` byte
array plus ``int64`` offsets, which avoids any per-string Python work.
"""

from __future__ import annotations

import ctypes
from typing import Iterable, Sequence

import numpy as np

from trex_amd import _lib
from trex_amd.device import DeviceBuffer, Stream
from trex_amd.tables import DfaByteTable

VARIANTS = {"auto": 0, "k1_generic": 1, "k1_16": 2, "k2_window": 3, "k2_window_f": 5, "k2_window_fr": 6}

# Staggered HBM layout: strings are kept in groups of one wave (64 lanes); group g
# is moved STAGGER_GAP * g bytes further than in the packed CSR layout, so that
# waves streaming strings of equal length start at different offsets modulo the
# HBM channel interleave instead of all hitting the same channels.
STAGGER_GROUP = 64
STAGGER_GAP = 256
# Interleaved HBM layout: the 64 strings of a wave are stored as rows of
# 64 x 16 bytes (vector k of lane j at group base + (64 k + j) * 16), so every
# wave load reads one contiguous 1 KiB row; groups are IL_GAP bytes apart.
IL_GROUP = 64
IL_GAP = 256
VARIANT_NAMES = {v: k for k, v in VARIANTS.items()}


def pack_strings(strings: Iterable) -> tuple[np.ndarray, np.ndarray]:
    """CSR-pack ``str`` (latin-1) / ``bytes`` items: returns (uint8 bytes, int64 offsets).

    A list of only ``str`` is joined first and encoded once, and a list of only
    ``bytes`` / ``bytearray`` is joined once (10 M characters in 10,000
    strings: about 3.5 ms, against 7.5 ms encoding item by item); mixed
    lists take the per-item path. This replaces the reference's latin-1 join
    plus Python offset loop (src/gpu_bridge_monoid_batch.py:89-101)."""
    items = strings if isinstance(strings, list) else list(strings)
    n = len(items)
    offsets = np.zeros(n + 1, dtype=np.int64)
    if n == 0:
        return np.zeros(0, np.uint8), offsets
    kinds = set(map(type, items))
    if kinds == {str}:
        joined = "".join(items).encode("latin-1")
    elif kinds <= {bytes, bytearray}:
        joined = b"".join(items)
    else:
        items = [s.encode("latin-1") if isinstance(s, str) else bytes(s) for s in items]
        joined = b"".join(items)
    np.cumsum(np.fromiter(map(len, items), dtype=np.int64, count=n), out=offsets[1:])
    return np.frombuffer(joined, dtype=np.uint8), offsets


def _as_csr(strings=None, data=None, offsets=None) -> tuple[np.ndarray, np.ndarray]:
    if strings is not None:
        return pack_strings(strings)
    data = np.ascontiguousarray(data, dtype=np.uint8)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    if offsets.ndim != 1 or len(offsets) < 1:
        raise ValueError("offsets must be a 1-D array with n_strings + 1 entries")
    if len(offsets) > 1 and (offsets[0] < 0 or offsets[-1] > len(data) or np.any(np.diff(offsets) < 0)):
        raise ValueError("offsets must be non-decreasing and lie within the byte array")
    return data, offsets


class DfaEngine:
    """One resident DFA byte table (``trex_dfa_create``)."""

    def __init__(self, table: DfaByteTable):
        self.table = table
        h = ctypes.c_void_p()
        _lib.call("trex_dfa_create",
                  table.delta.ctypes.data_as(_lib.u8p), table.accept.ctypes.data_as(_lib.u8p),
                  table.n_states, table.start, ctypes.byref(h))
        self.handle = h.value

    # -- configuration -----------------------------------------------------
    def info(self) -> dict:
        v, w, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _lib.call("trex_dfa_info", self.handle, ctypes.byref(v), ctypes.byref(w), ctypes.byref(d))
        return {"variant": VARIANT_NAMES.get(v.value, v.value), "window_base": w.value,
                "dead_state": d.value}

    def set_variant(self, name: str) -> None:
        _lib.call("trex_dfa_set_variant", self.handle, VARIANTS[name])

    def set_prefetch(self, vectors: int) -> None:
        _lib.call("trex_dfa_set_prefetch", self.handle, int(vectors))

    def set_rotation(self, mode: int) -> None:
        """Read order of contiguous strings: -1 (default) rotated per wave when
        the table synchronises, 0 never, 1 always (``trex_dfa_set_rotation``)."""
        _lib.call("trex_dfa_set_rotation", self.handle, int(mode))

    def rotation(self) -> tuple[bool, bool]:
        """(launches over contiguous strings use the rotated read order,
        the table synchronises)."""
        a, sy = ctypes.c_int(), ctypes.c_int()
        _lib.call("trex_dfa_rotation", self.handle, ctypes.byref(a), ctypes.byref(sy))
        return bool(a.value), bool(sy.value)

    # -- matching ------------------------------------------------------------
    def match_device(self, d_bytes: int, d_offsets: int, n_strings: int, d_final: int, d_accept: int,
                     stream: Stream | None = None) -> None:
        """Asynchronous dispatch on a device CSR batch (inputs resident in HBM)."""
        _lib.call("trex_dfa_match", self.handle, d_bytes, d_offsets, int(n_strings), d_final, d_accept,
                  stream.handle if stream else None)

    def match_batch_device(self, batch: "DeviceBatch", d_final: int, d_accept: int,
                           stream: Stream | None = None) -> None:
        """Asynchronous dispatch on a :class:`DeviceBatch` (any layout)."""
        fn = "trex_dfa_match_interleaved" if batch.layout == "interleaved" else "trex_dfa_match_ranges"
        _lib.call(fn, self.handle, batch.d_bytes.ptr, batch.d_begin.ptr, batch.d_end.ptr,
                  batch.n_strings, d_final, d_accept, stream.handle if stream else None)

    def match_plan(self, plan: "ChunkPlan", d_bytes: int, d_final: int, d_accept: int,
                   stream: Stream | None = None, max_spins: int = 0) -> None:
        """Asynchronous chunked dispatch with the decoupled look-back
        (strings of any length; <= 16-state tables)."""
        _lib.call("trex_dfa_match_plan", self.handle, plan.handle, d_bytes, d_final, d_accept, int(max_spins),
                  stream.handle if stream else None)

    def match_plan_image(self, plan: "ChunkPlan", d_image: int, image_bytes: int, d_final: int, d_accept: int,
                         stream: Stream | None = None, max_spins: int = 0) -> None:
        """``match_plan`` over the plan's chunk-interleaved image in device
        memory (``ChunkPlan.build_image``, uploaded; ``image_bytes`` = the
        buffer's size, checked against the plan's image size): one wave load
        reads 1 KiB contiguously."""
        _lib.call("trex_dfa_match_plan_image", self.handle, plan.handle, d_image, int(image_bytes), d_final,
                  d_accept, int(max_spins), stream.handle if stream else None)

    def match_arrays(self, data: np.ndarray, offsets: np.ndarray):
        """Host CSR batch -> (final_state uint8[n], accept uint8[n], kernel_ms, total_ms)."""
        data, offsets = _as_csr(data=data, offsets=offsets)
        n = len(offsets) - 1
        final = np.empty(n, dtype=np.uint8)
        acc = np.empty(n, dtype=np.uint8)
        kms, tms = ctypes.c_float(), ctypes.c_float()
        if n > 0:
            _lib.call("trex_dfa_match_host", self.handle, data.ctypes.data if len(data) else None,
                      offsets.ctypes.data, n, final.ctypes.data, acc.ctypes.data,
                      ctypes.byref(kms), ctypes.byref(tms))
        return final, acc, kms.value, tms.value

    def match_strings(self, strings: Sequence):
        data, offsets = pack_strings(strings)
        return self.match_arrays(data, offsets)

    def close(self) -> None:
        if self.handle:
            _lib.lib().trex_dfa_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def staggered_ranges(offsets: np.ndarray) -> tuple[np.ndarray, np.ndarray, int]:
    """begin/end of every string in the staggered layout, and its total size."""
    offsets = np.asarray(offsets, dtype=np.int64)
    n = len(offsets) - 1
    shift = (np.arange(n, dtype=np.int64) // STAGGER_GROUP) * STAGGER_GAP - offsets[0]
    begin = offsets[:-1] + shift
    end = offsets[1:] + shift
    n_groups = -(-n // STAGGER_GROUP)
    size = int(offsets[-1] - offsets[0]) + n_groups * STAGGER_GAP
    return begin, end, size


def interleaved_image(data: np.ndarray, offsets: np.ndarray) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(device image, begin, end) of a CSR batch in the interleaved layout:
    group g of IL_GROUP consecutive strings holds rows of IL_GROUP x 16 bytes,
    vector k of its lane j at base_g + (IL_GROUP * k + j) * 16; a group has as
    many rows as its longest string has (partly filled) vectors; begin[i] is
    the address of string i's vector 0 and end[i] - begin[i] its length."""
    data, offsets = _as_csr(data=data, offsets=offsets)
    n = len(offsets) - 1
    lengths = offsets[1:] - offsets[:-1]
    n_groups = -(-n // IL_GROUP)
    vecs = -(-lengths // 16)
    gvecs = np.zeros(n_groups, np.int64)
    if n:
        np.maximum.at(gvecs, np.arange(n) // IL_GROUP, vecs)
    block = gvecs * (16 * IL_GROUP)
    base = np.zeros(n_groups, np.int64)
    if n_groups > 1:
        np.cumsum(block[:-1] + IL_GAP, out=base[1:])
    size = int(base[-1] + block[-1]) if n_groups else 0
    image = np.zeros(size + IL_GAP, np.uint8)   # a trailing gap keeps every group's block + gap in range
    lane = np.arange(n, dtype=np.int64) % IL_GROUP
    begin = base[np.arange(n) // IL_GROUP] + 16 * lane if n else np.zeros(0, np.int64)
    end = begin + lengths
    uniform = n and n % IL_GROUP == 0 and np.all(lengths == lengths[0]) and lengths[0] % 16 == 0
    if uniform:   # one strided copy: [group][lane][vector][16] -> [group][vector][lane][16]
        L = int(lengths[0])
        if L:
            blk = IL_GROUP * L
            img = image[:n_groups * (blk + IL_GAP)].reshape(n_groups, blk + IL_GAP)
            src = data[offsets[0]:offsets[0] + n * L].reshape(n_groups, IL_GROUP, L // 16, 16)
            img[:, :blk] = src.transpose(0, 2, 1, 3).reshape(n_groups, blk)
    else:
        for g in range(n_groups):
            gv = int(gvecs[g])
            if gv == 0:
                continue
            rows = np.zeros((gv, IL_GROUP, 16), np.uint8)
            for j in range(min(IL_GROUP, n - g * IL_GROUP)):
                i = g * IL_GROUP + j
                L = int(lengths[i])
                if L:
                    pad = np.zeros(int(vecs[i]) * 16, np.uint8)
                    pad[:L] = data[offsets[i]:offsets[i + 1]]
                    rows[:int(vecs[i]), j, :] = pad.reshape(-1, 16)
            image[base[g]:base[g] + block[g]] = rows.reshape(-1)
    return image, np.ascontiguousarray(begin, np.int64), np.ascontiguousarray(end, np.int64)


MULTI_FORMS = {"k2f": 5, "k2w": 3, "k2fp": 7, "k2fc": 8}
MULTI_FORM_NAMES = {v: k for k, v in MULTI_FORMS.items()}


class DfaMultiEngine:
    """Several DFA-16 patterns matched in one pass over the bytes
    (``trex_dfa_multi_*``): each launch reads the batch once and advances up
    to ``group`` (default 4) patterns. Every pattern must take the window
    kernels (``DfaEngine.info()["window_base"] >= 0``). Results are
    pattern-major arrays ``[n_patterns, n_strings]`` equal to each pattern's
    own :class:`DfaEngine` results."""

    def __init__(self, engines: Sequence[DfaEngine]):
        engines = list(engines)
        if not engines:
            raise ValueError("no patterns")
        handles = (ctypes.c_void_p * len(engines))(*[e.handle for e in engines])
        h = ctypes.c_void_p()
        _lib.call("trex_dfa_multi_create", handles, len(engines), ctypes.byref(h))
        self.handle = h.value
        self.n_patterns = len(engines)

    def info(self) -> dict:
        n, f, g = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _lib.call("trex_dfa_multi_info", self.handle, ctypes.byref(n), ctypes.byref(f), ctypes.byref(g))
        return {"n_patterns": n.value, "form": MULTI_FORM_NAMES.get(f.value, f.value), "group": g.value}

    def set_form(self, name: str) -> None:
        _lib.call("trex_dfa_multi_set_form", self.handle, MULTI_FORMS[name])

    def set_group(self, group: int) -> None:
        _lib.call("trex_dfa_multi_set_group", self.handle, int(group))

    def match_batch_device(self, batch: "DeviceBatch", d_final: int, d_accept: int,
                           stream: Stream | None = None) -> None:
        """Asynchronous dispatch on a :class:`DeviceBatch` (any layout); the
        outputs hold ``n_patterns * n_strings`` bytes each."""
        fn = "trex_dfa_multi_match_interleaved" if batch.layout == "interleaved" else "trex_dfa_multi_match_ranges"
        _lib.call(fn, self.handle, batch.d_bytes.ptr, batch.d_begin.ptr, batch.d_end.ptr, batch.n_strings,
                  d_final, d_accept, stream.handle if stream else None)

    def match_arrays(self, data: np.ndarray, offsets: np.ndarray, layout: str = "csr"):
        """Host CSR batch -> (final_state uint8[P, n], accept uint8[P, n])."""
        data, offsets = _as_csr(data=data, offsets=offsets)
        n = len(offsets) - 1
        final = np.empty((self.n_patterns, n), dtype=np.uint8)
        acc = np.empty((self.n_patterns, n), dtype=np.uint8)
        if n == 0:
            return final, acc
        batch = DeviceBatch(data, offsets, layout=layout)
        d_f, d_a = DeviceBuffer(final.size), DeviceBuffer(acc.size)
        try:
            self.match_batch_device(batch, d_f.ptr, d_a.ptr)   # null stream; the downloads synchronise
            final[:] = d_f.download(np.uint8, final.size).reshape(final.shape)
            acc[:] = d_a.download(np.uint8, acc.size).reshape(acc.shape)
        finally:
            batch.free()
            d_f.free()
            d_a.free()
        return final, acc

    def match_strings(self, strings: Sequence):
        return self.match_arrays(*pack_strings(strings))

    def close(self) -> None:
        if self.handle:
            _lib.lib().trex_dfa_multi_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBatch:
    """A string batch resident in HBM: bytes plus per-string [begin, end) ranges.

    ``layout="csr"`` keeps the packed CSR bytes (begin = offsets[:-1],
    end = offsets[1:]); ``layout="staggered"`` moves each group of 64 strings
    256 bytes further than the previous one (at most 0.4% more memory for
    4 KiB strings), see STAGGER_GAP; ``layout="interleaved"`` stores each
    group of 64 strings row by row (:func:`interleaved_image`; DFA kernels
    only, the groups' strings padded to the longest one).
    """

    def __init__(self, data: np.ndarray, offsets: np.ndarray, layout: str = "staggered"):
        data, offsets = _as_csr(data=data, offsets=offsets)
        self.n_strings = len(offsets) - 1
        self.layout = layout
        self.payload_bytes = int(offsets[-1] - offsets[0])
        base = int(offsets[0])
        if layout == "csr":
            begin = offsets[:-1] - base
            end = offsets[1:] - base
            self.d_bytes = DeviceBuffer(max(16, -(-self.payload_bytes // 16) * 16))
            self.d_bytes.upload(data[base:base + self.payload_bytes])
        elif layout == "staggered":
            begin, end, size = staggered_ranges(offsets)
            self.d_bytes = DeviceBuffer(max(16, -(-size // 16) * 16))
            for g0 in range(0, self.n_strings, STAGGER_GROUP):
                g1 = min(g0 + STAGGER_GROUP, self.n_strings)
                lo, hi = int(offsets[g0]), int(offsets[g1])
                if hi > lo:
                    self.d_bytes.upload(data[lo:hi], offset=int(begin[g0]))
        elif layout == "interleaved":
            image, begin, end = interleaved_image(data, offsets)
            self.d_bytes = DeviceBuffer(-(-len(image) // 16) * 16)
            self.d_bytes.upload(image)
            del image
        else:
            raise ValueError(f"unknown layout {layout!r}")
        self.begin = np.ascontiguousarray(begin, dtype=np.int64)
        self.end = np.ascontiguousarray(end, dtype=np.int64)
        self.d_begin = DeviceBuffer.from_array(self.begin)
        self.d_end = DeviceBuffer.from_array(self.end)

    def free(self) -> None:
        for b in (self.d_bytes, self.d_begin, self.d_end):
            b.free()


NFA_VARIANTS = {"bytes": 1, "window": 2, "window5": 3, "window6": 4}
NFA_VARIANT_NAMES = {v: k for k, v in NFA_VARIANTS.items()}


# Strings longer than this are cut into chunks matched in parallel (the
# look-back scan for DFA-16 tables, speculative chunks for NFA-64).
LONG_STRING = 1 << 20


def dfa_as_nfa(table: DfaByteTable):
    """The NFA-64 form of a DFA byte table of at most 64 states, for the
    chunked NFA path: one symbol per live byte (a byte that sends some
    non-dead state elsewhere than the dead state), ``succ[c][s] = 1 <<
    delta[s][b_c]``, every other byte rejecting (the empty set stands for
    the dead state). Returns (succ, byte_to_sym, start mask, accept mask), or
    None when the table has more than 64 states, no dead state, or more
    live bytes than the NFA kernels hold."""
    n, dead = int(table.n_states), int(table.dead)
    if n > 64 or dead < 0:
        return None
    delta = np.asarray(table.delta, dtype=np.int64)[:n]
    alive = np.array([s for s in range(n) if s != dead], dtype=np.int64)
    live = np.flatnonzero((delta[alive] != dead).any(axis=0)) if alive.size else np.zeros(0, np.int64)
    if live.size > 62:
        return None
    succ = np.zeros((max(live.size, 1), 64), np.uint64)
    for c, b in enumerate(live):
        succ[c, :n] = np.uint64(1) << delta[:, b].astype(np.uint64)
    b2s = np.full(256, -1, np.int16)
    b2s[live] = np.arange(live.size, dtype=np.int16)
    accept = 0
    for s in np.flatnonzero(np.asarray(table.accept)[:n]):
        accept |= 1 << int(s)
    return succ, b2s, 1 << int(table.start), accept


def nfa_mask_to_state(mask: np.ndarray, dead: int) -> np.ndarray:
    """Final DFA states from the final sets of :func:`dfa_as_nfa`'s automaton
    (a set is one state, or empty for the dead state)."""
    mask = np.asarray(mask, dtype=np.uint64)
    out = np.full(mask.shape, dead, dtype=np.uint8)
    nz = mask != 0
    out[nz] = np.log2(mask[nz].astype(np.float64)).astype(np.uint8)
    return out


def nfa_byte_map(nm, unknown: str = "reject") -> np.ndarray:
    """byte -> symbol index of an :class:`NFAMatrices`; bytes outside the
    alphabet reject (``simulate_nfa``, src/nfa_matrices.py:149-152) or, with
    ``unknown="skip"``, leave the state set unchanged."""
    fill = -1 if unknown == "reject" else -2
    b2s = np.full(256, fill, dtype=np.int16)
    for i, ch in enumerate(nm.alphabet):
        if ord(ch) < 256:
            b2s[ord(ch)] = i
    return b2s


class NfaEngine:
    """One resident Boolean NFA of at most 64 states (``trex_nfa_create``)."""

    def __init__(self, succ: np.ndarray, byte_to_sym: np.ndarray, start: int, accept: int):
        self.succ = np.ascontiguousarray(succ, dtype=np.uint64)
        self.byte_to_sym = np.ascontiguousarray(byte_to_sym, dtype=np.int16)
        if self.succ.ndim != 2 or self.succ.shape[1] != 64:
            raise ValueError(f"succ must have shape [n_symbols, 64], got {self.succ.shape}")
        if self.byte_to_sym.shape != (256,):
            raise ValueError("byte_to_sym must have 256 entries")
        self.start = int(start)
        self.accept = int(accept)
        self._mix = None
        h = ctypes.c_void_p()
        _lib.call("trex_nfa_create", self.succ.ctypes.data if self.succ.size else None, self.succ.shape[0],
                  self.byte_to_sym.ctypes.data, self.start, self.accept, ctypes.byref(h))
        self.handle = h.value

    @classmethod
    def from_matrices(cls, nm, unknown: str = "reject") -> "NfaEngine":
        return cls(nm.successor_masks(), nfa_byte_map(nm, unknown), nm.start_mask(), nm.accept_bits())

    def info(self) -> dict:
        v, w = ctypes.c_int(), ctypes.c_int()
        _lib.call("trex_nfa_info", self.handle, ctypes.byref(v), ctypes.byref(w))
        return {"variant": NFA_VARIANT_NAMES.get(v.value, v.value), "window_base": w.value}

    def set_variant(self, name: str) -> None:
        _lib.call("trex_nfa_set_variant", self.handle, NFA_VARIANTS[name])

    def match_batch_device(self, batch: "DeviceBatch", d_mask: int, d_accept: int,
                           stream: Stream | None = None) -> None:
        if batch.layout == "interleaved":
            raise ValueError("the NFA kernels take CSR or staggered batches, not the interleaved layout")
        _lib.call("trex_nfa_match_ranges", self.handle, batch.d_bytes.ptr, batch.d_begin.ptr, batch.d_end.ptr,
                  batch.n_strings, d_mask, d_accept, stream.handle if stream else None)

    def can_split(self) -> bool:
        """Whether long strings can be cut into speculative chunks (the mixed
        queue's NFA roles, which take automata whose live bytes lie in one
        aligned 32-byte window)."""
        return self.info()["window_base"] >= 0

    def match_arrays(self, data: np.ndarray, offsets: np.ndarray):
        """Host CSR batch -> (final mask uint64[n], accept uint8[n], kernel_ms, total_ms).
        A batch holding a string over ``LONG_STRING`` bytes goes through the
        split path when the automaton allows it (:meth:`match_split`)."""
        data, offsets = _as_csr(data=data, offsets=offsets)
        n = len(offsets) - 1
        if n > 0 and int(np.diff(offsets).max()) > LONG_STRING and self.can_split():
            return self.match_split(data, offsets)
        mask = np.empty(n, dtype=np.uint64)
        acc = np.empty(n, dtype=np.uint8)
        kms, tms = ctypes.c_float(), ctypes.c_float()
        if n > 0:
            _lib.call("trex_nfa_match_host", self.handle, data.ctypes.data if len(data) else None,
                      offsets.ctypes.data, n, mask.ctypes.data, acc.ctypes.data, ctypes.byref(kms), ctypes.byref(tms))
        return mask, acc, kms.value, tms.value

    def match_split(self, data: np.ndarray, offsets: np.ndarray, params=None):
        """Host CSR batch through the mixed queue with this one pattern: strings
        over the split length are cut into chunks walked in parallel from the
        set of all states (the first from the start set), each with
        checkpoints, then one fix-up lane per string walks the chunks again
        from the true entry set until it meets a checkpoint; shorter strings
        are walked whole. The reference walks every string on one thread
        (cuda/nfa_tc_evolution.cu:683-745). Returns (final mask, accept,
        kernel_ms, total_ms)."""
        import time
        from trex_amd.mix import MixEngine, MixParams
        if not self.can_split():
            raise ValueError("the split path needs an automaton whose live bytes lie in one 32-byte window")
        if self._mix is None:
            self._mix = MixEngine([self])
        t0 = time.perf_counter()
        n = len(offsets) - 1
        mask, acc, kms = self._mix.match_arrays(data, offsets, np.zeros(n, np.int32), params or MixParams(),
                                                timed=True)
        return mask, acc, kms, (time.perf_counter() - t0) * 1e3

    def match_strings(self, strings: Sequence):
        return self.match_arrays(*pack_strings(strings))

    def close(self) -> None:
        if self._mix is not None:
            self._mix.close()
            self._mix = None
        if self.handle:
            _lib.lib().trex_nfa_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


PNFA_MODES = {"f16x2": 1, "bf16x2": 2, "f16": 3, "bf16": 4}
PNFA_MODE_NAMES = {v: k for k, v in PNFA_MODES.items()}


class ProbNfaEngine:
    """One resident probabilistic NFA of 64 states (``trex_pnfa_create``).

    ``T``: [n_symbols, 64, 64] float32, row = destination, column = source (the
    layout of the reference's ``T_matrices``, cuda/nfa_tc_evolution.cu:943-953);
    ``start``/``accept``: [64]. ``loglik`` returns per sequence
    ``LL = log(accept . T[c_L] ... T[c_1] start)`` and ``log(1 . T[c_L] ... start)``.
    """

    N = 64

    def __init__(self, T: np.ndarray, start: np.ndarray, accept: np.ndarray, mode: str | None = None):
        self.T = np.ascontiguousarray(T, dtype=np.float32)
        self.start = np.ascontiguousarray(start, dtype=np.float32)
        self.accept = np.ascontiguousarray(accept, dtype=np.float32)
        if self.T.ndim != 3 or self.T.shape[1:] != (self.N, self.N):
            raise ValueError(f"T must have shape [n_symbols, 64, 64], got {self.T.shape}")
        if self.start.shape != (self.N,) or self.accept.shape != (self.N,):
            raise ValueError("start and accept must have 64 entries")
        h = ctypes.c_void_p()
        _lib.call("trex_pnfa_create", self.T.ctypes.data, self.T.shape[0], self.start.ctypes.data,
                  self.accept.ctypes.data, ctypes.byref(h))
        self.handle = h.value
        self.n_symbols = self.T.shape[0]
        if mode is not None:
            self.set_mode(mode)

    def info(self) -> dict:
        m, n, ok, w = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _lib.call("trex_pnfa_info", self.handle, ctypes.byref(m), ctypes.byref(n), ctypes.byref(ok), ctypes.byref(w))
        return {"mode": PNFA_MODE_NAMES[m.value], "n_symbols": n.value, "f16_ok": bool(ok.value), "waves": w.value}

    def set_waves(self, waves: int) -> None:
        _lib.call("trex_pnfa_set_waves", self.handle, int(waves))

    def set_grid(self, workgroups: int) -> None:
        """Caps the workgroups of one kernel launch (0: one per 64 sequences);
        the groups of 64 sequences then go out in consecutive launches."""
        _lib.call("trex_pnfa_set_grid", self.handle, int(workgroups))

    def set_mode(self, mode: str) -> None:
        if mode not in PNFA_MODES:
            raise ValueError(f"unknown mode {mode!r}; choose from {sorted(PNFA_MODES)}")
        _lib.call("trex_pnfa_set_mode", self.handle, PNFA_MODES[mode])

    def rounded(self, mode: str):
        """(T_rounded [σ,64,64] float32, col_scale [σ,64] float64) of ``mode``: the
        operator s' = T_rounded (col_scale[c] * s) the kernels apply on every step but the last."""
        t = np.empty_like(self.T)
        sc = np.empty((self.n_symbols, self.N), dtype=np.float64)
        _lib.call("trex_pnfa_rounded", self.handle, PNFA_MODES[mode], t.ctypes.data, sc.ctypes.data)
        return t, sc

    def loglik_device(self, d_symbols: int, n_seq: int, length: int, d_ll: int, d_logmass: int,
                      stream: Stream | None = None) -> None:
        _lib.call("trex_pnfa_loglik", self.handle, d_symbols, n_seq, length, d_ll, d_logmass,
                  stream.handle if stream else None)

    def loglik_ranges_device(self, d_symbols: int, d_begin: int, d_end: int, n_seq: int, d_ll: int,
                             d_logmass: int, stream: Stream | None = None) -> None:
        """Ragged sequences: sequence j = symbols[begin[j]:end[j]] (device arrays)."""
        _lib.call("trex_pnfa_loglik_ranges", self.handle, d_symbols, d_begin, d_end, int(n_seq), d_ll, d_logmass,
                  stream.handle if stream else None)

    def loglik_chunked_device(self, d_symbols: int, begin: np.ndarray, end: np.ndarray, d_ll: int, d_logmass: int,
                              chunk_len: int = 0, stream: Stream | None = None) -> None:
        """Parallel in time: sequence j = symbols[begin[j]:end[j]] with begin/end host arrays;
        chunk products of chunk_len symbols (0: about four chunks per compute unit, at least
        256 symbols) run concurrently and are chained per sequence (latency set by chunk_len,
        not the length; the error grows with the number of chunks)."""
        begin = np.ascontiguousarray(begin, dtype=np.int64)
        end = np.ascontiguousarray(end, dtype=np.int64)
        if begin.shape != end.shape:
            raise ValueError("begin and end differ in length")
        _lib.call("trex_pnfa_loglik_chunked", self.handle, d_symbols, begin.ctypes.data, end.ctypes.data, len(begin),
                  int(chunk_len), d_ll, d_logmass, stream.handle if stream else None)

    def loglik_ragged(self, data: np.ndarray, offsets: np.ndarray, chunk_len: int | None = None):
        """Host CSR batch of symbol sequences -> (ll float64[n], logmass float64[n]). Sequences are
        grouped by length (longest first) for the launch and returned in input order. With
        ``chunk_len`` (0: automatic) the parallel-in-time path runs instead of the batch kernel."""
        data, offsets = _as_csr(data=data, offsets=offsets)
        n = len(offsets) - 1
        ll = np.empty(n, np.float64)
        lm = np.empty(n, np.float64)
        if n == 0:
            return ll, lm
        if chunk_len is not None:
            stream = Stream()
            d_sym = DeviceBuffer.from_array(data if len(data) else np.zeros(16, np.uint8), stream)
            d_ll, d_lm = DeviceBuffer(8 * n), DeviceBuffer(8 * n)
            self.loglik_chunked_device(d_sym.ptr, offsets[:-1], offsets[1:], d_ll.ptr, d_lm.ptr, chunk_len, stream)
            ll[:] = d_ll.download(np.float64, n, stream)
            lm[:] = d_lm.download(np.float64, n, stream)
            for buf in (d_sym, d_ll, d_lm):
                buf.free()
            return ll, lm
        order = np.argsort(-(offsets[1:] - offsets[:-1]), kind="stable")
        stream = Stream()
        d_sym = DeviceBuffer.from_array(data if len(data) else np.zeros(16, np.uint8), stream)
        d_b = DeviceBuffer.from_array(offsets[:-1][order], stream)
        d_e = DeviceBuffer.from_array(offsets[1:][order], stream)
        d_ll, d_lm = DeviceBuffer(8 * n), DeviceBuffer(8 * n)
        self.loglik_ranges_device(d_sym.ptr, d_b.ptr, d_e.ptr, n, d_ll.ptr, d_lm.ptr, stream)
        ll[order] = d_ll.download(np.float64, n, stream)
        lm[order] = d_lm.download(np.float64, n, stream)
        for buf in (d_sym, d_b, d_e, d_ll, d_lm):
            buf.free()
        return ll, lm

    def loglik(self, symbols: np.ndarray):
        """symbols [n_seq, L] uint8 -> (ll float64[n], logmass float64[n], kernel_ms, total_ms)."""
        symbols = np.ascontiguousarray(symbols, dtype=np.uint8)
        if symbols.ndim != 2:
            raise ValueError("symbols must be [n_seq, length]")
        n, L = symbols.shape
        ll = np.empty(n, dtype=np.float64)
        lm = np.empty(n, dtype=np.float64)
        kms, tms = ctypes.c_float(), ctypes.c_float()
        if n > 0:
            _lib.call("trex_pnfa_loglik_host", self.handle, symbols.ctypes.data if symbols.size else None, n, L,
                      ll.ctypes.data, lm.ctypes.data, ctypes.byref(kms), ctypes.byref(tms))
        return ll, lm, kms.value, tms.value

    def close(self) -> None:
        if self.handle:
            _lib.lib().trex_pnfa_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# Chunks the look-back scan aims for: one wave of 64 chunk lanes per SIMD
# (1,024 SIMDs on 256 compute units). A walking wave issues a whole vector's
# table reads at once, so one wave per SIMD keeps it busy; smaller chunks add
# per-chunk work (metadata, the composed prefix until the function turns
# constant, the status word) and more, shorter HBM streams.
LOOKBACK_TARGET_CHUNKS = 1024 * 64
# Automata that never synchronise compose every byte pair (about 5x the
# instructions of a walk), so they take twice the chunks: two waves per SIMD.
LOOKBACK_TARGET_CHUNKS_COMPOSING = 2 * 1024 * 64


def synchronizes(delta: np.ndarray, n_states: int, dead: int = -1) -> bool:
    """True when some word over the live bytes sends every pair of states
    to one state (then almost every chunk's function turns constant within a
    few bytes and the chunk skips the look-back). Live bytes: those that send
    some non-dead state elsewhere than the dead state; a byte that kills
    everything would synchronise any automaton but never occurs in matched
    text. Backward search from the diagonal over pairs of states; the dead
    state is one of them when a live byte can reach it (a pair with it
    merges once the other state dies)."""
    n = int(n_states)
    d = np.asarray(delta, dtype=np.int64)[:n]
    alive = [s for s in range(n) if s != dead]
    live = [b for b in range(d.shape[1]) if any(int(d[s, b]) != dead for s in alive)]
    if not live:
        return True
    reaches_dead = 0 <= dead < n and any(int(d[s, b]) == dead for s in alive for b in live)
    live_states = alive + ([dead] if reaches_dead else [])
    if len(live_states) <= 1:
        return True
    # pairs (p, q), p < q, that reach the diagonal
    good = np.zeros((n, n), bool)
    np.fill_diagonal(good, True)
    changed = True
    while changed:
        changed = False
        for p in live_states:
            for q in live_states:
                if p < q and not good[p, q]:
                    if any(good[d[p, b], d[q, b]] for b in live):
                        good[p, q] = good[q, p] = True
                        changed = True
    return all(good[p, q] for p in live_states for q in live_states)


def lookback_target_chunks(table) -> int:
    """Chunk count to aim for with this byte table (``DfaByteTable``): one
    wave per SIMD when the automaton synchronises, two when its chunks keep
    composing."""
    if synchronizes(table.delta, table.n_states, table.dead):
        return LOOKBACK_TARGET_CHUNKS
    return LOOKBACK_TARGET_CHUNKS_COMPOSING


def lookback_chunk_len(total_bytes: int, target_chunks: int = LOOKBACK_TARGET_CHUNKS) -> int:
    """Chunk length for ``total_bytes`` of long strings: the power of two
    nearest ``total_bytes / target_chunks``, within 1 KiB .. 64 KiB. 1 GiB
    gives 16 KiB (16 strings of 64 MiB on one MI355X: 0.25-0.27 ms per GiB,
    against 0.28-0.32 ms at 8 KiB and 0.80 ms at 64 KiB)."""
    want = max(int(total_bytes), 1) / target_chunks
    k = int(round(np.log2(max(want, 1.0))))
    return int(min(max(1 << k, 1 << 10), 1 << 16))


class ChunkPlan:
    """Strings cut into chunks of ``chunk_len`` bytes for the look-back scan
    (``trex_plan_create``). ``begin``/``end`` are byte ranges in the device
    buffer the plan is matched against; ``chunk_len=None`` picks
    ``lookback_chunk_len`` of the strings' total length."""

    def __init__(self, begin: np.ndarray, end: np.ndarray, chunk_len: int | None = None):
        begin = np.ascontiguousarray(begin, dtype=np.int64)
        end = np.ascontiguousarray(end, dtype=np.int64)
        if begin.shape != end.shape:
            raise ValueError("begin and end must have the same length")
        if chunk_len is None:
            chunk_len = lookback_chunk_len(int((end - begin).sum()) if len(begin) else 0)
        h = ctypes.c_void_p()
        _lib.call("trex_plan_create", begin.ctypes.data, end.ctypes.data, len(begin), int(chunk_len),
                  ctypes.byref(h))
        self.handle = h.value
        self.n_strings = len(begin)
        self.chunk_len = int(chunk_len)
        nc, ns = ctypes.c_int64(), ctypes.c_int64()
        _lib.call("trex_plan_info", self.handle, ctypes.byref(nc), ctypes.byref(ns))
        self.n_chunks = nc.value

    @classmethod
    def for_batch(cls, batch: "DeviceBatch", chunk_len: int | None = None) -> "ChunkPlan":
        return cls(batch.begin, batch.end, chunk_len)

    def set_chains(self, chains: int) -> None:
        """Dependent chains per lane in the scan's first phase (1 or 2: the
        two halves of each chunk walked side by side)."""
        _lib.call("trex_plan_set_chains", self.handle, int(chains))

    def set_image_gap(self, gap: int) -> None:
        """Bytes after each tile of the chunk-interleaved image (a multiple
        of 16; 256 by default). Build the image again after changing it."""
        _lib.call("trex_plan_set_image_gap", self.handle, int(gap))

    def set_hold(self, hold: int) -> None:
        """Test hook: later launches hold back each tile's publication,
        in reverse order within groups of 8 tiles (0 = off)."""
        _lib.call("trex_plan_set_hold", self.handle, int(hold))

    def image_size(self) -> int:
        """Bytes of the plan's chunk-interleaved image."""
        size = ctypes.c_int64()
        _lib.call("trex_plan_image_size", self.handle, ctypes.byref(size))
        return size.value

    def build_image(self, data: np.ndarray) -> np.ndarray:
        """The plan's chunk-interleaved image of ``data`` (the host bytes its
        ranges index): tile t = chunks 64t .. 64t+63 in processing order,
        vector i of the tile's k-th chunk at tile_off[t] + (64 i + k) * 16,
        each tile followed by the plan's image gap."""
        data = np.ascontiguousarray(data, dtype=np.uint8)
        out = np.empty(max(self.image_size(), 16), np.uint8)
        _lib.call("trex_plan_build_image", self.handle, data.ctypes.data, data.nbytes, out.ctypes.data, out.nbytes)
        return out

    def error(self, stream: Stream | None = None) -> int:
        err = ctypes.c_uint32()
        _lib.call("trex_plan_error", self.handle, ctypes.byref(err), stream.handle if stream else None)
        return err.value

    def close(self) -> None:
        if self.handle:
            _lib.lib().trex_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def upload_batch(data: np.ndarray, offsets: np.ndarray) -> tuple[DeviceBuffer, DeviceBuffer]:
    """Copy a CSR batch into device memory (bytes padded to a 16-byte multiple)."""
    data, offsets = _as_csr(data=data, offsets=offsets)
    d_bytes = DeviceBuffer(max(16, -(-len(data) // 16) * 16))
    d_bytes.upload(data)
    d_offsets = DeviceBuffer.from_array(offsets)
    return d_bytes, d_offsets
