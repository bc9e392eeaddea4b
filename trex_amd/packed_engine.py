"""PackedEngine -- several regexes over one batch, the reference's API on the
multi-pattern kernels.

Same constructor, methods and results as the reference
(``src/packed_engine.py:25-480``): ``PackedEngine(regexes, use_gpu=...,
gpu_mode=...)``, ``match_batch(strings) -> results[pattern][string]``,
``match_batch_timed`` and ``config_info``. The reference packs the DFAs
block-diagonally into one state vector and evolves it with matrix products;
here each pattern keeps its own DFA-16 pair table resident and one pass of
``dfa16_multi_*`` over the bytes advances every pattern (one HBM read of the
batch per four patterns).

Semantics (the reference's, ``src/packed_engine.py:91-105, 383-389``): a
character outside a pattern's alphabet leaves that pattern's state unchanged;
a character outside every pattern's alphabet leaves every state unchanged.
Built as:

* each pattern's byte table sends the characters of the other patterns'
  alphabets to the identity;
* bytes outside the union of the alphabets are deleted on the host before
  upload (deleting identity steps changes no result), so the tables may send
  them to an absorbing dead state. This is what lets every table take the
  32-byte window kernels.

Matching always runs on the GPU (``use_gpu`` and ``gpu_mode`` are accepted
for compatibility): through the multi-pattern pass when every table has at
most 16 states and the union of the alphabets lies in one aligned 32-byte
window, else through one launch per pattern of the general byte walk.
"""

from __future__ import annotations

import time
from typing import Sequence

import numpy as np

from trex_amd.regex_to_dfa import DFA, compile_regex
from trex_amd.tables import DfaByteTable


def packed_byte_table(dfa: DFA, union: Sequence[str]) -> DfaByteTable:
    """One pattern's byte table under PackedEngine semantics: its own
    characters follow the DFA, the other characters of ``union`` are the
    identity, every other byte goes to an absorbing dead state (such bytes are
    deleted before matching). The DFA's own sink serves as the dead state when
    it has one."""
    n = dfa.n_states
    own = set(dfa.alphabet)
    sink = dfa.dead_state
    dead = sink if sink is not None else n
    n_tab = n if sink is not None else n + 1
    delta = np.full((n_tab, 256), dead, dtype=np.uint8)
    for ch in union:
        b = ord(ch)
        if ch in own:
            for s in range(n):
                delta[s, b] = dfa.transitions[s][ch]
        else:
            delta[:n, b] = np.arange(n, dtype=np.uint8)
    accept = np.zeros(n_tab, np.uint8)
    for s in dfa.accept_states:
        accept[s] = 1
    return DfaByteTable(delta, accept, dfa.start, dead, n_tab)


class PackedEngine:
    def __init__(self, regexes: list[str], use_gpu: bool = True, gpu_mode: str = "auto"):
        if gpu_mode not in ("auto", "sparse", "dense"):
            raise ValueError(f"gpu_mode must be 'auto', 'sparse' or 'dense', got {gpu_mode!r}")
        self._regexes = list(regexes)
        self._n_patterns = len(self._regexes)
        self._gpu_mode = gpu_mode
        self._dfas = [compile_regex(r) for r in self._regexes]
        union = set()
        for d in self._dfas:
            union.update(d.alphabet)
        self._unified_alphabet = sorted(union)
        bad = [c for c in self._unified_alphabet if ord(c) > 255]
        if bad:
            raise ValueError(f"characters outside latin-1 cannot be matched as bytes: {bad[:5]}")
        self._tables = [packed_byte_table(d, self._unified_alphabet) for d in self._dfas]
        self._state_counts = [d.n_states for d in self._dfas]
        self._offsets = [int(x) for x in np.concatenate([[0], np.cumsum(self._state_counts)[:-1]])] \
            if self._dfas else []
        self._NP = -(-sum(self._state_counts) // 16) * 16
        self._keep = np.zeros(256, dtype=bool)
        for c in self._unified_alphabet:
            self._keep[ord(c)] = True
        self._engines = None
        self._multi = None
        self._stream = None
        if use_gpu:
            self._init_gpu()

    # -- device state (created on first use) ----------------------------------
    def _init_gpu(self):
        if self._engines is not None:
            return
        from trex_amd.engine import DfaEngine, DfaMultiEngine
        self._engines = [DfaEngine(t) for t in self._tables]
        if self._engines and all(e.info()["window_base"] >= 0 for e in self._engines):
            self._multi = DfaMultiEngine(self._engines)

    @property
    def _backend(self) -> str:
        if self._engines is None:
            return "gpu (not initialised)"
        return "gpu-multi" if self._multi is not None else "gpu"

    def _prepare(self, strings: Sequence):
        """Pack the strings and delete bytes outside every pattern's alphabet."""
        from trex_amd.engine import pack_strings
        return self._delete_outside(*pack_strings(strings))

    def _delete_outside(self, data: np.ndarray, offsets: np.ndarray):
        keep = self._keep[data]
        if keep.all():
            return data, offsets
        kept = np.zeros(len(data) + 1, np.int64)
        np.cumsum(keep, out=kept[1:])
        return np.ascontiguousarray(data[keep]), kept[offsets]

    def _run(self, data: np.ndarray, offsets: np.ndarray) -> tuple:
        """(accept uint8 [P, n], kernel_ms) on the GPU."""
        from trex_amd.device import DeviceBuffer, Event, Stream
        from trex_amd.engine import DeviceBatch
        self._init_gpu()
        n = len(offsets) - 1
        P = self._n_patterns
        batch = DeviceBatch(data, offsets, layout="csr")
        d_f, d_a = DeviceBuffer(P * n), DeviceBuffer(P * n)
        if self._stream is None:
            self._stream = Stream()
        stream = self._stream
        e0, e1 = Event(), Event()
        try:
            e0.record(stream)
            if self._multi is not None:
                self._multi.match_batch_device(batch, d_f.ptr, d_a.ptr, stream)
            else:
                for p, e in enumerate(self._engines):
                    e.match_batch_device(batch, d_f.ptr + p * n, d_a.ptr + p * n, stream)
            e1.record(stream)
            acc = d_a.download(np.uint8, P * n, stream=stream).reshape(P, n)
            kms = e0.elapsed_ms(e1)
        finally:
            batch.free()
            d_f.free()
            d_a.free()
        return acc, kms

    # -- public API (src/packed_engine.py:285-480) ----------------------------
    def match_batch(self, strings: list[str]) -> list[list[bool]]:
        """results[pattern][string]."""
        return self.match_batch_timed(strings)[0]

    def match_batch_timed(self, strings: list[str]) -> tuple[list[list[bool]], dict]:
        P = self._n_patterns
        B = len(strings)
        t0 = time.perf_counter()
        if B == 0 or P == 0:
            return [[] for _ in range(P)], {"kernel_ms": 0.0, "total_ms": 0.0, "n_strings": B,
                                            "n_patterns": P, "NP": self._NP}
        data, offsets = self._prepare(strings)
        acc, kms = self._run(data, offsets)
        results = [[bool(x) for x in row] for row in acc]
        return results, {"kernel_ms": kms, "total_ms": (time.perf_counter() - t0) * 1e3, "n_strings": B,
                         "n_patterns": P, "NP": self._NP}

    def match_arrays(self, data: np.ndarray, offsets: np.ndarray) -> np.ndarray:
        """CSR batch (uint8 bytes, int64 offsets) -> accept uint8 [P, n]."""
        from trex_amd.engine import _as_csr
        data, offsets = self._delete_outside(*_as_csr(data=data, offsets=offsets))
        if len(offsets) == 1:
            return np.zeros((self._n_patterns, 0), np.uint8)
        return self._run(data, offsets)[0]

    @property
    def config_info(self) -> dict:
        return {"n_patterns": self._n_patterns, "state_counts": list(self._state_counts),
                "offsets": list(self._offsets), "NP": self._NP,
                "unified_alphabet_size": len(self._unified_alphabet), "regexes": list(self._regexes),
                "backend": self._backend}

    def close(self) -> None:
        if self._multi is not None:
            self._multi.close()
            self._multi = None
        for e in self._engines or []:
            e.close()
        self._engines = None
