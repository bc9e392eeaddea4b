"""OptimizedEngine -- the reference's Python matching API on the HIP engine.

Same constructor, configuration vocabulary and public methods as the
reference (``src/optimized_engine.py:29-425``): ``OptimizedEngine(regex,
config=None, dfa_state_cap=64, monoid_cap=65536)``, ``config_info``,
``match_batch``, ``match_batch_timed``, plus ``match`` (the north star's name
for ``match_batch``) and the array form ``match_arrays``.

Every configuration matches on the GPU through the C ABI; none falls back to
a CPU simulation, and matching raises if the native library or the device is
missing. Each configuration keeps the reference behaviour for characters
outside the pattern's alphabet:

* ``None`` (auto), ``baseline``, ``batched+gpu``, ``kgram+gpu``,
  ``fp16_tc+gpu``: the DFA (``DFA.simulate``, src/regex_to_dfa.py:224-233):
  such a character rejects (dead state);
* ``monoid``, ``monoid+kgram``, ``monoid+gpu``, ``monoid_batch+gpu``: the
  monoid walk (``simulate_monoid``, src/monoid.py:185-190): identity;
* ``prefix+gpu``: the tmap walk (``precompute_tmap``, src/simulation.py:283-289):
  identity;
* ``nfa``: the NFA (``simulate_nfa``, src/nfa_matrices.py:149-152): rejects.

GPU engines are created on first use, so construction and ``config_info``
work on a host without a GPU.

The compiled result (byte table or NFA successor masks plus the selection
fields) is a :class:`trex_amd.table_cache.CompiledPattern`. With a table
cache (the ``cache`` argument, or the directory named by
``$TREX_AMD_TABLE_CACHE``) a pattern compiled once under the same options is
loaded instead of compiled again.
"""

from __future__ import annotations

import time
from typing import Optional, Sequence

import numpy as np

from trex_amd.engine import LONG_STRING, nfa_mask_to_state
from trex_amd.kgram import auto_k
from trex_amd.monoid import compute_monoid, monoid_byte_table
from trex_amd.nfa_matrices import compile_nfa_matrices
from trex_amd.regex_to_dfa import compile_regex
from trex_amd.simulation import DFAMatrices, precompute_tmap
from trex_amd.table_cache import CompiledPattern, TableCache, default_cache, pattern_key
from trex_amd.tables import MAX_TABLE_STATES, dfa_byte_table, tmap_byte_table

CONFIGS = (None, "monoid", "monoid+kgram", "baseline", "nfa", "monoid+gpu", "batched+gpu", "kgram+gpu",
           "monoid_batch+gpu", "prefix+gpu", "fp16_tc+gpu")


class OptimizedEngine:
    def __init__(self, regex: str, config: Optional[str] = None, dfa_state_cap: int = 64,
                 monoid_cap: int = 65536, cache=None):
        """``cache``: a :class:`TableCache`, a directory for one, ``False`` for
        none, or ``None`` for the default (``$TREX_AMD_TABLE_CACHE`` if set)."""
        if config not in CONFIGS:
            raise ValueError(f"Unknown config: {config!r}. Choose from None, " +
                             ", ".join(repr(c) for c in CONFIGS[1:]) + ".")
        self._regex = regex
        self._config = config
        self._dfa_state_cap = dfa_state_cap
        self._monoid_cap = monoid_cap
        self._dfa = None
        self._dm = None
        self._md = None
        self._nm = None
        self._table = None          # DfaByteTable walked on the GPU (DFA, monoid or tmap form)
        self._engine = None         # DfaEngine / NfaEngine, created on first match
        self._nfa_engine = None     # NFA-64 form of a wider DFA for long strings (False: none)
        self._lookback_target = None  # look-back chunk target of the table
        self._representation = None
        self._scan_backend = None
        self._selection_reason = None
        self._kgram_k = None
        if cache is None:
            cache = default_cache()
        elif cache is False:
            cache = None
        elif not isinstance(cache, TableCache):
            cache = TableCache(cache)
        key = pattern_key(regex, config, dfa_state_cap, monoid_cap)
        cp = cache.get(key) if cache is not None else None
        if cp is None:
            self._select(config)
            cp = self._record()
            if cache is not None:
                cache.put(key, cp)
        self._compiled = cp
        self._table = cp.table
        self._representation = cp.representation
        self._scan_backend = cp.scan_backend
        self._selection_reason = cp.selection_reason
        self._kgram_k = cp.kgram_k

    def _select(self, config):
        {
            None: self._auto_select,
            "monoid": lambda: self._select_monoid("monoid", "forced config='monoid'"),
            "monoid+kgram": self._select_monoid_kgram,
            "baseline": lambda: self._select_dfa("sequential", "forced config='baseline'"),
            "nfa": lambda: self._select_nfa("forced config='nfa'"),
            "monoid+gpu": lambda: self._select_monoid("monoid+gpu", "GPU monoid scan"),
            "batched+gpu": lambda: self._select_dfa("batched+gpu", "GPU batched DFA walk"),
            "kgram+gpu": self._select_kgram_gpu,
            "monoid_batch+gpu": lambda: self._select_monoid("monoid_batch+gpu", None, limit=255),
            "prefix+gpu": self._select_prefix,
            "fp16_tc+gpu": lambda: self._select_dfa("fp16_tc+gpu", "GPU DFA walk (fp16_tc+gpu name kept)"),
        }[config]()

    # -- selection (host only) ------------------------------------------------
    def _build_dfa(self):
        if self._dfa is None:
            self._dfa = compile_regex(self._regex)
            self._dm = DFAMatrices(self._dfa)

    def _select_dfa(self, backend: str, reason: str):
        self._build_dfa()
        self._table = dfa_byte_table(self._dfa)
        self._representation = "dfa"
        self._scan_backend = backend
        self._selection_reason = reason

    def _select_nfa(self, reason: str):
        self._nm = compile_nfa_matrices(self._regex)
        if self._nm.n_states_raw > 64:
            raise RuntimeError(f"NFA has {self._nm.n_states_raw} states; the bit-parallel engine holds 64")
        self._representation = "nfa"
        self._scan_backend = "nfa"
        self._selection_reason = reason

    def _select_monoid(self, backend: str, reason: Optional[str], limit: int = MAX_TABLE_STATES):
        self._build_dfa()
        self._md = compute_monoid(self._dm, max_size=self._monoid_cap)
        if self._md is None:
            raise RuntimeError("compute_monoid returned None (monoid too large). "
                               "Increase monoid_cap or use a different config.")
        if self._md.size > limit:
            raise RuntimeError(f"Monoid size {self._md.size} > {limit}; use another config")
        self._table = monoid_byte_table(self._md, self._dm)
        self._representation = "dfa"
        self._scan_backend = backend
        self._selection_reason = reason or (f"GPU monoid batch (M={self._md.size}, "
                                            f"sigma={len(self._dm.alphabet)})")

    def _select_monoid_kgram(self):
        self._select_monoid("monoid+kgram", "forced config='monoid+kgram'")
        self._kgram_k = auto_k(len(self._dfa.alphabet))

    def _select_kgram_gpu(self):
        self._select_dfa("kgram+gpu", None)
        self._kgram_k = 2
        self._selection_reason = f"GPU K-gram DFA walk (two bytes per table step, N={self._table.n_states})"

    def _select_prefix(self):
        self._build_dfa()
        dm = self._dm
        accept = np.zeros(dm.n_states, np.uint8)
        for s in self._dfa.accept_states:
            accept[s] = 1
        self._table = tmap_byte_table(precompute_tmap(dm), accept, self._dfa.start, dm.n_states)
        self._representation = "dfa"
        self._scan_backend = "prefix+gpu"
        self._selection_reason = f"GPU prefix compose (N={dm.n_states})"

    def _auto_select(self):
        """DFA size decides: a byte table of at most 256 states (dead state
        included) runs as a DFA, the K-gram kernels when it has at most 16;
        larger DFAs, or DFAs over ``dfa_state_cap``, run as a bit-parallel
        NFA (at most 64 NFA states)."""
        self._build_dfa()
        n_states = self._dfa.n_states
        try:
            table = dfa_byte_table(self._dfa)
        except ValueError:          # more states than a byte table holds
            table = None
        if n_states > self._dfa_state_cap or table is None:
            self._select_nfa(f"DFA has {n_states} states > cap {self._dfa_state_cap}; using NFA")
            return
        self._table = table
        self._representation = "dfa"
        if table.n_states <= 16:
            self._scan_backend = "dfa16+gpu"
            self._kgram_k = 2
            self._selection_reason = (f"DFA has {n_states} states ({table.n_states} with its dead state) "
                                      f"<= 16; auto-selected dfa16+gpu (two bytes per table step)")
        else:
            self._scan_backend = "dfa+gpu"
            self._selection_reason = (f"DFA has {n_states} states ({table.n_states} with its dead state) "
                                      f"> 16; auto-selected dfa+gpu (byte table walk)")

    def _record(self) -> CompiledPattern:
        """The selection's compiled form (what the table cache stores)."""
        info = {}
        if self._dfa is not None:
            info["alphabet_size"] = len(self._dfa.alphabet)
            info["dfa_states"] = self._dfa.n_states
        elif self._nm is not None:
            info["alphabet_size"] = len(self._nm.alphabet)
            info["nfa_states"] = self._nm.n_states_raw
        if self._md is not None:
            info["monoid_size"] = self._md.size
        if self._nm is not None and "nfa_states" not in info:
            info["nfa_states"] = self._nm.n_states_raw
        common = dict(representation=self._representation, scan_backend=self._scan_backend,
                      selection_reason=self._selection_reason, kgram_k=self._kgram_k, info=info)
        if self._representation == "nfa":
            from trex_amd.engine import nfa_byte_map
            nm = self._nm
            return CompiledPattern(succ=nm.successor_masks(), byte_to_sym=nfa_byte_map(nm),
                                   start_mask=nm.start_mask(), accept_mask=nm.accept_bits(), **common)
        return CompiledPattern(table=self._table, **common)

    # -- public API -------------------------------------------------------------
    @property
    def compiled(self) -> CompiledPattern:
        return self._compiled

    @property
    def config_info(self) -> dict:
        cp = self._compiled
        info = {"representation": cp.representation, "scan_backend": cp.scan_backend,
                "kgram_k": cp.kgram_k, "selection_reason": cp.selection_reason}
        info.update(cp.info)
        return info

    def _gpu(self):
        if self._engine is None:
            self._engine = self._compiled.engine()
        return self._engine

    def match_arrays(self, data: np.ndarray, offsets: np.ndarray):
        """CSR batch (uint8 bytes, int64 offsets) -> (final state, accept) arrays.
        Final states are DFA table states, or 64-bit state masks for NFAs."""
        final, acc, _ = self._match(data, offsets)
        return final, acc

    def _match(self, data, offsets):
        """(final, accept, kernel ms). Batches with a string over
        ``LONG_STRING`` bytes are cut into chunks matched in parallel: DFAs of
        at most 16 states by the look-back scan, NFAs (and wider DFAs in their
        NFA-64 form) by speculative chunks and fix-ups; other automata walk
        each string on one lane."""
        from trex_amd.engine import DfaEngine, _as_csr
        data, offsets = _as_csr(data=data, offsets=offsets)
        eng = self._gpu()
        lengths = np.diff(offsets)
        long = len(lengths) > 0 and int(lengths.max()) > LONG_STRING
        if long and isinstance(eng, DfaEngine):
            if self._table.n_states <= 16:
                return self._match_long(eng, data, offsets)
            nfa = self._nfa_form()
            if nfa is not None:
                mask, acc, kms, _ = nfa.match_split(data, offsets)
                return nfa_mask_to_state(mask, self._table.dead), acc, kms
        final, acc, kms, _ = eng.match_arrays(data, offsets)
        return final, acc, kms

    def _nfa_form(self):
        """NfaEngine of the DFA table's NFA-64 form when it can be split, else None (cached)."""
        if self._nfa_engine is None:
            from trex_amd.engine import NfaEngine, dfa_as_nfa
            form = dfa_as_nfa(self._table)
            eng = NfaEngine(*form) if form is not None else None
            self._nfa_engine = eng if eng is not None and eng.can_split() else False
        return self._nfa_engine or None

    def _match_long(self, eng, data, offsets):
        from trex_amd.device import DeviceBuffer, Event, Stream
        from trex_amd.engine import ChunkPlan, DeviceBatch, lookback_chunk_len, lookback_target_chunks
        if self._lookback_target is None:       # a property of the table: computed once
            self._lookback_target = lookback_target_chunks(self._table)
        stream = Stream()
        batch = DeviceBatch(data, offsets, layout="csr")
        total = int(offsets[-1] - offsets[0]) if len(offsets) else 0
        plan = ChunkPlan.for_batch(batch, lookback_chunk_len(total, self._lookback_target))
        n = batch.n_strings
        d_f, d_a = DeviceBuffer(n), DeviceBuffer(n)
        e0, e1 = Event(), Event()
        e0.record(stream)
        eng.match_plan(plan, batch.d_bytes.ptr, d_f.ptr, d_a.ptr, stream)
        e1.record(stream)
        if plan.error(stream):
            raise RuntimeError("look-back scan reported a wait timeout")
        return d_f.download(np.uint8, n, stream), d_a.download(np.uint8, n, stream), e0.elapsed_ms(e1)

    def match_batch(self, strings: Sequence) -> list:
        """Match a list of strings (``str`` as latin-1, or bytes); returns list[bool]."""
        if not len(strings):
            return []
        from trex_amd.engine import pack_strings
        _, acc = self.match_arrays(*pack_strings(strings))
        return [bool(x) for x in acc]

    def match(self, strings: Sequence) -> list:
        return self.match_batch(strings)

    def match_batch_timed(self, strings: Sequence) -> tuple:
        """(results, {'kernel_ms', 'total_ms', ...}) like the reference's GPU configs."""
        from trex_amd.engine import pack_strings
        if not len(strings):
            return [], {"kernel_ms": 0.0, "total_ms": 0.0, "n_strings": 0}
        data, offsets = pack_strings(strings)
        t0 = time.perf_counter()
        _, acc, kms = self._match(data, offsets)
        tms = (time.perf_counter() - t0) * 1e3
        return [bool(x) for x in acc], {"kernel_ms": kms, "total_ms": tms,
                                        "total_seconds": time.perf_counter() - t0, "n_strings": len(strings)}
