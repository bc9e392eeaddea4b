"""Compiled-table cache and on-disk table format.

Compiling a pattern (regex -> NFA -> DFA -> minimised byte table, or the
monoid / tmap forms, ``src/optimized_engine.py:113-196``) is host work done
once per pattern; a server holding thousands of resident patterns repeats it
on every start. This module stores the compiled result so a restart loads
tables instead of compiling them:

* :class:`CompiledPattern` -- what :class:`trex_amd.optimized_engine.OptimizedEngine`
  selects for a pattern: the DFA byte table (DFA, monoid or tmap form) or the
  NFA successor masks, plus the selection fields ``config_info`` reports;
* :class:`ProbModel` -- a probabilistic NFA-64 (``T``, start, accept, mode);
* :func:`save_tables` / :func:`load_tables` -- a table file holding any number
  of these records (a serving set is one file);
* :class:`TableCache` -- a directory of table files keyed by a hash of the
  pattern text and every option that changes the compiled table.

File format (version 1, extension ``.trxt``), little-endian:

======  ==================================================================
0       magic ``TREXTBL\\0`` (8 bytes)
8       uint32 format version, uint32 header length ``H``
16      UTF-8 JSON header: ``{"meta": {"records": [...]}, "arrays": {...}}``
        -- one object per record (kind, scalar fields, first row in its
        kind's arrays), and per array its dtype, shape and byte offset
16+H    the arrays, each at a 64-byte aligned offset:
        ``dfa.delta`` uint8 [rows, 256], ``dfa.accept`` uint8 [rows],
        ``nfa.succ`` uint64 [rows, 64], ``nfa.byte_to_sym`` int16 [n, 256],
        ``pnfa.T`` float32 [rows, 64, 64], ``pnfa.start`` / ``pnfa.accept``
        float32 [n, 64]
======  ==================================================================

Records of a kind share one array per field (rows in record order), so a
file holding thousands of patterns is read with one ``read`` and seven
``frombuffer`` views; nothing in it is executed or unpickled, and only the
four dtypes above are accepted.

Every array is checked against its record's fields on load (dtype, shape,
state and symbol ranges, finiteness); a file that fails raises
:class:`TableFileError`, and the cache treats it as a miss and compiles again.
"""

from __future__ import annotations

import hashlib
import json
import os
import threading
from dataclasses import dataclass, field
from pathlib import Path
from typing import Iterable, Optional

import numpy as np

from trex_amd.tables import DfaByteTable

FORMAT_VERSION = 1
ENV_DIR = "TREX_AMD_TABLE_CACHE"


class TableFileError(ValueError):
    """A table file that is unreadable, of another format version or
    inconsistent with its own metadata."""


@dataclass
class CompiledPattern:
    """The compiled form of one pattern under one configuration."""
    representation: str                 # "dfa" (byte table walk) or "nfa" (bit-parallel)
    scan_backend: str
    selection_reason: str
    kgram_k: Optional[int] = None
    info: dict = field(default_factory=dict)   # alphabet_size, dfa_states / nfa_states, monoid_size
    table: Optional[DfaByteTable] = None       # representation "dfa"
    succ: Optional[np.ndarray] = None          # representation "nfa": uint64 [n_symbols, 64]
    byte_to_sym: Optional[np.ndarray] = None   # int16 [256]
    start_mask: int = 0
    accept_mask: int = 0
    key: Optional[dict] = None                 # the cache key fields it was stored under

    def __post_init__(self):
        if self.representation == "dfa":
            if self.table is None:
                raise ValueError("a DFA pattern needs its byte table")
        elif self.representation == "nfa":
            self.succ = np.ascontiguousarray(self.succ, dtype=np.uint64)
            self.byte_to_sym = np.ascontiguousarray(self.byte_to_sym, dtype=np.int16)
            _check_nfa(self.succ, self.byte_to_sym, self.start_mask, self.accept_mask)
        else:
            raise ValueError(f"unknown representation {self.representation!r}")

    def engine(self):
        """A resident GPU engine for this pattern (needs the device)."""
        from trex_amd.engine import DfaEngine, NfaEngine
        if self.representation == "nfa":
            return NfaEngine(self.succ, self.byte_to_sym, self.start_mask, self.accept_mask)
        eng = DfaEngine(self.table)
        if self.scan_backend == "sequential":
            eng.set_variant("k1_generic" if self.table.n_states > 16 else "k1_16")
        return eng

    def same_tables(self, other: "CompiledPattern") -> bool:
        if self.representation != other.representation:
            return False
        if self.representation == "dfa":
            a, b = self.table, other.table
            return (a.n_states == b.n_states and a.start == b.start and a.dead == b.dead
                    and np.array_equal(a.delta, b.delta) and np.array_equal(a.accept, b.accept))
        return (np.array_equal(self.succ, other.succ) and np.array_equal(self.byte_to_sym, other.byte_to_sym)
                and self.start_mask == other.start_mask and self.accept_mask == other.accept_mask)


@dataclass
class ProbModel:
    """A probabilistic NFA-64: ``T[c]`` (row = destination, column = source),
    start and accept weights, and the engine's arithmetic mode (None = the
    engine's default)."""
    T: np.ndarray
    start: np.ndarray
    accept: np.ndarray
    mode: Optional[str] = None

    def __post_init__(self):
        self.T = np.ascontiguousarray(self.T, dtype=np.float32)
        self.start = np.ascontiguousarray(self.start, dtype=np.float32)
        self.accept = np.ascontiguousarray(self.accept, dtype=np.float32)
        if self.T.ndim != 3 or self.T.shape[1:] != (64, 64) or self.T.shape[0] < 1:
            raise ValueError(f"T must have shape [n_symbols, 64, 64], got {self.T.shape}")
        if self.start.shape != (64,) or self.accept.shape != (64,):
            raise ValueError("start and accept must have 64 entries")
        for name, a in (("T", self.T), ("start", self.start), ("accept", self.accept)):
            if not np.isfinite(a).all():
                raise ValueError(f"{name} has non-finite entries")

    def engine(self):
        from trex_amd.engine import ProbNfaEngine
        return ProbNfaEngine(self.T, self.start, self.accept, mode=self.mode)

    def digest(self) -> str:
        h = hashlib.sha256()
        for a in (self.T, self.start, self.accept):
            h.update(str(a.shape).encode())
            h.update(a.tobytes())
        return h.hexdigest()


def _check_nfa(succ, b2s, start_mask, accept_mask):
    if succ.ndim != 2 or succ.shape[1] != 64:
        raise ValueError(f"succ must have shape [n_symbols, 64], got {succ.shape}")
    if b2s.shape != (256,):
        raise ValueError("byte_to_sym must have 256 entries")
    if len(b2s) and (int(b2s.min()) < -2 or int(b2s.max()) >= succ.shape[0]):
        raise ValueError("byte_to_sym names a symbol outside -2..n_symbols-1")
    for name, m in (("start", start_mask), ("accept", accept_mask)):
        if not 0 <= int(m) < 1 << 64:
            raise ValueError(f"{name} mask outside 64 bits")


# -- records <-> engines -------------------------------------------------------
def record_of(engine):
    """Host record of a resident engine (DfaEngine, NfaEngine or ProbNfaEngine)."""
    from trex_amd.engine import DfaEngine, NfaEngine, ProbNfaEngine
    if isinstance(engine, DfaEngine):
        t = engine.table
        return CompiledPattern("dfa", "dfa16+gpu" if t.n_states <= 16 else "dfa+gpu", "resident engine",
                               info={"table_states": t.n_states}, table=t)
    if isinstance(engine, NfaEngine):
        return CompiledPattern("nfa", "nfa", "resident engine", succ=engine.succ,
                               byte_to_sym=engine.byte_to_sym, start_mask=engine.start, accept_mask=engine.accept,
                               info={"alphabet_size": int(engine.succ.shape[0])})
    if isinstance(engine, ProbNfaEngine):
        return ProbModel(engine.T, engine.start, engine.accept, mode=engine.info()["mode"])
    raise TypeError(f"no table record for {type(engine).__name__}")


def engines_of(records: Iterable) -> list:
    """Resident engines for a list of records (needs the device)."""
    return [r.engine() for r in records]


# -- file format -----------------------------------------------------------------
MAGIC = b"TREXTBL\0"
_ALIGN = 64
_DTYPES = {"|u1": np.uint8, "<i2": np.int16, "<u8": np.uint64, "<f4": np.float32}


def _pack(meta: dict, arrays: dict) -> bytes:
    """Container: magic, uint32 version, uint32 header length, JSON header
    (the records' fields and every array's dtype, shape and byte offset),
    then the arrays, each at a 64-byte aligned offset."""
    layout, chunks, pos = {}, [], 0
    for name, a in arrays.items():
        a = np.ascontiguousarray(a)
        dt = a.dtype.newbyteorder("<") if a.dtype.byteorder not in ("|", "<", "=") else a.dtype
        key = np.dtype(dt).str
        if key not in _DTYPES:
            raise TypeError(f"{name}: dtype {a.dtype} is not stored")
        raw = a.astype(dt, copy=False).tobytes()
        layout[name] = {"dtype": key, "shape": list(a.shape), "offset": pos}
        chunks.append(raw + b"\0" * (-len(raw) % _ALIGN))
        pos += len(chunks[-1])
    base = 0   # first array offset; header length and offsets settle in a few passes
    while True:
        absolute = {k: dict(v, offset=v["offset"] + base) for k, v in layout.items()}
        header = json.dumps({"meta": meta, "arrays": absolute}).encode()
        need = 16 + len(header)
        need += -need % _ALIGN
        if need <= base:
            header += b" " * (base - 16 - len(header))
            break
        base = need
    return MAGIC + np.array([FORMAT_VERSION, len(header)], "<u4").tobytes() + header + b"".join(chunks)


def _unpack(buf: bytes) -> tuple:
    if len(buf) < 16 or buf[:8] != MAGIC:
        raise TableFileError("not a table file")
    version, head_len = (int(x) for x in np.frombuffer(buf, "<u4", 2, 8))
    if version != FORMAT_VERSION:
        raise TableFileError(f"format version {version}, this build reads {FORMAT_VERSION}")
    if 16 + head_len > len(buf):
        raise TableFileError("truncated header")
    try:
        head = json.loads(buf[16:16 + head_len].decode())
        meta, layout = head["meta"], head["arrays"]
        arrays = {}
        for name, v in layout.items():
            if v["dtype"] not in _DTYPES:
                raise TableFileError(f"{name}: dtype {v['dtype']!r} is not a table dtype")
            dt = np.dtype(v["dtype"])
            shape = tuple(int(x) for x in v["shape"])
            if any(x < 0 for x in shape):
                raise TableFileError(f"{name}: negative extent")
            count = int(np.prod(shape, dtype=np.int64))
            off = int(v["offset"])
            if off < 16 + head_len or off % _ALIGN or off + count * dt.itemsize > len(buf):
                raise TableFileError(f"{name}: extent outside the file")
            arrays[name] = np.frombuffer(buf, dt, count, off).reshape(shape)
    except TableFileError:
        raise
    except (ValueError, KeyError, TypeError, UnicodeDecodeError) as e:
        raise TableFileError(f"header is not valid: {e}") from None
    return meta, arrays


def _encode(records: list) -> tuple:
    """Records -> (meta, arrays). Records of a kind share one array per field
    (rows concatenated in record order), so a file of thousands of patterns
    is a handful of arrays."""
    cols = {k: [] for k in ("dfa.delta", "dfa.accept", "nfa.succ", "nfa.byte_to_sym",
                            "pnfa.T", "pnfa.start", "pnfa.accept")}
    rows = {"dfa": 0, "nfa": 0, "pnfa": 0}
    metas = []
    for r in records:
        if isinstance(r, CompiledPattern):
            m = {"kind": r.representation, "scan_backend": r.scan_backend, "selection_reason": r.selection_reason,
                 "kgram_k": r.kgram_k, "info": r.info, "key": r.key}
            if r.representation == "dfa":
                t = r.table
                m.update(row=rows["dfa"], n_states=t.n_states, start=t.start, dead=t.dead)
                cols["dfa.delta"].append(t.delta)
                cols["dfa.accept"].append(t.accept)
                rows["dfa"] += t.n_states
            else:
                m.update(row=rows["nfa"], index=len(cols["nfa.byte_to_sym"]), n_symbols=int(r.succ.shape[0]),
                         start_mask=int(r.start_mask), accept_mask=int(r.accept_mask))
                cols["nfa.succ"].append(r.succ)
                cols["nfa.byte_to_sym"].append(r.byte_to_sym[None])
                rows["nfa"] += int(r.succ.shape[0])
        elif isinstance(r, ProbModel):
            m = {"kind": "pnfa", "row": rows["pnfa"], "index": len(cols["pnfa.start"]),
                 "n_symbols": int(r.T.shape[0]), "mode": r.mode, "digest": r.digest()}
            cols["pnfa.T"].append(r.T)
            cols["pnfa.start"].append(r.start[None])
            cols["pnfa.accept"].append(r.accept[None])
            rows["pnfa"] += int(r.T.shape[0])
        else:
            raise TypeError(f"cannot store {type(r).__name__}")
        metas.append(m)
    empty = {"dfa.delta": (np.uint8, (0, 256)), "dfa.accept": (np.uint8, (0,)),
             "nfa.succ": (np.uint64, (0, 64)), "nfa.byte_to_sym": (np.int16, (0, 256)),
             "pnfa.T": (np.float32, (0, 64, 64)), "pnfa.start": (np.float32, (0, 64)),
             "pnfa.accept": (np.float32, (0, 64))}
    arrays = {k: (np.concatenate(v) if v else np.zeros(empty[k][1], empty[k][0])) for k, v in cols.items()}
    return {"records": metas}, arrays


def _column(arrays, name, dtype, tail):
    if name not in arrays:
        raise TableFileError(f"missing array {name}")
    a = arrays[name]
    if a.dtype != np.dtype(dtype) or a.shape[1:] != tuple(tail):
        raise TableFileError(f"{name}: expected {np.dtype(dtype)} [*, {tail}], got {a.dtype} {a.shape}")
    return a


def _rows(a, name, lo, n):
    if not 0 <= lo <= lo + n <= len(a):
        raise TableFileError(f"{name}: rows {lo}..{lo + n} outside 0..{len(a)}")
    return a[lo:lo + n]


def _decode(meta: dict, arrays: dict) -> list:
    delta = _column(arrays, "dfa.delta", np.uint8, (256,))
    dacc = _column(arrays, "dfa.accept", np.uint8, ())
    succ = _column(arrays, "nfa.succ", np.uint64, (64,))
    b2s = _column(arrays, "nfa.byte_to_sym", np.int16, (256,))
    T = _column(arrays, "pnfa.T", np.float32, (64, 64))
    pst = _column(arrays, "pnfa.start", np.float32, (64,))
    pacc = _column(arrays, "pnfa.accept", np.float32, (64,))
    if len(delta) != len(dacc):
        raise TableFileError("dfa.delta and dfa.accept differ in rows")
    out = []
    try:
        metas = meta["records"]
        for i, m in enumerate(metas):
            kind = m["kind"]
            if kind == "dfa":
                n = int(m["n_states"])
                if not 1 <= n <= 256:
                    raise TableFileError(f"record {i}: n_states={n}")
                row = int(m["row"])
                t = DfaByteTable(_rows(delta, "dfa.delta", row, n), _rows(dacc, "dfa.accept", row, n),
                                 int(m["start"]), int(m["dead"]), n)
                if not -1 <= t.dead < n:
                    raise TableFileError(f"record {i}: dead state {t.dead} outside the table")
                out.append(CompiledPattern("dfa", m["scan_backend"], m["selection_reason"], m["kgram_k"],
                                           dict(m["info"]), table=t, key=m.get("key")))
            elif kind == "nfa":
                s = int(m["n_symbols"])
                out.append(CompiledPattern("nfa", m["scan_backend"], m["selection_reason"], m["kgram_k"],
                                           dict(m["info"]), succ=_rows(succ, "nfa.succ", int(m["row"]), s),
                                           byte_to_sym=_rows(b2s, "nfa.byte_to_sym", int(m["index"]), 1)[0],
                                           start_mask=int(m["start_mask"]), accept_mask=int(m["accept_mask"]),
                                           key=m.get("key")))
            elif kind == "pnfa":
                s, k = int(m["n_symbols"]), int(m["index"])
                r = ProbModel(_rows(T, "pnfa.T", int(m["row"]), s), _rows(pst, "pnfa.start", k, 1)[0],
                              _rows(pacc, "pnfa.accept", k, 1)[0], mode=m["mode"])
                if r.digest() != m["digest"]:
                    raise TableFileError(f"record {i}: probabilistic tables do not match their digest")
                out.append(r)
            else:
                raise TableFileError(f"record {i}: unknown kind {kind!r}")
    except TableFileError:
        raise
    except (KeyError, TypeError, ValueError) as e:
        raise TableFileError(f"record {len(out)}: {e}") from None
    return out


def save_tables(path, records: Iterable) -> None:
    """Write records (CompiledPattern / ProbModel, or resident engines) to one
    table file, atomically (a reader sees the old file or the new one)."""
    recs = [r if isinstance(r, (CompiledPattern, ProbModel)) else record_of(r) for r in records]
    buf = _pack(*_encode(recs))
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    tmp = path.with_name(f".{path.name}.{os.getpid()}.{threading.get_ident()}.tmp")
    try:
        with open(tmp, "wb") as f:
            f.write(buf)
        os.replace(tmp, path)
    finally:
        if tmp.exists():
            tmp.unlink()


def load_tables(path) -> list:
    """Records of a table file; raises :class:`TableFileError` on a bad file.
    The arrays of the records are read-only views of the file's bytes."""
    try:
        buf = Path(path).read_bytes()
    except OSError as e:
        raise TableFileError(f"{path}: {e}") from None
    return _decode(*_unpack(buf))


# -- the cache ----------------------------------------------------------------------
def pattern_key(regex: str, config: Optional[str], dfa_state_cap: int, monoid_cap: int) -> dict:
    """Every input that changes what OptimizedEngine compiles for a pattern."""
    return {"format": FORMAT_VERSION, "regex": regex, "config": config,
            "dfa_state_cap": int(dfa_state_cap), "monoid_cap": int(monoid_cap)}


def _digest(key: dict) -> str:
    return hashlib.sha256(json.dumps(key, sort_keys=True).encode()).hexdigest()[:40]


class TableCache:
    """Compiled patterns keyed by pattern hash: an in-process map in front of
    a directory of table files (``<hash>.trxt``, one record each).

    A file whose stored key differs from the requested one (a hash
    collision), or that fails to load, is a miss; :meth:`get` never raises
    for a bad file. Writers replace files atomically, so processes may share
    a directory."""

    def __init__(self, root=None, memory: bool = True):
        self.root = Path(root) if root is not None else None
        self._mem: dict = {} if memory else None
        self._lock = threading.Lock()
        self.hits = self.misses = self.bad_files = 0

    def path_of(self, key: dict) -> Optional[Path]:
        return self.root / f"{_digest(key)}.trxt" if self.root is not None else None

    def get(self, key: dict) -> Optional[CompiledPattern]:
        d = _digest(key)
        with self._lock:
            if self._mem is not None and d in self._mem and self._mem[d].key == key:
                self.hits += 1
                return self._mem[d]
        p = self.path_of(key)
        if p is not None and p.exists():
            try:
                recs = load_tables(p)
            except TableFileError:
                recs = None
                with self._lock:
                    self.bad_files += 1
            if recs and len(recs) == 1 and isinstance(recs[0], CompiledPattern) and recs[0].key == key:
                with self._lock:
                    self.hits += 1
                    if self._mem is not None:
                        self._mem[d] = recs[0]
                return recs[0]
        with self._lock:
            self.misses += 1
        return None

    def put(self, key: dict, cp: CompiledPattern) -> None:
        cp.key = key
        d = _digest(key)
        with self._lock:
            if self._mem is not None:
                self._mem[d] = cp
        p = self.path_of(key)
        if p is not None:
            try:
                save_tables(p, [cp])
            except OSError:
                pass   # a read-only directory only loses the on-disk copy

    def clear_memory(self) -> None:
        with self._lock:
            if self._mem is not None:
                self._mem.clear()


_default: Optional[TableCache] = None
_default_root: Optional[str] = None


def default_cache() -> Optional[TableCache]:
    """The cache ``OptimizedEngine`` uses when given none: a directory cache at
    ``$TREX_AMD_TABLE_CACHE`` when that is set, else no cache."""
    global _default, _default_root
    root = os.environ.get(ENV_DIR) or None
    if root is None:
        return None
    if _default is None or _default_root != root:
        _default, _default_root = TableCache(root), root
    return _default
