"""Autotuner: picks launch parameters by timing candidates on the device and
caches the choice.

* :func:`tune_mix` -- split lengths, chunk lengths, checkpoint spacing and
  NFA table form (which sets the workgroup shape) of a mixed DFA/NFA work
  queue (:class:`trex_amd.mix.MixParams`), timed on a seeded sample of the
  batch;
* :func:`tune_dfa` -- kernel variant and loads in flight of a DFA-16 engine;
* :func:`tune_pnfa` -- waves per workgroup of a probabilistic NFA engine.

Choices are cached in a JSON file (``$TREX_AMD_TUNE_CACHE``, default
``~/.cache/trex_amd/tune.json``) under a key made of the device, the kind of
work and a coarse summary of the workload (pattern kinds, log2 length
histogram), so a repeated workload shape is not tuned again.
"""

from __future__ import annotations

import hashlib
import itertools
import json
import os
from pathlib import Path

import numpy as np

from trex_amd import _lib
from trex_amd.device import DeviceBuffer, Event, Stream
from trex_amd.mix import MixEngine, MixParams

MIX_GRID = {
    "dfa_split": (64 << 10, 256 << 10, 1 << 20),
    "dfa_chunk": (16 << 10, 64 << 10),
    "nfa_split": (8 << 10, 32 << 10),
    "nfa_chunk": (8 << 10, 32 << 10),
    "ck": (64, 256),
    "form": ("nibble", "slice"),   # NFA tables and workgroup shape: 512 or 1024 threads
}
NEAR = 0.02   # relative timing difference on the sample treated as noise


def cache_path() -> Path:
    return Path(os.environ.get("TREX_AMD_TUNE_CACHE", Path.home() / ".cache" / "trex_amd" / "tune.json"))


def _load() -> dict:
    try:
        return json.loads(cache_path().read_text())
    except (OSError, ValueError):
        return {}


def _store(key: str, value: dict) -> None:
    data = _load()
    data[key] = value
    p = cache_path()
    try:
        p.parent.mkdir(parents=True, exist_ok=True)
        tmp = p.with_suffix(".tmp")
        tmp.write_text(json.dumps(data, indent=1, sort_keys=True))
        tmp.replace(p)
    except OSError:
        pass   # a read-only home only loses the cache


def _key(kind: str, *parts) -> str:
    h = hashlib.sha1(json.dumps([kind, *parts], sort_keys=True, default=str).encode()).hexdigest()[:16]
    return f"{kind}:{h}"


def _time(fn, stream: Stream, reps: int) -> float:
    fn()
    stream.synchronize()
    a, b = Event(), Event()
    a.record(stream)
    for _ in range(reps):
        fn()
    b.record(stream)
    stream.synchronize()
    return a.elapsed_ms(b) / reps


def _length_summary(lengths: np.ndarray) -> list:
    bins = np.bincount(np.floor(np.log2(np.maximum(lengths, 1))).astype(np.int64), minlength=25)
    return (bins * 64 // max(1, len(lengths))).tolist()   # coarse histogram (1/64 resolution)


def tune_mix(engines, begin, end, pattern_of, d_bytes: int, stream: Stream, sample: int = 4096,
             grid: dict | None = None, seed: int = 0, use_cache: bool = True, log=None):
    """Best :class:`MixParams` for this batch, and a dict describing the choice."""
    begin = np.asarray(begin, np.int64)
    end = np.asarray(end, np.int64)
    pattern_of = np.asarray(pattern_of, np.int32)
    mix = MixEngine(engines)
    kinds = mix.kinds.tolist()
    lengths = end - begin
    key = _key("mix", sorted(MIX_GRID.items()), _lib.device_name(), kinds, _length_summary(lengths),
               int(np.log2(max(len(begin), 1))))
    if use_cache:
        hit = _load().get(key)
        if hit:
            return MixParams(**hit["params"]), {"source": "cache", "key": key, "ms": hit.get("ms")}
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(len(begin), size=min(sample, len(begin)), replace=False))
    sb, se, sp = begin[idx], end[idx], pattern_of[idx]
    n = len(idx)
    d_mask, d_acc = DeviceBuffer(8 * max(n, 1)), DeviceBuffer(max(n, 1))
    grid = grid or MIX_GRID
    names = list(grid)
    results = []
    for values in itertools.product(*(grid[k] for k in names)):
        params = MixParams(**dict(zip(names, values)))
        if params.nfa_chunk % params.ck or params.dfa_split < params.dfa_chunk or params.nfa_split < params.nfa_chunk:
            continue
        try:
            mix.plan(sb, se, sp, params)
        except _lib.TrexError:   # a form some pattern cannot take
            continue
        ms = _time(lambda: mix.run(d_bytes, d_mask.ptr, d_acc.ptr, stream), stream, 2)
        results.append((ms, params))
        if log:
            log(f"tune_mix: {params} {ms:.3f} ms")
    d_mask.free()
    d_acc.free()
    mix.close()
    # Parameters whose effect on the sample is within its timing noise are
    # left at their defaults: among the candidates within NEAR of the fastest,
    # the one with the fewest fields off the defaults wins (then the fastest).
    fastest = min(r[0] for r in results)
    default = MixParams()
    off = lambda p: sum(getattr(p, f) != getattr(default, f) for f in ("dfa_split", "dfa_chunk", "nfa_split",
                                                                        "nfa_chunk", "ck"))
    ms, best = min((r for r in results if r[0] <= fastest * (1 + NEAR)), key=lambda r: (off(r[1]), r[0]))
    _store(key, {"params": best.__dict__, "ms": ms, "sample_strings": n, "candidates": len(results)})
    return best, {"source": "tuned", "key": key, "ms": ms, "candidates": len(results), "sample_strings": n}


def tune_dfa(engine, batch, d_final: int, d_accept: int, stream: Stream, use_cache: bool = True) -> dict:
    """Kernel variant and prefetch depth of one DFA engine on a device batch."""
    mean_len = batch.payload_bytes // max(1, batch.n_strings)
    key = _key("dfa", _lib.device_name(), engine.table.n_states, engine.info()["window_base"] >= 0,
               int(np.log2(max(batch.n_strings, 1))), int(np.log2(max(mean_len, 1))))
    if use_cache:
        hit = _load().get(key)
        if hit:
            engine.set_variant(hit["variant"])
            engine.set_prefetch(hit["prefetch"])
            return {"source": "cache", **hit}
    options = []
    for variant in ("k2_window_f", "k2_window", "k1_16", "k1_generic"):
        for pf in (8, 16):
            try:
                engine.set_variant(variant)
                engine.set_prefetch(pf)
            except _lib.TrexError:
                continue
            ms = _time(lambda: engine.match_batch_device(batch, d_final, d_accept, stream), stream, 3)
            options.append((ms, variant, pf))
    ms, variant, pf = min(options)
    engine.set_variant(variant)
    engine.set_prefetch(pf)
    choice = {"variant": variant, "prefetch": pf, "ms": ms}
    _store(key, choice)
    return {"source": "tuned", **choice}


def tune_pnfa(engine, d_symbols: int, n_seq: int, length: int, d_ll: int, d_lm: int, stream: Stream,
              use_cache: bool = True) -> dict:
    """Waves per workgroup of a probabilistic NFA engine (its mode is kept)."""
    mode = engine.info()["mode"]
    key = _key("pnfa", _lib.device_name(), mode, engine.n_symbols, int(np.log2(max(n_seq, 1))),
               int(np.log2(max(length, 1))))
    if use_cache:
        hit = _load().get(key)
        if hit:
            engine.set_waves(hit["waves"])
            return {"source": "cache", **hit}
    options = []
    for waves in (4, 8):
        engine.set_waves(waves)
        options.append((_time(lambda: engine.loglik_device(d_symbols, n_seq, length, d_ll, d_lm, stream), stream, 2),
                        waves))
    ms, waves = min(options)
    engine.set_waves(waves)
    choice = {"waves": waves, "ms": ms}
    _store(key, choice)
    return {"source": "tuned", **choice}
