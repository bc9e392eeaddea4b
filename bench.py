"""Benchmark: DFA-16 K-gram-fused batch over 1 GiB on one MI355X.

Workload (one "step"): 8 seeded random regexes over the 32-symbol alphabet
0x40-0x5F, each compiled to a <= 16-state DFA (bytes outside the alphabet go
to the dead state); 262,144 strings x 4,096 bytes = 1 GiB of i.i.d. uniform
symbols, resident in HBM before timing; every pattern walks every string
(8 kernel launches per step); outputs are the final state and accept bit per
string. Gchars/s = 8 x 2^30 / step time.

The batch is held in the engine's staggered HBM layout (groups of 64 strings,
each group 256 bytes further than the last; --layout csr keeps the packed
CSR bytes). The dominant kernel's roofline figure uses its algorithmic bytes
per launch (input bytes + per-string begin/end + final-state and accept
bytes) over its average launch duration, measured with HIP events on the
launch stream. The results of the last step are checked bit-exact against the
multithreaded C oracle over the whole batch; the oracle's time on that check
is the CPU baseline.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
N_STRINGS = 262_144
LENGTH = 4096
N_PATTERNS = 8
SEED = 2
KERNEL_NAMES = {("k2_window", 16): "dfa16_k2w_pf16", ("k2_window", 8): "dfa16_k2w_pf8",
                ("k2_window_f", 16): "dfa16_k2f_pf16", ("k2_window_f", 8): "dfa16_k2f_pf8",
                ("k1_16", 16): "dfa16_k1", ("k1_16", 8): "dfa16_k1"}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def time_launches(launch, stream, reps: int) -> float:
    """Average ms of `launch()` over `reps` back-to-back launches (after one warm-up)."""
    from trex_amd.device import Event
    launch()
    a, b = Event(), Event()
    a.record(stream)
    for _ in range(reps):
        launch()
    b.record(stream)
    stream.synchronize()
    return a.elapsed_ms(b) / reps


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-verify", action="store_true")
    ap.add_argument("--variant", default="auto")
    ap.add_argument("--prefetch", type=int, default=8, help="16-byte vectors in flight per lane (8 or 16)")
    ap.add_argument("--layout", default="staggered", choices=["staggered", "csr"])
    ap.add_argument("--sweep", action="store_true", help="time every kernel variant and layout, then exit")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also time a plain 1 GiB device copy (HBM traffic counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}), flush=True)
        return 2

    from trex_amd import _lib
    from trex_amd.device import DeviceBuffer, Event, Stream
    from trex_amd.engine import DeviceBatch, DfaEngine
    from trex_amd.tables import dfa_byte_table
    from trex_amd.workloads import ALPHABET_32, random_dfa_pattern, uniform_batch

    _lib.check(_lib.lib().trex_device_check(), "trex_device_check")
    rng = np.random.default_rng(SEED)
    patterns = [random_dfa_pattern(rng, ALPHABET_32) for _ in range(N_PATTERNS)]
    tables = [dfa_byte_table(d) for _, d in patterns]
    log(f"patterns: {[p for p, _ in patterns]}; states {[t.n_states for t in tables]}")
    t0 = time.perf_counter()
    data, offsets = uniform_batch(rng, N_STRINGS, LENGTH, 0x40, 0x60)
    log(f"generated {data.nbytes / 2**30:.2f} GiB in {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    batch = DeviceBatch(data, offsets, layout=args.layout)
    log(f"uploaded ({args.layout} layout, {batch.d_bytes.nbytes} device bytes) in {time.perf_counter() - t0:.1f}s")
    n = N_STRINGS
    outs = [(DeviceBuffer(n), DeviceBuffer(n)) for _ in range(N_PATTERNS)]
    engines = [DfaEngine(t) for t in tables]
    for e in engines:
        if args.variant != "auto":
            e.set_variant(args.variant)
        e.set_prefetch(args.prefetch)
    variants = [e.info()["variant"] for e in engines]
    log(f"variants: {variants}, prefetch {args.prefetch}")
    stream = Stream()

    copy_gbs = None
    if args.copy_probe:
        d_copy = DeviceBuffer(data.nbytes)
        ms = time_launches(lambda: _lib.call("trex_copy", d_copy.ptr, batch.d_bytes.ptr, data.nbytes, stream.handle),
                           stream, 5)
        copy_gbs = 2 * data.nbytes / (ms * 1e-3) / 1e9
        log(f"copy probe: {data.nbytes / 2**30:.2f} GiB copy {ms:.4f} ms = {copy_gbs:.0f} GB/s (read+write)")
        d_copy.free()

    if args.sweep:
        return sweep(args, data, offsets, batch, engines, outs, stream)

    evs = [[(Event(), Event()) for _ in range(N_PATTERNS)] for _ in range(args.steps)]
    step_ev = (Event(), Event())

    def step(k: int | None):
        for i, eng in enumerate(engines):
            if k is not None:
                evs[k][i][0].record(stream)
            eng.match_batch_device(batch, outs[i][0].ptr, outs[i][1].ptr, stream)
            if k is not None:
                evs[k][i][1].record(stream)

    for _ in range(args.warmup):
        step(None)
    stream.synchronize()
    step_ev[0].record(stream)
    wall0 = time.perf_counter()
    for k in range(args.steps):
        step(k)
    step_ev[1].record(stream)
    stream.synchronize()
    wall = time.perf_counter() - wall0
    kernel_ms = [a.elapsed_ms(b) for row in evs for a, b in row]
    total_ms = step_ev[0].elapsed_ms(step_ev[1])
    ms_per_step = total_ms / args.steps
    chars = float(N_PATTERNS) * data.nbytes
    gchars = chars / (ms_per_step * 1e-3) / 1e9
    avg_kernel_ms = float(np.mean(kernel_ms))
    alg_bytes = data.nbytes + 16 * n + 2 * n
    achieved = alg_bytes / (avg_kernel_ms * 1e-3) / 1e9
    log(f"step {ms_per_step:.3f} ms (wall {wall / args.steps * 1e3:.3f} ms), kernel avg {avg_kernel_ms:.4f} ms "
        f"min {min(kernel_ms):.4f} max {max(kernel_ms):.4f}, {gchars:.1f} Gchars/s, {achieved:.0f} GB/s")

    cpu = None
    verified = None
    if not args.no_verify:
        from oracle import cpu as ocpu
        from oracle.dfa import dfa_simulate
        threads = ocpu.default_threads()
        t_cpu = 0.0
        ok = True
        for i, t in enumerate(tables):
            t0 = time.perf_counter()
            rf, ra = ocpu.dfa_walk(t.delta, t.accept, t.start, data, offsets, threads=threads)
            t_cpu += time.perf_counter() - t0
            gf = outs[i][0].download(np.uint8, n)
            ga = outs[i][1].download(np.uint8, n)
            same = bool(np.array_equal(gf, rf) and np.array_equal(ga, ra))
            ok &= same
            log(f"pattern {i}: bit-exact={same} accept-rate={float(ra.mean()):.4f}")
        verified = ok
        # pure-Python restatement of DFA.simulate on a 256-string sample (1 MiB)
        sample = [bytes(data[offsets[j]:offsets[j + 1]]).decode("latin-1") for j in range(256)]
        t0 = time.perf_counter()
        for s in sample:
            dfa_simulate(patterns[0][1], s)
        t_py = time.perf_counter() - t0
        cpu = {"value": round(chars / t_cpu / 1e9, 3), "unit": "Gchars/s", "cores": threads,
               "kind": "port",
               "sample": f"full workload (8 patterns x 1 GiB) through the C oracle byte walk, {threads} threads, "
                         f"{t_cpu:.2f} s; host {cpu_model()}, os.cpu_count()={os.cpu_count()}",
               "python_dfa_simulate_mchars_s": round(len(sample) * LENGTH / t_py / 1e6, 2)}
        if not ok:
            log("PARITY FAILURE")

    kernel = KERNEL_NAMES.get((variants[0], args.prefetch), variants[0])
    result = {
        "metric": "Gchars/s (DFA-16, NFA-64) + %HBM roofline; p50/p99 latency served vs plain",
        "value": round(gchars, 2),
        "unit": "Gchars/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded PCG64, uniform over 0x40-0x5F)",
        "config": {"workload": "dfa16_k4_fused_1GiB", "patterns": N_PATTERNS, "strings": N_STRINGS,
                   "length": LENGTH, "alphabet": "0x40-0x5F", "kernel_variants": variants,
                   "prefetch": args.prefetch, "layout": args.layout, "launches_per_step": N_PATTERNS},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": kernel, "kernel_avg_ms": round(avg_kernel_ms, 5),
                     "alg_bytes_per_launch": alg_bytes, "copy_probe_gbs": copy_gbs},
        "cpu_baseline": cpu,
        "verified_bit_exact": verified,
    }
    print(json.dumps(result), flush=True)
    return 0 if verified in (None, True) else 1


def sweep(args, data, offsets, batch, engines, outs, stream) -> int:
    """Kernel variants x prefetch x layout, plus the read probes."""
    from trex_amd import _lib
    from trex_amd.device import DeviceBuffer
    from trex_amd.engine import DeviceBatch
    n = len(offsets) - 1
    other = DeviceBatch(data, offsets, layout="csr" if batch.layout == "staggered" else "staggered")
    layouts = {batch.layout: batch, other.layout: other}
    ref = []
    for i, eng in enumerate(engines):
        eng.match_batch_device(batch, outs[i][0].ptr, outs[i][1].ptr, stream)
        ref.append(outs[i][0].download(np.uint8, n, stream))
    sink = DeviceBuffer(max(8192 * 256, n) * 4)
    ms = time_launches(lambda: _lib.call("trex_read_probe", batch.d_bytes.ptr, data.nbytes, None, None, 0, 0,
                                         sink.ptr, stream.handle), stream, 5)
    log(f"read probe coalesced stream      : {ms:.4f} ms = {data.nbytes / (ms * 1e-3) / 1e9:.0f} GB/s")
    for name, b in layouts.items():
        for mode in (1, 2):
            ms = time_launches(lambda: _lib.call("trex_read_probe", b.d_bytes.ptr, 16, b.d_begin.ptr, b.d_end.ptr,
                                                 n, mode, sink.ptr, stream.handle), stream, 5)
            log(f"read probe {name:9s} pf{8 * mode:2d}          : {ms:.4f} ms = "
                f"{data.nbytes / (ms * 1e-3) / 1e9:.0f} GB/s")
    for name, b in layouts.items():
        for v, pf in [("k1_16", 16), ("k2_window", 8), ("k2_window", 16), ("k2_window_f", 8), ("k2_window_f", 16)]:
            times, same = [], True
            for i, eng in enumerate(engines):
                eng.set_variant(v)
                eng.set_prefetch(pf)
                times.append(time_launches(lambda: eng.match_batch_device(b, outs[i][0].ptr, outs[i][1].ptr, stream),
                                           stream, 3))
                same &= bool(np.array_equal(outs[i][0].download(np.uint8, n, stream), ref[i]))
            ms = float(np.mean(times))
            log(f"sweep {name:9s} {v:12s} pf{pf:2d}: kernel avg {ms:.4f} ms ({min(times):.4f}-{max(times):.4f}) "
                f"= {(data.nbytes + 18 * n) / (ms * 1e-3) / 1e9:.0f} GB/s, same={same}")
    other.free()
    return 0


if __name__ == "__main__":
    sys.exit(main())
