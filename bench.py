"""Benchmark: DFA-16 K-gram-fused batch over 1 GiB on one MI355X, with the
Boolean NFA-64 and the probabilistic NFA-64 as secondary legs.

Headline workload (one "step"): 8 seeded random regexes over the 32-symbol
alphabet 0x40-0x5F, each compiled to a <= 16-state DFA (bytes outside the
alphabet go to the dead state); 262,144 strings x 4,096 bytes = 1 GiB of
i.i.d. uniform symbols, resident in HBM before timing; every pattern walks
every string (8 kernel launches per step); outputs are the final state and
accept bit per string. Gchars/s = 8 x 2^30 / step time.

The batch is held in the engine's staggered HBM layout (groups of 64 strings,
each group 256 bytes further than the last; --layout csr keeps the packed
CSR bytes). The dominant kernel's roofline figure uses its algorithmic bytes
per launch (input bytes + per-string begin/end + final-state and accept
bytes) over its average launch duration, measured with HIP events on the
launch stream. The results of the last step are checked bit-exact against the
multithreaded C oracle over the whole batch; the oracle's time on that check
is the CPU baseline.

Secondary legs (reported under "secondary"):
  * nfa64: one Boolean NFA-64 (64 states, each per-symbol edge present with
    p = 0.06, 32 symbols, state 0 start, 4 accept states) over the same 1 GiB
    batch, bit-exact against the C oracle's set walk;
  * prob_nfa64: probabilistic NFA-64, 32 symbols, T[c] column-stochastic with
    Dirichlet(0.1) columns, uniform start, 16 accept states; 16,384 sequences
    x 2,048 uniform symbols; every mode's relative log-likelihood error
    against the FP64 C oracle (whose time is that leg's CPU baseline) and the
    MFMA roofline (2 * 64^2 FLOP per symbol).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--legs dfa,nfa,prob]
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
MFMA_F16_PEAK_TFLOPS = 2500.0
N_STRINGS = 262_144
LENGTH = 4096
N_PATTERNS = 8
SEED = 2
PROB_SEQ = 16_384
PROB_LEN = 2048
KERNEL_NAMES = {("k2_window", 16): "dfa16_k2w_pf16", ("k2_window", 8): "dfa16_k2w_pf8",
                ("k2_window_f", 16): "dfa16_k2f_pf16", ("k2_window_f", 8): "dfa16_k2f_pf8",
                ("k1_16", 16): "dfa16_k1", ("k1_16", 8): "dfa16_k1"}
PROB_WAVES = (4, 8)


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
    ap.add_argument("--legs", default="dfa,nfa,prob", help="comma list of dfa, nfa, prob")
    ap.add_argument("--prob-modes", default="f16x2,bf16x2,f16,bf16")
    ap.add_argument("--sweep", action="store_true", help="time every kernel variant and layout, then exit")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also time a plain 1 GiB device copy (HBM traffic counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}), flush=True)
        return 2
    legs = [x for x in args.legs.split(",") if x]

    from trex_amd import _lib
    from trex_amd.device import DeviceBuffer, Stream
    from trex_amd.engine import DeviceBatch
    from trex_amd.workloads import uniform_batch

    _lib.check(_lib.lib().trex_device_check(), "trex_device_check")
    stream = Stream()
    result: dict = {}
    secondary: dict = {}
    ok = True

    batch = data = offsets = None
    if "dfa" in legs or "nfa" in legs or args.sweep:
        rng = np.random.default_rng(SEED)
        t0 = time.perf_counter()
        data, offsets = uniform_batch(rng, N_STRINGS, LENGTH, 0x40, 0x60)
        log(f"generated {data.nbytes / 2**30:.2f} GiB in {time.perf_counter() - t0:.1f}s")
        t0 = time.perf_counter()
        batch = DeviceBatch(data, offsets, layout=args.layout)
        log(f"uploaded ({args.layout} layout, {batch.d_bytes.nbytes} device bytes) in {time.perf_counter() - t0:.1f}s")

    if args.copy_probe and batch is not None:
        d_copy = DeviceBuffer(data.nbytes)
        ms = time_launches(lambda: _lib.call("trex_copy", d_copy.ptr, batch.d_bytes.ptr, data.nbytes, stream.handle),
                           stream, 5)
        copy_gbs = 2 * data.nbytes / (ms * 1e-3) / 1e9
        log(f"copy probe: {data.nbytes / 2**30:.2f} GiB copy {ms:.4f} ms = {copy_gbs:.0f} GB/s (read+write)")
        d_copy.free()
    else:
        copy_gbs = None

    if "dfa" in legs or args.sweep:
        rc = dfa_leg(args, data, offsets, batch, stream, result, copy_gbs)
        if args.sweep:
            return rc
        ok &= rc == 0
    if "nfa" in legs:
        secondary["nfa64"] = nfa_leg(args, data, offsets, batch, stream)
        ok &= secondary["nfa64"].get("verified_bit_exact") in (None, True)
    if batch is not None:
        batch.free()
    if "prob" in legs:
        secondary["prob_nfa64"] = prob_leg(args, stream)
        ok &= secondary["prob_nfa64"].get("within_tolerance") in (None, True)

    if not result:   # a run without the headline leg: report what ran
        result = {"metric": "secondary legs only", "value": None, "n_gpus": 1, "steps": args.steps,
                  "warmup": args.warmup}
    result["secondary"] = secondary
    print(json.dumps(result), flush=True)
    return 0 if ok else 1


def dfa_leg(args, data, offsets, batch, stream, result: dict, copy_gbs) -> int:
    from trex_amd.device import DeviceBuffer, Event
    from trex_amd.engine import DfaEngine
    from trex_amd.tables import dfa_byte_table
    from trex_amd.workloads import ALPHABET_32, random_dfa_pattern

    rng = np.random.default_rng(SEED)
    patterns = [random_dfa_pattern(rng, ALPHABET_32) for _ in range(N_PATTERNS)]
    tables = [dfa_byte_table(d) for _, d in patterns]
    log(f"patterns: {[p for p, _ in patterns]}; states {[t.n_states for t in tables]}")
    n = N_STRINGS
    outs = [(DeviceBuffer(n), DeviceBuffer(n)) for _ in range(N_PATTERNS)]
    engines = [DfaEngine(t) for t in tables]
    for e in engines:
        if args.variant != "auto":
            e.set_variant(args.variant)
        e.set_prefetch(args.prefetch)
    variants = [e.info()["variant"] for e in engines]
    log(f"variants: {variants}, prefetch {args.prefetch}")

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
    log(f"dfa step {ms_per_step:.3f} ms (wall {wall / args.steps * 1e3:.3f} ms), kernel avg {avg_kernel_ms:.4f} ms "
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
            log("PARITY FAILURE (dfa)")
    for a, b in outs:
        a.free()
        b.free()

    kernel = KERNEL_NAMES.get((variants[0], args.prefetch), variants[0])
    result.update({
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
    })
    return 0 if verified in (None, True) else 1


def nfa_leg(args, data, offsets, batch, stream) -> dict:
    from trex_amd.device import DeviceBuffer, Event
    from trex_amd.engine import NfaEngine
    from trex_amd.workloads import random_nfa64

    nm = random_nfa64(np.random.default_rng(SEED + 1))
    eng = NfaEngine.from_matrices(nm)
    n = N_STRINGS
    d_mask, d_acc = DeviceBuffer(8 * n), DeviceBuffer(n)
    reps = max(3, args.steps // 2)
    launch = lambda: eng.match_batch_device(batch, d_mask.ptr, d_acc.ptr, stream)
    for _ in range(args.warmup):
        launch()
    evs = [(Event(), Event()) for _ in range(reps)]
    for a, b in evs:
        a.record(stream)
        launch()
        b.record(stream)
    stream.synchronize()
    ms = [a.elapsed_ms(b) for a, b in evs]
    avg = float(np.mean(ms))
    alg_bytes = data.nbytes + 16 * n + 9 * n
    gchars = data.nbytes / (avg * 1e-3) / 1e9
    achieved = alg_bytes / (avg * 1e-3) / 1e9
    log(f"nfa64 ({eng.info()['variant']}): kernel avg {avg:.4f} ms min {min(ms):.4f} max {max(ms):.4f}, "
        f"{gchars:.1f} Gchars/s, {achieved:.0f} GB/s")
    out = {"workload": "nfa64_bitpar_1GiB", "variant": eng.info()["variant"], "strings": n, "length": LENGTH,
           "value": round(gchars, 2), "unit": "Gchars/s", "kernel_avg_ms": round(avg, 5),
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "alg_bytes_per_launch": alg_bytes,
                        "note": "16 LDS table lookups per byte bound the kernel before HBM does"},
           "verified_bit_exact": None, "cpu_baseline": None}
    if not args.no_verify:
        from oracle import cpu as ocpu
        from trex_amd.engine import nfa_byte_map
        threads = ocpu.default_threads()
        t0 = time.perf_counter()
        rm, ra = ocpu.nfa_walk(nm.successor_masks(), nfa_byte_map(nm, "reject"), nm.start_mask(), nm.accept_bits(),
                               data, offsets, threads=threads)
        t_cpu = time.perf_counter() - t0
        gm = d_mask.download(np.uint64, n)
        ga = d_acc.download(np.uint8, n)
        same = bool(np.array_equal(gm, rm) and np.array_equal(ga, ra))
        log(f"nfa64: bit-exact={same} accept-rate={float(ra.mean()):.4f}, C oracle {t_cpu:.2f} s")
        out["verified_bit_exact"] = same
        out["cpu_baseline"] = {"value": round(data.nbytes / t_cpu / 1e9, 3), "unit": "Gchars/s", "cores": threads,
                               "kind": "port", "sample": "full 1 GiB batch through the C oracle set walk"}
        if not same:
            log("PARITY FAILURE (nfa64)")
    d_mask.free()
    d_acc.free()
    eng.close()
    return out


def prob_leg(args, stream) -> dict:
    from trex_amd.device import DeviceBuffer, Event
    from trex_amd.engine import ProbNfaEngine
    from trex_amd.workloads import dirichlet_pnfa

    rng = np.random.default_rng(SEED + 2)
    T, start, accept = dirichlet_pnfa(rng, 32)
    sym = rng.integers(0, 32, size=(PROB_SEQ, PROB_LEN), dtype=np.uint8)
    d_sym = DeviceBuffer.from_array(sym)
    d_ll, d_lm = DeviceBuffer(8 * PROB_SEQ), DeviceBuffer(8 * PROB_SEQ)
    eng = ProbNfaEngine(T, start, accept)
    flops = 2.0 * 64 * 64 * PROB_SEQ * PROB_LEN
    modes = [m for m in args.prob_modes.split(",") if m]
    reps = max(3, args.steps // 2)
    per_mode = {}
    for mode in modes:
        eng.set_mode(mode)
        launch = lambda: eng.loglik_device(d_sym.ptr, PROB_SEQ, PROB_LEN, d_ll.ptr, d_lm.ptr, stream)
        timed = {}
        for waves in PROB_WAVES:
            eng.set_waves(waves)
            for _ in range(max(1, args.warmup)):
                launch()
            evs = [(Event(), Event()) for _ in range(reps)]
            for a, b in evs:
                a.record(stream)
                launch()
                b.record(stream)
            stream.synchronize()
            ms = [a.elapsed_ms(b) for a, b in evs]
            timed[waves] = (float(np.mean(ms)), d_ll.download(np.float64, PROB_SEQ, stream),
                            d_lm.download(np.float64, PROB_SEQ, stream))
            log(f"prob {mode} w{waves}: kernel avg {timed[waves][0]:.4f} ms (min {min(ms):.4f})")
        waves = min(timed, key=lambda w: timed[w][0])
        avg, ll, lm = timed[waves]
        per_mode[mode] = {"kernel": f"pnfa64_{mode}_w{waves}", "kernel_avg_ms": round(avg, 4),
                          "kernel_avg_ms_by_waves": {str(w): round(v[0], 4) for w, v in timed.items()},
                          "gchars_s": round(PROB_SEQ * PROB_LEN / (avg * 1e-3) / 1e9, 3),
                          "tflops": round(flops / (avg * 1e-3) / 1e12, 2), "_ll": ll, "_lm": lm}
        log(f"prob {mode}: best w{waves} {avg:.4f} ms, "
            f"{per_mode[mode]['gchars_s']:.2f} Gchars/s, {per_mode[mode]['tflops']:.1f} TFLOP/s")

    cpu = None
    within = None
    if not args.no_verify:
        from oracle import cpu as ocpu
        threads = ocpu.default_threads()
        t0 = time.perf_counter()
        ref_ll, ref_lm = ocpu.pnfa_forward(T, start, accept, sym, threads=threads)
        t_cpu = time.perf_counter() - t0
        cpu = {"value": round(PROB_SEQ * PROB_LEN / t_cpu / 1e9, 4), "unit": "Gchars/s", "cores": threads,
               "kind": "port", "sample": f"full workload through the FP64 C oracle, {threads} threads, {t_cpu:.2f} s"}
        within = True
        for mode, r in per_mode.items():
            err = np.abs(r["_ll"] - ref_ll) / np.abs(ref_ll)
            r["rel_ll_err_vs_fp64"] = {"median": float(f"{np.median(err):.3e}"), "max": float(f"{err.max():.3e}")}
            r["logmass_abs_err_max"] = float(f"{np.abs(r['_lm'] - ref_lm).max():.3e}")
            if mode.endswith("x2"):
                tq, sc = eng.rounded(mode)
                q_ll, _ = ocpu.pnfa_forward(tq, start, accept, sym, col_scale=sc, T_last=T, threads=threads)
                e2 = np.abs(r["_ll"] - q_ll) / np.abs(q_ll)
                r["rel_ll_err_vs_fp64_rounded_T"] = {"median": float(f"{np.median(e2):.3e}"),
                                                     "max": float(f"{e2.max():.3e}")}
                within &= bool(err.max() <= 1e-3)
            log(f"prob {mode}: rel LL err vs FP64 median {np.median(err):.2e} max {err.max():.2e}")
    for r in per_mode.values():
        r.pop("_ll")
        r.pop("_lm")
    head = per_mode.get("f16x2") or next(iter(per_mode.values()))
    head_ms = head["kernel_avg_ms"]
    achieved = flops / (head_ms * 1e-3) / 1e12
    d_sym.free()
    d_ll.free()
    d_lm.free()
    eng.close()
    return {"workload": "prob_nfa64_dirichlet0.1", "sequences": PROB_SEQ, "length": PROB_LEN, "symbols": 32,
            "value": head["gchars_s"], "unit": "Gchars/s", "dtype": "f16 MFMA, f32 accumulate, f64 log-scale",
            "roofline": {"bound": "mfma", "achieved": round(achieved, 2), "peak": MFMA_F16_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(achieved / MFMA_F16_PEAK_TFLOPS, 4),
                         "kernel": head["kernel"], "alg_flops_per_launch": flops,
                         "note": "algorithmic FLOPs = 2*64^2 per symbol (one matvec); the split mode issues 2x"},
            "modes": per_mode, "cpu_baseline": cpu, "within_tolerance": within}


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
