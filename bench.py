"""Benchmark: DFA-16 K-gram-fused batch over 1 GiB on one MI355X, with the
Boolean NFA-64 and the probabilistic NFA-64 as secondary legs.

Headline workload (one "step"): 8 seeded random regexes over the 32-symbol
alphabet 0x40-0x5F, each compiled to a <= 16-state DFA (bytes outside the
alphabet go to the dead state); 262,144 strings x 4,096 bytes = 1 GiB of
i.i.d. uniform symbols, resident in HBM before timing; every pattern walks
every string (8 kernel launches per step); outputs are the final state and
accept bit per string. Gchars/s = 8 x 2^30 / step time.

The batch is held in the engine's interleaved HBM layout (each group of 64
strings stored as rows of 64 x 16 bytes, so a wave's load reads one 1 KiB
row; --layout staggered keeps each string contiguous with groups 256 bytes
apart, --layout csr the packed CSR bytes). Building that layout is part of
the upload, outside the timed region, like the upload itself. The dominant kernel's roofline figure uses its algorithmic bytes
per launch (input bytes + per-string begin/end + final-state and accept
bytes) over its average launch duration, measured with HIP events on the
launch stream. The results of the last step are checked bit-exact against the
multithreaded C oracle over the whole batch; the oracle's time on that check
is the CPU baseline.

Beside the headline, "multi_pattern" times the same 8 patterns over the same
batch in one launch of the multi-pattern kernel (one read of the bytes
advances all 8), checked against the same oracle results.

Secondary legs (reported under "secondary"):
  * nfa64: one Boolean NFA-64 (64 states, each per-symbol edge present with
    p = 0.06, 32 symbols, state 0 start, 4 accept states) over the same 1 GiB
    batch, bit-exact against the C oracle's set walk;
  * lookback_scan: the same 1 GiB as 16 strings of 64 MiB, the DFA leg's
    first pattern, cut into 64 KiB chunks chained by the decoupled look-back
    (one launch), bit-exact against the C oracle;
  * prob_nfa64: probabilistic NFA-64, 32 symbols, T[c] column-stochastic with
    Dirichlet(0.1) columns, uniform start, 16 accept states; 16,384 sequences
    x 2,048 uniform symbols; every mode's relative log-likelihood error
    against the FP64 C oracle (whose time is that leg's CPU baseline) and the
    MFMA roofline (2 * 64^2 FLOP per symbol);
  * uneven_mix: 64 patterns (32 DFA-16, 32 NFA-64) over 65,536 strings of
    log-uniform lengths in [32 B, 16 MiB] in one persistent autotuned launch,
    against one launch per pattern on a subset;
  * serving: Poisson arrivals over 128 resident patterns, p50/p99 latency,
    and a saturated burst served vs one synchronous launch per request.
  * host_input: the Python API end to end on host strings, 10,000 str
    strings x 1,000 characters through OptimizedEngine.match_batch (packing,
    copies, kernel, results as list[bool]), split into its stages;
  * plain_dfa_cpu_path: the reference's CPU path on its plain DFA workload
    (seeded regex over abcd, 1,024 x 1,024 B): the Python DFA.simulate
    restatement, the C oracle on one and on all threads, and the GPU beside
    them, bit-exact.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--legs dfa,nfa,scan,prob,mix,serve,host,plain]
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
LDS_PEAK_TBS = 157.3   # 256 CUs x 256 B per clock x 2.4 GHz
# HBM bytes per launch of the headline kernel on this workload, from two rocprofv3
# counter runs (FETCH_SIZE x 2.000 + WRITE_SIZE, both calibrated against a plain
# 1 GiB copy); the command lines and counts are in the profile named with each.
PMC_TRAFFIC = {
    ("dfa16_k2f_il_pf8", "interleaved"): (1_078_713_000, "profiles/dfa16_k2f_il_pf8_1GiB_8patterns.md"),
    ("dfa16_k2f_pf8", "staggered"): (1_092_300_000, "profiles/dfa16_k2f_pf8_1GiB_8patterns.md"),
}
MFMA_F16_PEAK_TFLOPS = 2500.0
N_STRINGS = 262_144
LENGTH = 4096
N_PATTERNS = 8
SEED = 2
PROB_SEQ = 16_384
PROB_LEN = 2048
KERNEL_FORMS = {"k2_window": "dfa16_k2w", "k2_window_f": "dfa16_k2f", "k2_window_fr": "dfa16_k2fr"}


def kernel_name(variant: str, prefetch: int, layout: str) -> str:
    """Name of the code-object kernel a DFA dispatch launches."""
    il = "_il" if layout == "interleaved" else ""
    if variant in KERNEL_FORMS:
        return f"{KERNEL_FORMS[variant]}{il}_pf{16 if prefetch == 16 else 8}"
    return ("dfa16_k1" if variant == "k1_16" else "dfa_k1") + il


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
    ap.add_argument("--layout", default="interleaved", choices=["interleaved", "staggered", "csr"],
                    help="HBM layout of the DFA batch (the NFA leg reads a staggered copy when this is interleaved)")
    ap.add_argument("--legs", default="dfa,nfa,scan,prob,mix,serve,host,plain",
                    help="comma list of dfa, nfa, scan, prob, mix, serve, host, plain")
    ap.add_argument("--prob-modes", default="f16x2,bf16x2,f16,bf16")
    ap.add_argument("--mix-strings", type=int, default=MIX_STRINGS, help="strings of the uneven-mix leg")
    ap.add_argument("--mix-tune", type=int, default=1, help="autotune the uneven-mix plan (0: defaults)")
    ap.add_argument("--mix-max-spins", type=int, default=0, help="bound of every wait in the mixed queue (0: default)")
    ap.add_argument("--serve-qps", type=float, default=100.0, help="Poisson arrival rate of the serving leg")
    ap.add_argument("--serve-seconds", type=float, default=20.0)
    ap.add_argument("--serve-burst", type=int, default=24, help="requests of the served-vs-plain burst")
    ap.add_argument("--serve-small-qps", type=float, default=40.0,
                    help="Poisson rate of the second arrival stream, requests of at most 1 MiB")
    ap.add_argument("--serve-hbm-gb", type=int, default=64, help="HBM budget of the serving arena")
    ap.add_argument("--serve-prob-cap", type=int, default=1 << 16,
                    help="longest probabilistic sequence in serving requests (each symbol is a sequential step)")
    ap.add_argument("--sweep", action="store_true", help="time every kernel variant and layout, then exit")
    ap.add_argument("--serve-mix-form", default="nibble", choices=["nibble", "slice"],
                    help="NFA table form of the server's mixed-queue lanes")
    ap.add_argument("--scan-sweep", action="store_true", help="also time the look-back scan for several chunk lengths")
    ap.add_argument("--scan-chains", type=int, default=0,
                    help="dependent chains per lane in the look-back scan's first phase (0: library default)")
    ap.add_argument("--scan-gaps", default="",
                    help="comma-separated tile gaps (bytes) to also time the look-back scan's chunk image with")
    ap.add_argument("--multi-sweep", action="store_true",
                    help="also time the multi-pattern pass for every table form and patterns-per-launch")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also time a plain 1 GiB device copy (HBM traffic counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}), flush=True)
        return 2
    legs = [x for x in args.legs.split(",") if x]
    if "serve" in legs:
        # one hardware queue per server stream (a copy stream, two DFA/NFA lanes, four
        # probabilistic slots), the bench's own and the default stream, set before the
        # HIP runtime starts: streams that share a queue run in order, so a small
        # request's tick would wait behind a long probabilistic launch. Set outright:
        # an environment that exports HIP's default of 4 would otherwise keep it.
        os.environ["GPU_MAX_HW_QUEUES"] = str(max(12, int(os.environ.get("GPU_MAX_HW_QUEUES", "0") or 0)))

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
    if "dfa" in legs or "nfa" in legs or "scan" in legs or args.sweep:
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
        if batch.layout == "interleaved":   # the NFA kernels read contiguous strings
            batch.free()
            batch = DeviceBatch(data, offsets, layout="staggered")
        secondary["nfa64"] = nfa_leg(args, data, offsets, batch, stream)
        ok &= secondary["nfa64"].get("verified_bit_exact") in (None, True)
    if batch is not None:
        batch.free()
    if "scan" in legs:
        secondary["lookback_scan"] = scan_leg(args, data, stream)
        ok &= secondary["lookback_scan"].get("verified_bit_exact") in (None, True)
    if "prob" in legs:
        secondary["prob_nfa64"] = prob_leg(args, stream)
        ok &= secondary["prob_nfa64"].get("within_tolerance") in (None, True)
    if "mix" in legs:
        secondary["uneven_mix"] = mix_leg(args, stream)
        ok &= secondary["uneven_mix"].get("verified_bit_exact") in (None, True)
    if "serve" in legs:
        secondary["serving"] = serve_leg(args, stream)
        ok &= secondary["serving"]["burst"]["same_results"]
    if "host" in legs:
        secondary["host_input"] = host_leg(args)
        ok &= secondary["host_input"].get("verified") in (None, True)
    if "plain" in legs:
        secondary["plain_dfa_cpu_path"] = plain_leg(args, stream)
        ok &= secondary["plain_dfa_cpu_path"]["verified_bit_exact"]

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

    from trex_amd.device import Graph

    def timed_eager():
        step_ev[0].record(stream)
        w0 = time.perf_counter()
        for k in range(args.steps):
            step(k)
        step_ev[1].record(stream)
        stream.synchronize()
        return time.perf_counter() - w0

    for _ in range(args.warmup):
        step(None)
    stream.synchronize()
    # the timed region: the K steps recorded as one HIP graph and launched
    # once (events on the stream around it), as a serving loop replays it.
    # Events recorded inside a graph cannot be timed, so the per-launch
    # kernel times come from the same K steps launched one by one right after.
    launch_mode = "graph"
    try:
        with Graph.capture(stream) as steps_graph:
            for k in range(args.steps):
                step(None)
        stream.synchronize()
        step_ev[0].record(stream)
        wall0 = time.perf_counter()
        steps_graph.launch(stream)
        step_ev[1].record(stream)
        stream.synchronize()
        wall = time.perf_counter() - wall0
        total_ms = step_ev[0].elapsed_ms(step_ev[1])
        steps_graph.close()
    except Exception as e:   # a runtime that cannot capture: time the launches one by one
        log(f"graph of the timed steps unavailable ({e}); timing eager launches")
        launch_mode = "eager"
        stream.synchronize()
        wall = timed_eager()
        kernel_ms = [a.elapsed_ms(b) for row in evs for a, b in row]
        total_ms = step_ev[0].elapsed_ms(step_ev[1])
    ms_per_step = total_ms / args.steps
    eager_res = None
    if launch_mode == "graph":
        e_wall = timed_eager()
        kernel_ms = [a.elapsed_ms(b) for row in evs for a, b in row]
        e_ms = step_ev[0].elapsed_ms(step_ev[1]) / args.steps
        eager_res = {"ms_per_step": round(e_ms, 4), "kernel_avg_ms": round(float(np.mean(kernel_ms)), 5),
                     "wall_ms_per_step": round(e_wall / args.steps * 1e3, 4),
                     "note": "the same K steps launched one by one after the timed graph; roofline kernel times "
                             "come from these launches' events"}
        log(f"dfa step, launches one by one: {e_ms:.3f} ms, kernel avg {np.mean(kernel_ms):.4f} ms")
    chars = float(N_PATTERNS) * data.nbytes
    gchars = chars / (ms_per_step * 1e-3) / 1e9
    avg_kernel_ms = float(np.mean(kernel_ms))
    alg_bytes = data.nbytes + 16 * n + 2 * n
    achieved = alg_bytes / (avg_kernel_ms * 1e-3) / 1e9
    log(f"dfa step ({launch_mode}) {ms_per_step:.3f} ms (wall {wall / args.steps * 1e3:.3f} ms), kernel avg {avg_kernel_ms:.4f} ms "
        f"min {min(kernel_ms):.4f} max {max(kernel_ms):.4f}, {gchars:.1f} Gchars/s, {achieved:.0f} GB/s")

    # the same kernel family on the packed CSR layout (what OptimizedEngine and
    # the reference's bridges hand over): per-launch time and results
    csr_res = None
    if args.layout == "interleaved":
        from trex_amd.engine import DeviceBatch
        csr_batch = DeviceBatch(data, offsets, layout="csr")
        csr_outs = (DeviceBuffer(n), DeviceBuffer(n))
        csr_res = {"layout": "csr"}
        # read order of the packed strings: as stored (rotation 0), then the
        # library's default (rotated per wave for tables that synchronise)
        for mode, key in ((0, "unrotated"), (-1, "default")):
            c_evs = [(Event(), Event()) for _ in range(N_PATTERNS * 2)]
            csr_same = True
            rotated = []
            for eng in engines:
                eng.set_rotation(mode)
                rotated.append(eng.rotation()[0])
            for r in range(2):
                for i, eng in enumerate(engines):
                    a, b = c_evs[r * N_PATTERNS + i]
                    a.record(stream)
                    eng.match_batch_device(csr_batch, csr_outs[0].ptr, csr_outs[1].ptr, stream)
                    b.record(stream)
                    if r == 1:
                        csr_same &= bool(np.array_equal(csr_outs[0].download(np.uint8, n, stream),
                                                        outs[i][0].download(np.uint8, n, stream))
                                         and np.array_equal(csr_outs[1].download(np.uint8, n, stream),
                                                            outs[i][1].download(np.uint8, n, stream)))
            stream.synchronize()
            for eng in engines:
                eng.set_rotation(-1)
            c_ms = float(np.mean([a.elapsed_ms(b) for a, b in c_evs[N_PATTERNS:]]))
            c_ach = alg_bytes / (c_ms * 1e-3) / 1e9
            kname = "dfa16_k2f_rot_pf8" if all(rotated) else kernel_name(variants[0], args.prefetch, "csr")
            csr_res[key] = {"kernel": kname, "patterns_rotated": int(sum(rotated)),
                            "kernel_avg_ms": round(c_ms, 5), "achieved_gbs": round(c_ach, 1),
                            "frac": round(c_ach / HBM_PEAK_GBS, 4), "same_results_as_interleaved": csr_same}
            log(f"dfa on packed CSR ({key} read order, {sum(rotated)} of {len(engines)} patterns rotated): "
                f"kernel avg {c_ms:.4f} ms, {c_ach:.0f} GB/s ({c_ach / HBM_PEAK_GBS:.1%}), same results: {csr_same}")
        csr_res.update({k: csr_res["default"][k] for k in ("kernel", "kernel_avg_ms", "achieved_gbs", "frac")})
        csr_res["same_results_as_interleaved"] = (csr_res["default"]["same_results_as_interleaved"]
                                                  and csr_res["unrotated"]["same_results_as_interleaved"])
        csr_outs[0].free()
        csr_outs[1].free()
        csr_batch.free()

    multi_res, multi_outs = multi_leg(args, engines, batch, stream, ms_per_step, data.nbytes)

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
            if multi_outs is not None:
                mf = multi_outs[0].download(np.uint8, n, offset=i * n)
                ma = multi_outs[1].download(np.uint8, n, offset=i * n)
                m_same = bool(np.array_equal(mf, rf) and np.array_equal(ma, ra))
                multi_res["verified_bit_exact"] = multi_res.get("verified_bit_exact", True) and m_same
                ok &= m_same
                if not m_same:
                    log(f"pattern {i}: multi-pattern pass differs from the oracle")
        if csr_res is not None:   # the CSR launches' results equal the interleaved ones checked above
            ok &= bool(csr_res["same_results_as_interleaved"])
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
    for a, b in outs + ([multi_outs] if multi_outs is not None else []):
        a.free()
        b.free()

    kernel = kernel_name(variants[0], args.prefetch, args.layout)
    if (args.layout != "interleaved" and kernel == "dfa16_k2f_pf8"
            and all(eng.rotation()[0] for eng in engines)):
        kernel = "dfa16_k2f_rot_pf8"
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
        "config": {"workload": "dfa16_kgram_1GiB_8patterns", "patterns": N_PATTERNS, "strings": N_STRINGS,
                   "length": LENGTH, "alphabet": "0x40-0x5F", "kernel_variants": variants,
                   "kgram": "4 bytes per word step as two K=2 pair-function lookups (a K=4 table would be 16 MiB)",
                   "prefetch": args.prefetch, "layout": args.layout, "launches_per_step": N_PATTERNS},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": PMC_TRAFFIC.get((kernel, args.layout), (None, None))[0],
                     "traffic_source": ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, one pass each: "
                                        + PMC_TRAFFIC[(kernel, args.layout)][1])
                     if (kernel, args.layout) in PMC_TRAFFIC else None,
                     "kernel": kernel, "kernel_avg_ms": round(avg_kernel_ms, 5),
                     "alg_bytes_per_launch": alg_bytes, "copy_probe_gbs": copy_gbs},
        "cpu_baseline": cpu,
        "verified_bit_exact": verified,
        "multi_pattern": multi_res,
        "launch": launch_mode,
        "eager_launches": eager_res,
        "csr_layout": csr_res,
    })
    return 0 if verified in (None, True) else 1


def scan_leg(args, data, stream) -> dict:
    """Long strings through the decoupled look-back scan: the 1 GiB batch read
    as 16 strings of 64 MiB, cut into chunks of lookback_chunk_len (16 KiB
    here; one launch for all chunks), with the DFA leg's first pattern,
    directly and over the plan's chunk-interleaved image."""
    from trex_amd.device import DeviceBuffer
    from trex_amd.engine import ChunkPlan, DeviceBatch, DfaEngine
    from trex_amd.tables import dfa_byte_table
    from trex_amd.workloads import ALPHABET_32, random_dfa_pattern

    n_long = 16
    per = data.nbytes // n_long
    offsets = np.arange(n_long + 1, dtype=np.int64) * per
    table = dfa_byte_table(random_dfa_pattern(np.random.default_rng(SEED), ALPHABET_32)[1])
    eng = DfaEngine(table)
    batch = DeviceBatch(data, offsets, layout="csr")
    d_f, d_a = DeviceBuffer(n_long), DeviceBuffer(n_long)
    if args.scan_sweep:
        for clen in (4096, 8192, 12288, 16384, 16400, 16640, 20480, 24576, 32768, 65536):
            pl = ChunkPlan.for_batch(batch, clen)
            if args.scan_chains:
                pl.set_chains(args.scan_chains)
            t = time_launches(lambda: eng.match_plan(pl, batch.d_bytes.ptr, d_f.ptr, d_a.ptr, stream), stream,
                              max(3, args.steps))
            img = pl.build_image(data)
            d_im = DeviceBuffer(img.nbytes)
            d_im.upload(img, stream=stream)
            stream.synchronize()
            del img
            ti = time_launches(lambda: eng.match_plan_image(pl, d_im.ptr, d_im.nbytes, d_f.ptr, d_a.ptr, stream),
                               stream, max(3, args.steps))
            log(f"look-back scan sweep: {clen:6d} B chunks ({pl.n_chunks} chunks): {t:.4f} ms "
                f"({data.nbytes / (t * 1e-3) / 1e9:.0f} Gchars/s), chunk image {ti:.4f} ms "
                f"({data.nbytes / (ti * 1e-3) / 1e9:.0f} Gchars/s), error word {pl.error(stream)}")
            d_im.free()
            pl.close()
    plan = ChunkPlan.for_batch(batch)
    if args.scan_chains:
        plan.set_chains(args.scan_chains)
    ms = time_launches(lambda: eng.match_plan(plan, batch.d_bytes.ptr, d_f.ptr, d_a.ptr, stream), stream,
                       max(3, args.steps))
    err = plan.error(stream)
    alg_bytes = data.nbytes + 2 * n_long
    gchars = data.nbytes / (ms * 1e-3) / 1e9
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    log(f"look-back scan: {n_long} x {per >> 20} MiB, {plan.n_chunks} chunks of {plan.chunk_len >> 10} KiB, "
        f"{ms:.4f} ms per launch, "
        f"{gchars:.1f} Gchars/s, {achieved:.0f} GB/s, error word {err}")
    out = {"workload": f"{n_long} strings x {per >> 20} MiB, {plan.chunk_len >> 10} KiB chunks chained by the look-back",
           "kernel": "dfa16_chunks_window", "chunks": plan.n_chunks, "chunk_len": plan.chunk_len, "kernel_avg_ms": round(ms, 4),
           "value": round(gchars, 1), "unit": "Gchars/s", "error_word": int(err),
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "alg_bytes_per_launch": alg_bytes},
           "verified_bit_exact": None}
    rf = ra = None
    if not args.no_verify:
        from oracle import cpu as ocpu
        threads = min(ocpu.default_threads(), n_long)
        t0 = time.perf_counter()
        rf, ra = ocpu.dfa_walk(table.delta, table.accept, table.start, data, offsets, threads=threads)
        t_cpu = time.perf_counter() - t0
        same = bool(np.array_equal(d_f.download(np.uint8, n_long), rf)
                    and np.array_equal(d_a.download(np.uint8, n_long), ra) and err == 0)
        out["verified_bit_exact"] = same
        out["cpu_baseline"] = {"value": round(data.nbytes / t_cpu / 1e9, 3), "unit": "Gchars/s", "cores": threads,
                               "kind": "port",
                               "sample": f"the same 16 x 64 MiB through the C oracle byte walk (one thread per "
                                         f"string, {threads} threads), {t_cpu:.2f} s; host {cpu_model()}"}
        log(f"look-back scan: bit-exact={same}; C oracle {data.nbytes / t_cpu / 1e9:.2f} Gchars/s on {threads} threads")
        if not same:
            log("PARITY FAILURE (look-back scan)")

    # the same plan over its chunk-interleaved image (one wave load = 1 KiB)
    t0 = time.perf_counter()
    image = plan.build_image(data)
    t_build = time.perf_counter() - t0
    d_img = DeviceBuffer(image.nbytes)
    d_img.upload(image, stream=stream)
    stream.synchronize()
    del image
    ims = time_launches(lambda: eng.match_plan_image(plan, d_img.ptr, d_img.nbytes, d_f.ptr, d_a.ptr, stream),
                        stream, max(3, args.steps))
    ierr = plan.error(stream)
    iach = alg_bytes / (ims * 1e-3) / 1e9
    log(f"look-back scan over the chunk image ({plan.image_size() >> 20} MiB, built on the host in "
        f"{t_build:.2f} s): {ims:.4f} ms per launch, {data.nbytes / (ims * 1e-3) / 1e9:.1f} Gchars/s, "
        f"{iach:.0f} GB/s, error word {ierr}")
    out["csr"] = {"kernel": "dfa16_chunks_window", "kernel_avg_ms": out["kernel_avg_ms"], "value": out["value"],
                  "roofline_frac": out["roofline"]["frac"]}
    out.update({"kernel": "dfa16_chunks_window_img", "kernel_avg_ms": round(ims, 4),
                "value": round(data.nbytes / (ims * 1e-3) / 1e9, 1), "error_word": int(ierr),
                "layout": "chunk-interleaved image of the plan (tile of 64 chunks, 16-byte vectors side by side)"})
    out["roofline"].update({"achieved": round(iach, 1), "frac": round(iach / HBM_PEAK_GBS, 4)})
    if not args.no_verify:
        same = bool(np.array_equal(d_f.download(np.uint8, n_long), rf)
                    and np.array_equal(d_a.download(np.uint8, n_long), ra) and ierr == 0)
        out["verified_bit_exact"] = bool(out["verified_bit_exact"]) and same
        log(f"look-back scan (chunk image): bit-exact={same}")
        if not same:
            log("PARITY FAILURE (look-back scan, chunk image)")
    if args.scan_gaps:
        # the image under other gaps after each tile (0 = tiles back to back)
        gaps = {}
        for gap in (int(g) for g in args.scan_gaps.split(",")):
            plan.set_image_gap(gap)
            d_img.free()
            image = plan.build_image(data)
            d_img = DeviceBuffer(image.nbytes)
            d_img.upload(image, stream=stream)
            stream.synchronize()
            del image
            gms = time_launches(lambda: eng.match_plan_image(plan, d_img.ptr, d_img.nbytes, d_f.ptr, d_a.ptr,
                                                             stream), stream, max(3, args.steps))
            gerr = plan.error(stream)
            ok = None
            if not args.no_verify:
                ok = bool(np.array_equal(d_f.download(np.uint8, n_long), rf)
                          and np.array_equal(d_a.download(np.uint8, n_long), ra) and gerr == 0)
            gaps[gap] = {"kernel_avg_ms": round(gms, 4), "frac": round(alg_bytes / (gms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                         "error_word": int(gerr), "verified_bit_exact": ok}
            log(f"look-back scan over the chunk image, tile gap {gap} B: {gms:.4f} ms per launch, "
                f"{alg_bytes / (gms * 1e-3) / 1e9:.0f} GB/s, error word {gerr}, bit-exact={ok}")
        out["image_gaps"] = gaps
        plan.set_image_gap(256)
    # an automaton that never synchronises: every chunk keeps a permutation as
    # its function and resolves through the look-back chain
    from trex_amd.tables import DfaByteTable
    delta = np.full((16, 256), 15, np.uint8)
    for b in range(0x40, 0x60):
        delta[:15, b] = (np.arange(15) + (b & 15)) % 15
    accept = np.zeros(16, np.uint8)
    accept[0] = 1
    perm = DfaByteTable(delta, accept, 0, 15, 16)
    peng = DfaEngine(perm)
    from trex_amd.engine import lookback_chunk_len, lookback_target_chunks
    pplan = ChunkPlan.for_batch(batch, lookback_chunk_len(data.nbytes, lookback_target_chunks(perm)))
    if args.scan_chains:
        pplan.set_chains(args.scan_chains)
    d_img.free()
    image = pplan.build_image(data)
    d_img = DeviceBuffer(image.nbytes)
    d_img.upload(image, stream=stream)
    stream.synchronize()
    del image
    pms = time_launches(lambda: peng.match_plan_image(pplan, d_img.ptr, d_img.nbytes, d_f.ptr, d_a.ptr, stream),
                        stream, max(3, args.steps))
    perr = pplan.error(stream)
    log(f"look-back scan over the chunk image, sum of symbols mod 15 (never synchronises; "
        f"{pplan.chunk_len >> 10} KiB chunks): {pms:.4f} ms per launch, "
        f"{data.nbytes / (pms * 1e-3) / 1e9:.1f} Gchars/s, error word {perr}")
    out["permutation_automaton"] = {"automaton": "sum of (byte & 15) mod 15, 16 states with the dead state",
                                    "chunk_len": pplan.chunk_len, "kernel_avg_ms": round(pms, 4),
                                    "value": round(data.nbytes / (pms * 1e-3) / 1e9, 1), "error_word": int(perr),
                                    "verified_bit_exact": None}
    if not args.no_verify:
        from oracle import cpu as ocpu
        rf, ra = ocpu.dfa_walk(perm.delta, perm.accept, perm.start, data, offsets)
        same = bool(np.array_equal(d_f.download(np.uint8, n_long), rf)
                    and np.array_equal(d_a.download(np.uint8, n_long), ra) and perr == 0)
        out["permutation_automaton"]["verified_bit_exact"] = same
        log(f"look-back scan (mod 15): bit-exact={same}")
        if not same:
            log("PARITY FAILURE (look-back scan, mod 15)")
    peng.close()
    d_img.free()
    pplan.close()
    plan.close()
    batch.free()
    eng.close()
    return out


def host_leg(args) -> dict:
    """OptimizedEngine.match_batch on 10,000 host str strings x 1,000
    characters (10 M), the size the reference quotes its host packing at
    (88 ms per 10 M characters for the latin-1 join and offset loop)."""
    from trex_amd.engine import pack_strings
    from trex_amd.optimized_engine import OptimizedEngine
    from trex_amd.workloads import ALPHABET_32

    rng = np.random.default_rng(SEED + 7)
    codes = rng.integers(0x40, 0x60, size=(10_000, 1_000), dtype=np.uint8)
    strings = [bytes(row).decode("latin-1") for row in codes]
    pattern = "(" + "|".join(ALPHABET_32[:16]) + ")*" + ALPHABET_32[16] + ALPHABET_32[17]
    eng = OptimizedEngine(pattern)
    eng.match_batch(strings[:100])                       # device tables and kernels loaded
    reps = 5
    pack_ms = min(_wall(lambda: pack_strings(strings)) for _ in range(reps))
    data, offsets = pack_strings(strings)
    arrays_ms = min(_wall(lambda: eng.match_arrays(data, offsets)) for _ in range(reps))
    total_ms = min(_wall(lambda: eng.match_batch(strings)) for _ in range(reps))
    res, timing = eng.match_batch_timed(strings)
    verified = None
    if not args.no_verify:
        from oracle.dfa import dfa_simulate
        from trex_amd.regex_to_dfa import compile_regex
        dfa = compile_regex(pattern)
        idx = rng.choice(len(strings), size=200, replace=False)
        verified = all(res[i] == dfa_simulate(dfa, strings[i]) for i in idx)
    out = {"workload": "10,000 str x 1,000 chars through OptimizedEngine.match_batch",
           "pattern": pattern, "backend": eng.config_info["scan_backend"],
           "pack_ms": round(pack_ms, 3), "match_arrays_ms": round(arrays_ms, 3),
           "match_batch_ms": round(total_ms, 3), "kernel_ms": round(timing["kernel_ms"], 4),
           "mchars_s_end_to_end": round(10e6 / (total_ms * 1e-3) / 1e6, 1),
           "reference_pack_ms_quoted": 88.0, "verified": verified}
    log(f"host input: pack {pack_ms:.2f} ms, match_arrays (copies + kernel) {arrays_ms:.2f} ms, "
        f"match_batch end to end {total_ms:.2f} ms (kernel {timing['kernel_ms']:.4f} ms), verified={verified}")
    return out


def plain_leg(args, stream) -> dict:
    """The reference's CPU path on its own plain DFA workload: a seeded regex
    over abcd (<= 16-state minimal DFA, bytes outside the alphabet -> dead
    state), 1,024 strings x 1,024 bytes; final state and accept bit per
    string. Times the Python restatement of DFA.simulate (one thread), the C
    oracle's byte walk on one thread and on all threads (median of 3 after a
    warm-up, the reference's method), and beside them the GPU on the same
    batch (device-resident kernel time, and end to end from host arrays)."""
    from oracle import cpu as ocpu
    from oracle.dfa import dfa_simulate
    from trex_amd.device import DeviceBuffer
    from trex_amd.engine import DeviceBatch, DfaEngine
    from trex_amd.tables import dfa_byte_table
    from trex_amd.workloads import PLAIN_DFA_SEED, plain_dfa_workload

    pattern, dfa, data, offsets = plain_dfa_workload()
    tab = dfa_byte_table(dfa)
    n = len(offsets) - 1
    chars = int(offsets[-1])
    strings = [bytes(data[offsets[i]:offsets[i + 1]]).decode("latin-1") for i in range(n)]

    def median3(fn):
        fn()
        return float(np.median([_wall(fn) for _ in range(3)]))

    py_res = []
    py_ms = median3(lambda: py_res.__setitem__(slice(None), [dfa_simulate(dfa, s) for s in strings]))
    threads = ocpu.default_threads()
    c1_ms = median3(lambda: ocpu.dfa_walk(tab.delta, tab.accept, tab.start, data, offsets, threads=1))
    cn_ms = median3(lambda: ocpu.dfa_walk(tab.delta, tab.accept, tab.start, data, offsets, threads=threads))
    rf, ra = ocpu.dfa_walk(tab.delta, tab.accept, tab.start, data, offsets)

    eng = DfaEngine(tab)
    batch = DeviceBatch(data, offsets, layout="csr")
    d_f, d_a = DeviceBuffer(n), DeviceBuffer(n)
    kms = time_launches(lambda: eng.match_batch_device(batch, d_f.ptr, d_a.ptr, stream), stream, 20)
    gf, ga = d_f.download(np.uint8, n, stream), d_a.download(np.uint8, n, stream)
    e2e = [eng.match_arrays(data, offsets) for _ in range(4)][1:]
    e2e_ms = float(np.median([x[3] for x in e2e]))
    same = bool(np.array_equal(gf, rf) and np.array_equal(ga, ra) and np.array_equal(e2e[-1][0], rf)
                and [bool(x) for x in ra] == py_res)
    for b in (d_f, d_a):
        b.free()
    batch.free()
    eng.close()
    out = {"workload": f"plain DFA, seed {PLAIN_DFA_SEED}: regex over abcd, {tab.n_states}-state table, "
                       f"{n} x {chars // n} B",
           "pattern": pattern, "accept_rate": round(float(ra.mean()), 4),
           "python_dfa_simulate_mchars_s": round(chars / (py_ms * 1e-3) / 1e6, 2),
           "c_oracle_1_thread_mchars_s": round(chars / (c1_ms * 1e-3) / 1e6, 1),
           "c_oracle_threads": threads,
           "c_oracle_all_threads_mchars_s": round(chars / (cn_ms * 1e-3) / 1e6, 1),
           "cpu": cpu_model(), "os_cpu_count": os.cpu_count(),
           "gpu_kernel_ms": round(kms, 4), "gpu_kernel_gchars_s": round(chars / (kms * 1e-3) / 1e9, 2),
           "gpu_end_to_end_ms": round(e2e_ms, 4),
           "reference_dfa_simulate_mchars_s_quoted": [18.4, 20.8],
           "verified_bit_exact": same}
    log(f"plain DFA (CPU path): Python DFA.simulate {out['python_dfa_simulate_mchars_s']} Mchar/s, C oracle "
        f"{out['c_oracle_1_thread_mchars_s']} Mchar/s (1 thread) / {out['c_oracle_all_threads_mchars_s']} Mchar/s "
        f"({threads} threads); GPU kernel {kms:.4f} ms, end to end {e2e_ms:.3f} ms; bit-exact={same}")
    return out


def _wall(fn) -> float:
    t0 = time.perf_counter()
    fn()
    return (time.perf_counter() - t0) * 1e3


def multi_leg(args, engines, batch, stream, per_pattern_ms: float, nbytes: int):
    """The same 8 patterns over the same batch in one pass (trex_dfa_multi):
    one read of the bytes advances all patterns. Returns its summary and its
    output buffers (checked against the oracle with the per-pattern ones)."""
    from trex_amd.device import DeviceBuffer, Event
    from trex_amd.engine import DfaMultiEngine

    if any(e.info()["window_base"] < 0 for e in engines):
        log("multi-pattern pass skipped: a pattern has no 32-byte window")
        return None, None
    n = batch.n_strings
    P = len(engines)
    multi = DfaMultiEngine(engines)
    d_f, d_a = DeviceBuffer(P * n), DeviceBuffer(P * n)

    def timed(reps: int) -> list:
        launch = lambda: multi.match_batch_device(batch, d_f.ptr, d_a.ptr, stream)
        for _ in range(args.warmup):
            launch()
        evs = [(Event(), Event()) for _ in range(reps)]
        for a, b in evs:
            a.record(stream)
            launch()
            b.record(stream)
        stream.synchronize()
        return [a.elapsed_ms(b) for a, b in evs]

    if args.multi_sweep:
        for form in ("k2f", "k2fp", "k2fc", "k2w"):
            multi.set_form(form)
            for group in (1, 2, 3, 4, 5, 6, 7, 8):
                multi.set_group(group)
                ms = float(np.median(timed(max(3, args.steps))))
                log(f"multi sweep: form {form} group {group}: {ms:.4f} ms per 8 patterns, "
                    f"{P * nbytes / (ms * 1e-3) / 1e9:.0f} Gchars/s, {per_pattern_ms / ms:.2f}x per-pattern launches")
        multi.close()
        multi = DfaMultiEngine(engines)   # back to the library's default form and group
    ms = timed(max(3, args.steps))
    avg = float(np.mean(ms))
    info = multi.info()
    alg_bytes = nbytes + 16 * n + 2 * P * n
    launches = -(-P // info["group"])
    per_launch = -(-P // launches)
    il_tag = "_il" if batch.layout == "interleaved" else ""
    if info["form"] == "k2fc":
        small = sum(1 for e in engines if e.table.n_states <= 4)
        kname = f"dfa16_multi_k2fc_f{P - small}_b{small}{il_tag}" if launches == 1 else f"dfa16_multi_k2fc_*{il_tag}"
    else:
        kname = f"dfa16_multi_{info['form']}_p{per_launch}{il_tag}"
    res = {"kernel": kname, "form": info["form"],
           "patterns_of_at_most_4_states": sum(1 for e in engines if e.table.n_states <= 4),
           "patterns_per_launch": per_launch, "launches_per_step": launches,
           "ms_per_step": round(avg, 4), "ms_min": round(min(ms), 4), "ms_max": round(max(ms), 4),
           "gchars_s": round(P * nbytes / (avg * 1e-3) / 1e9, 1),
           "speedup_vs_per_pattern_launches": round(per_pattern_ms / avg, 2),
           "hbm_gbs": round(alg_bytes / (avg * 1e-3) / 1e9, 1),
           "hbm_frac": round(alg_bytes / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    log(f"multi-pattern pass ({res['kernel']}): {avg:.4f} ms for {P} patterns = {res['gchars_s']} Gchars/s, "
        f"{res['speedup_vs_per_pattern_launches']}x the per-pattern launches, {res['hbm_gbs']} GB/s of HBM")
    # the paired form in launches of four (the default before the compact
    # tables), for comparison; its results are checked against the default's
    if info["form"] == "k2fc":
        multi.set_form("k2fp")
        multi.set_group(4)
        d_f2, d_a2 = DeviceBuffer(P * n), DeviceBuffer(P * n)
        cms = []
        for _ in range(args.warmup):
            multi.match_batch_device(batch, d_f2.ptr, d_a2.ptr, stream)
        for _ in range(max(3, args.steps)):
            a, b = Event(), Event()
            a.record(stream)
            multi.match_batch_device(batch, d_f2.ptr, d_a2.ptr, stream)
            b.record(stream)
            stream.synchronize()
            cms.append(a.elapsed_ms(b))
        cavg = float(np.mean(cms))
        same = bool(np.array_equal(d_f2.download(np.uint8, P * n, stream), d_f.download(np.uint8, P * n, stream))
                    and np.array_equal(d_a2.download(np.uint8, P * n, stream), d_a.download(np.uint8, P * n, stream)))
        res["paired_form"] = {"form": "k2fp", "patterns_per_launch": 4, "ms_per_step": round(cavg, 4),
                              "gchars_s": round(P * nbytes / (cavg * 1e-3) / 1e9, 1),
                              "speedup_vs_per_pattern_launches": round(per_pattern_ms / cavg, 2),
                              "same_results_as_default_form": same}
        log(f"multi-pattern pass, paired tables in launches of four: {cavg:.4f} ms = "
            f"{res['paired_form']['gchars_s']} Gchars/s; same results as the compact form: {same}")
        d_f2.free()
        d_a2.free()
    multi.close()
    return res, (d_f, d_a)


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

    def timed(variant: str) -> list:
        eng.set_variant(variant)
        for _ in range(args.warmup):
            launch()
        evs = [(Event(), Event()) for _ in range(reps)]
        for a, b in evs:
            a.record(stream)
            launch()
            b.record(stream)
        stream.synchronize()
        return [a.elapsed_ms(b) for a, b in evs]

    default = eng.info()["variant"]
    per_variant = {v: timed(v) for v in ("window", "window5", "window6")}
    for v, t in per_variant.items():
        log(f"nfa64 ({v}): kernel avg {np.mean(t):.4f} ms min {min(t):.4f} max {max(t):.4f}")
    eng.set_variant(default)
    launch()                                   # the default variant's results are the ones checked
    stream.synchronize()
    ms = per_variant[default]
    avg = float(np.mean(ms))
    alg_bytes = data.nbytes + 16 * n + 9 * n
    gchars = data.nbytes / (avg * 1e-3) / 1e9
    achieved = alg_bytes / (avg * 1e-3) / 1e9
    log(f"nfa64 ({default}): kernel avg {avg:.4f} ms min {min(ms):.4f} max {max(ms):.4f}, "
        f"{gchars:.1f} Gchars/s, {achieved:.0f} GB/s")
    out = {"workload": "nfa64_bitpar_1GiB", "variant": default, "strings": n, "length": LENGTH,
           "value": round(gchars, 2), "unit": "Gchars/s", "kernel_avg_ms": round(avg, 5),
           "variants_ms": {v: round(float(np.mean(t)), 5) for v, t in per_variant.items()},
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "alg_bytes_per_launch": alg_bytes,
                        "note": "LDS table lookups (12 per byte with 6/5-bit slices, 13 with 5-bit, 16 with nibbles) bound "
                                "the kernel before HBM does"},
           "verified_bit_exact": None, "cpu_baseline": None}
    reads = {"window6": 12, "window5": 13, "window": 16}.get(default)
    if reads:
        # the bound that holds: LDS bytes per character against the LDS's
        # 256 B per clock per CU (256 CUs at 2.4 GHz)
        lds_tbs = reads * 8 * gchars / 1e3
        out["lds_roofline"] = {"reads_per_char": reads, "bytes_per_read": 8, "achieved": round(lds_tbs, 1),
                               "peak": LDS_PEAK_TBS, "unit": "TB/s", "frac": round(lds_tbs / LDS_PEAK_TBS, 4),
                               "note": "bank conflicts add about 2.1 cycles to each 2-cycle wave read "
                                       "(rocprof SQ_LDS_BANK_CONFLICT / SQ_LDS_IDX_ACTIVE)"}
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
    default_mode = eng.info()["mode"]
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
    head = per_mode.get(default_mode) or next(iter(per_mode.values()))
    head_ms = head["kernel_avg_ms"]
    achieved = flops / (head_ms * 1e-3) / 1e12
    d_sym.free()
    d_ll.free()
    d_lm.free()
    eng.set_mode(default_mode)
    eng.set_waves(8)
    long_seq = prob_long_leg(args, eng, T, start, accept, stream)
    if long_seq.get("within_tolerance") is not None:
        within = bool(within is not False and long_seq["within_tolerance"])
    eng.close()
    return {"workload": "prob_nfa64_dirichlet0.1", "sequences": PROB_SEQ, "length": PROB_LEN, "symbols": 32,
            "default_mode": default_mode, "long_sequences": long_seq,
            "value": head["gchars_s"], "unit": "Gchars/s", "dtype": "f16 MFMA, f32 accumulate, f64 log-scale",
            "roofline": {"bound": "mfma", "achieved": round(achieved, 2), "peak": MFMA_F16_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(achieved / MFMA_F16_PEAK_TFLOPS, 4),
                         "kernel": head["kernel"], "alg_flops_per_launch": flops,
                         "note": "algorithmic FLOPs = 2*64^2 per symbol (one matvec); the split mode issues 2x"},
            "modes": per_mode, "cpu_baseline": cpu, "within_tolerance": within}


PROB_LONG = (4, 1 << 18)            # sequences x symbols of the parallel-in-time comparison


def prob_long_leg(args, eng, T, start, accept, stream) -> dict:
    """Few long sequences: the batch kernel walks each one step after step;
    the parallel-in-time path multiplies chunks concurrently and chains them."""
    from trex_amd.device import DeviceBuffer, Event
    n, L = PROB_LONG
    rng = np.random.default_rng(SEED + 5)
    sym = rng.integers(0, 32, size=(n, L), dtype=np.uint8)
    d_sym = DeviceBuffer.from_array(sym)
    begin = np.arange(n, dtype=np.int64) * L
    end = begin + L
    d_b, d_e = DeviceBuffer.from_array(begin), DeviceBuffer.from_array(end)
    d_ll, d_lm = DeviceBuffer(8 * n), DeviceBuffer(8 * n)
    out = {"sequences": n, "length": L, "mode": eng.info()["mode"]}
    runs = {"sequential": lambda: eng.loglik_ranges_device(d_sym.ptr, d_b.ptr, d_e.ptr, n, d_ll.ptr, d_lm.ptr, stream),
            "chunked": lambda: eng.loglik_chunked_device(d_sym.ptr, begin, end, d_ll.ptr, d_lm.ptr, 0, stream)}
    lls = {}
    for name, fn in runs.items():
        fn()
        a, b = Event(), Event()
        t0 = time.perf_counter()
        a.record(stream)
        fn()
        b.record(stream)
        stream.synchronize()
        out[f"{name}_ms"] = round(a.elapsed_ms(b), 3)
        out[f"{name}_wall_ms"] = round((time.perf_counter() - t0) * 1e3, 3)
        lls[name] = d_ll.download(np.float64, n, stream)
    out["speedup"] = round(out["sequential_ms"] / out["chunked_ms"], 1)
    log(f"prob long: {n} x {L} symbols, sequential {out['sequential_ms']:.2f} ms, chunked {out['chunked_ms']:.2f} ms "
        f"({out['speedup']}x)")
    if not args.no_verify:
        from oracle import cpu as ocpu
        ref, _ = ocpu.pnfa_forward(T, start, accept, sym, threads=n)
        for name, ll in lls.items():
            err = np.abs(ll - ref) / np.abs(ref)
            out[f"{name}_rel_ll_err_max"] = float(f"{err.max():.3e}")
        out["within_tolerance"] = bool(out["sequential_rel_ll_err_max"] <= 1e-3 and out["chunked_rel_ll_err_max"] <= 1e-3)
        log(f"prob long: rel LL err vs FP64 sequential {out['sequential_rel_ll_err_max']:.2e}, "
            f"chunked {out['chunked_rel_ll_err_max']:.2e}")
    for buf in (d_sym, d_b, d_e, d_ll, d_lm):
        buf.free()
    return out


MIX_STRINGS = 65_536
MIX_PATTERNS = 64                  # 32 DFA-16 + 32 Boolean NFA-64
MIX_LEN = (32, 1 << 24)            # log-uniform string lengths
MIX_SUBSET = (0, 1, 2, 3, 32, 33, 34, 35)   # patterns of the batched-vs-plain comparison
MIX_GEN_SEED = 0x5EED03


def mix_leg(args, stream) -> dict:
    """Uneven mix: one persistent launch vs one launch per pattern."""
    from trex_amd import _lib
    from trex_amd.autotune import tune_mix
    from trex_amd.device import DeviceBuffer, Event
    from trex_amd.engine import ChunkPlan, DfaEngine, NfaEngine, lookback_chunk_len, lookback_target_chunks
    from trex_amd.mix import MixEngine, MixParams
    from trex_amd.tables import dfa_byte_table
    from trex_amd.workloads import ALPHABET_32, random_dfa_pattern, random_nfa64

    rng = np.random.default_rng(SEED + 3)
    dfa_tables = [dfa_byte_table(random_dfa_pattern(rng, ALPHABET_32)[1]) for _ in range(MIX_PATTERNS // 2)]
    nfas = [random_nfa64(rng) for _ in range(MIX_PATTERNS // 2)]
    engines = [DfaEngine(t) for t in dfa_tables] + [NfaEngine.from_matrices(nm) for nm in nfas]
    n_strings = args.mix_strings
    lengths = np.floor(np.exp(rng.uniform(np.log(MIX_LEN[0]), np.log(MIX_LEN[1]), size=n_strings))).astype(np.int64)
    pattern_of = rng.integers(0, MIX_PATTERNS, size=n_strings).astype(np.int32)
    offsets = np.zeros(n_strings + 1, np.int64)
    np.cumsum(lengths, out=offsets[1:])
    total = int(offsets[-1])
    t0 = time.perf_counter()
    d_bytes = DeviceBuffer(-(-total // 16) * 16)
    _lib.call("trex_gen_symbols", d_bytes.ptr, d_bytes.nbytes, MIX_GEN_SEED, 0x40, 32, stream.handle)
    stream.synchronize()
    log(f"mix: {n_strings} strings, {total / 1e9:.1f} GB generated in HBM in {time.perf_counter() - t0:.2f}s")
    n = n_strings
    d_mask, d_acc = DeviceBuffer(8 * n), DeviceBuffer(n)
    begin, end = offsets[:-1], offsets[1:]

    def timed(fn, reps):
        fn()
        stream.synchronize()
        evs = [(Event(), Event()) for _ in range(reps)]
        for a, b in evs:
            a.record(stream)
            fn()
            b.record(stream)
        stream.synchronize()
        return [a.elapsed_ms(b) for a, b in evs]

    # autotuned plan on a sample of the batch, then the full batch
    if args.mix_tune:
        params, tuning = tune_mix(engines, begin, end, pattern_of, d_bytes.ptr, stream, log=log)
    else:
        params, tuning = MixParams(), {"source": "defaults"}
    log(f"mix: parameters {params} ({tuning.get('source')})")
    mix = MixEngine(engines)
    t0 = time.perf_counter()
    info = mix.plan(begin, end, pattern_of, params)
    log(f"mix: planned {info} in {time.perf_counter() - t0:.2f}s")
    t0 = time.perf_counter()
    mix.run(d_bytes.ptr, d_mask.ptr, d_acc.ptr, stream, args.mix_max_spins)
    stream.synchronize()
    log(f"mix: first run {time.perf_counter() - t0:.3f}s, error word {mix.error(stream)}")
    ms = timed(lambda: mix.run(d_bytes.ptr, d_mask.ptr, d_acc.ptr, stream, args.mix_max_spins),
               max(2, args.steps // 10))
    err = mix.error(stream)
    mix.capture(d_bytes.ptr, d_mask.ptr, d_acc.ptr, stream)
    ms_graph = timed(lambda: mix.replay(stream), max(2, args.steps // 10))
    full_mask = d_mask.download(np.uint64, n, stream)
    full_acc = d_acc.download(np.uint8, n, stream)
    avg = float(np.median(ms))
    gchars = total / (avg * 1e-3) / 1e9
    log(f"mix: one launch {avg:.2f} ms (graph replay {np.median(ms_graph):.2f} ms), {gchars:.1f} Gchars/s, "
        f"plan {info}, error word {err}")

    # the same plan parameters with the other NFA table form
    other = "nibble" if params.form == "slice" else "slice"
    mix_o = MixEngine(engines)
    mix_o.plan(begin, end, pattern_of, MixParams(**{**params.__dict__, "form": other}))
    mix_o.run(d_bytes.ptr, d_mask.ptr, d_acc.ptr, stream, args.mix_max_spins)
    ms_o = timed(lambda: mix_o.run(d_bytes.ptr, d_mask.ptr, d_acc.ptr, stream, args.mix_max_spins),
                max(2, args.steps // 10))
    err_o = mix_o.error(stream)
    same_o = bool(np.array_equal(d_mask.download(np.uint64, n, stream), full_mask)
                 and np.array_equal(d_acc.download(np.uint8, n, stream), full_acc))
    avg_o = float(np.median(ms_o))
    log(f"mix ({other} NFA tables): one launch {avg_o:.2f} ms, {total / (avg_o * 1e-3) / 1e9:.1f} Gchars/s, "
        f"error word {err_o}, results equal to the tuned form's: {same_o}")
    other_res = {"form": other, "kernel_ms": round(avg_o, 3), "gchars_s": round(total / (avg_o * 1e-3) / 1e9, 1),
              "error_word": int(err_o), "same_results": same_o}
    del mix_o

    # batched vs plain on the subset. Plain = one launch per pattern of that
    # pattern's best single-pattern form: the look-back plan for a DFA (chunk
    # length picked for the table and the bytes), the persistent queue with
    # that one NFA for an NFA (speculative chunks and fix-ups for its long
    # strings). The lane-per-string NFA launch is timed beside it for reference.
    sub = np.flatnonzero(np.isin(pattern_of, MIX_SUBSET))
    sub_bytes = int(lengths[sub].sum())
    sub_mix = MixEngine(engines)
    sub_mix.plan(begin[sub], end[sub], pattern_of[sub], params)
    d_smask, d_sacc = DeviceBuffer(8 * len(sub)), DeviceBuffer(len(sub))
    ms_sub = float(np.median(timed(lambda: sub_mix.run(d_bytes.ptr, d_smask.ptr, d_sacc.ptr, stream), 3)))
    plain_ms = 0.0
    lane_ms = 0.0
    plain_mask = np.zeros(len(sub), np.uint64)
    plain_acc = np.zeros(len(sub), np.uint8)
    per_pattern = []
    for p in MIX_SUBSET:
        idx = np.flatnonzero(pattern_of[sub] == p)
        sb, se = begin[sub][idx], end[sub][idx]
        d_pm, d_pa = DeviceBuffer(8 * len(idx)), DeviceBuffer(len(idx))
        e = engines[p]
        if isinstance(e, DfaEngine):
            tab = dfa_tables[p]
            plan = ChunkPlan(sb, se, lookback_chunk_len(int((se - sb).sum()), lookback_target_chunks(tab)))
            d_state = DeviceBuffer(len(idx))
            t = float(np.median(timed(lambda: e.match_plan(plan, d_bytes.ptr, d_state.ptr, d_pa.ptr, stream), 3)))
            plain_mask[idx] = d_state.download(np.uint8, len(idx), stream)
            form = f"look-back plan, {plan.chunk_len} B chunks"
            d_state.free()
            plan.close()
        else:
            one = MixEngine([e])
            one.plan(sb, se, np.zeros(len(idx), np.int32), params)
            t = float(np.median(timed(lambda: one.run(d_bytes.ptr, d_pm.ptr, d_pa.ptr, stream), 3)))
            plain_mask[idx] = d_pm.download(np.uint64, len(idx), stream)
            form = "persistent queue with this NFA alone"
            del one
            d_b, d_e = DeviceBuffer.from_array(sb, stream), DeviceBuffer.from_array(se, stream)
            lane_ms += timed(lambda: _lib.call("trex_nfa_match_ranges", e.handle, d_bytes.ptr, d_b.ptr, d_e.ptr,
                                               len(idx), d_pm.ptr, d_pa.ptr, stream.handle), 1)[0]
            d_b.free()
            d_e.free()
        plain_acc[idx] = d_pa.download(np.uint8, len(idx), stream)
        plain_ms += t
        per_pattern.append({"pattern": p, "strings": len(idx), "bytes": int((se - sb).sum()), "ms": round(t, 3),
                            "form": form})
        log(f"mix plain: pattern {p} ({type(e).__name__}) {len(idx)} strings, {t:.2f} ms ({form})")
        d_pm.free()
        d_pa.free()
    sub_mask = d_smask.download(np.uint64, len(sub), stream)
    sub_acc = d_sacc.download(np.uint8, len(sub), stream)
    same_sub = bool(np.array_equal(sub_mask, plain_mask) and np.array_equal(sub_acc, plain_acc)
                    and np.array_equal(full_mask[sub], plain_mask) and np.array_equal(full_acc[sub], plain_acc))
    ratio = plain_ms / ms_sub
    log(f"mix subset {MIX_SUBSET}: {sub_bytes / 1e9:.2f} GB, batched {ms_sub:.2f} ms vs plain {plain_ms:.2f} ms "
        f"= {ratio:.2f}x; batched == plain: {same_sub}; NFA lane-per-string launches {lane_ms:.1f} ms")

    verified = None
    cpu = None
    if not args.no_verify:
        from oracle import cpu as ocpu
        from trex_amd.engine import nfa_byte_map
        vr = np.random.default_rng(7)
        short = np.flatnonzero(lengths <= 1 << 16)
        long_ = np.flatnonzero(lengths > 1 << 20)
        pick = np.concatenate([vr.choice(short, min(4096, len(short)), replace=False),
                               vr.choice(long_, min(24, len(long_)), replace=False)])
        ok = True
        t_cpu = 0.0
        n_checked = 0
        for p in range(MIX_PATTERNS):
            idx = pick[pattern_of[pick] == p]
            if len(idx) == 0:
                continue
            sub_off = np.zeros(len(idx) + 1, np.int64)
            np.cumsum(lengths[idx], out=sub_off[1:])
            data = np.concatenate([ocpu.gen_symbols(begin[i], lengths[i], MIX_GEN_SEED) for i in idx])
            t0 = time.perf_counter()
            if p < MIX_PATTERNS // 2:
                t = dfa_tables[p]
                rf, ra = ocpu.dfa_walk(t.delta, t.accept, t.start, data, sub_off)
            else:
                nm = nfas[p - MIX_PATTERNS // 2]
                rf, ra = ocpu.nfa_walk(nm.successor_masks(), nfa_byte_map(nm, "reject"), nm.start_mask(),
                                       nm.accept_bits(), data, sub_off)
            t_cpu += time.perf_counter() - t0
            n_checked += int(sub_off[-1])
            ok &= bool(np.array_equal(full_mask[idx], rf.astype(np.uint64)) and np.array_equal(full_acc[idx], ra))
        verified = ok and same_sub and err == 0 and other_res["same_results"] and other_res["error_word"] == 0
        cpu = {"value": round(n_checked / t_cpu / 1e9, 3), "unit": "Gchars/s", "cores": ocpu.default_threads(),
               "kind": "port", "sample": f"{len(pick)} strings ({n_checked / 1e9:.2f} GB: 4096 <= 64 KiB and 24 "
                                         f"> 1 MiB) through the C oracle walks, {t_cpu:.2f} s"}
        log(f"mix: oracle check of {len(pick)} strings: {ok}")
    for b in (d_bytes, d_mask, d_acc, d_smask, d_sacc):
        b.free()
    return {"workload": "uneven_mix_64_patterns", "strings": n, "bytes": total, "patterns": "32 DFA-16 + 32 NFA-64",
            "value": round(gchars, 2), "unit": "Gchars/s", "kernel_ms": round(avg, 3),
            "graph_replay_ms": round(float(np.median(ms_graph)), 3), "other_form": other_res,
            "plan": info, "params": params.__dict__,
            "tuning": tuning,
            "roofline": {"bound": "hbm (DFA strings) / lds (NFA strings)",
                         "achieved": round(total / (avg * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(total / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
            "batched_vs_plain": {"patterns": list(MIX_SUBSET), "bytes": sub_bytes, "batched_ms": round(ms_sub, 3),
                                 "plain_ms": round(plain_ms, 3), "speedup": round(ratio, 2),
                                 "plain": "one launch per pattern of its best single-pattern form: look-back plan "
                                          "(DFA), persistent queue with the one NFA (NFA)",
                                 "per_pattern": per_pattern,
                                 "nfa_lane_per_string_ms": round(lane_ms, 3)},
            "verified_bit_exact": verified, "cpu_baseline": cpu}


SERVE_PATTERNS = (77, 38, 13)      # DFA-16, Boolean NFA-64, probabilistic NFA-64
SERVE_POOL = 3 << 30               # host bytes each request's strings are sliced from
SERVE_MEDIAN, SERVE_SIGMA, SERVE_CAP = 4096, 2.0, 1 << 24


def serve_leg(args, stream) -> dict:
    """Continuous batching: Poisson arrivals into the server, latency per
    request; a saturated burst served vs one launch per request."""
    from trex_amd.engine import (ChunkPlan, DfaEngine, NfaEngine, ProbNfaEngine, lookback_chunk_len,
                                 lookback_target_chunks)
    from trex_amd.serve import Request, Server
    from trex_amd.tables import dfa_byte_table
    from trex_amd.workloads import ALPHABET_32, dirichlet_pnfa, random_dfa_pattern, random_nfa64

    rng = np.random.default_rng(SEED + 4)
    n_dfa, n_nfa, n_prob = SERVE_PATTERNS
    engines = [DfaEngine(dfa_byte_table(random_dfa_pattern(rng, ALPHABET_32)[1])) for _ in range(n_dfa)]
    engines += [NfaEngine.from_matrices(random_nfa64(rng)) for _ in range(n_nfa)]
    engines += [ProbNfaEngine(*dirichlet_pnfa(rng, 32)) for _ in range(n_prob)]
    t0 = time.perf_counter()
    from trex_amd.device import PinnedBuffer
    pinned, pinned_sym = PinnedBuffer(SERVE_POOL), PinnedBuffer(SERVE_POOL)   # requests arrive in pinned memory
    pool, pool_sym = pinned.array, pinned_sym.array
    pool[:] = rng.integers(0x40, 0x60, size=SERVE_POOL, dtype=np.uint8)
    np.bitwise_and(pool, np.uint8(31), out=pool_sym)
    log(f"serve: {len(engines)} resident patterns, {SERVE_POOL >> 30} GiB pinned request pool in "
        f"{time.perf_counter() - t0:.1f}s")

    def make_request(r: np.random.Generator) -> Request:
        p = int(r.integers(0, len(engines)))
        while True:
            n = int(r.integers(1, 4097))
            cap = args.serve_prob_cap if isinstance(engines[p], ProbNfaEngine) else SERVE_CAP
            lengths = np.minimum(np.floor(np.exp(r.normal(np.log(SERVE_MEDIAN), SERVE_SIGMA, size=n))),
                                 cap).astype(np.int64)
            total = int(lengths.sum())
            if total <= SERVE_POOL:
                break
        off = np.zeros(n + 1, np.int64)
        np.cumsum(lengths, out=off[1:])
        start = int(r.integers(0, SERVE_POOL - total + 1))
        src = pool_sym if isinstance(engines[p], ProbNfaEngine) else pool
        return Request(p, src[start:start + total], off)

    def pct(x, q):
        return round(float(np.percentile(x, q)), 3) if len(x) else None

    def make_small_request(r: np.random.Generator) -> Request:
        """A request of at most 1 MiB: 1-64 strings, the same length law."""
        p = int(r.integers(0, len(engines)))
        while True:
            n = int(r.integers(1, 65))
            lengths = np.floor(np.exp(r.normal(np.log(SERVE_MEDIAN), SERVE_SIGMA, size=n))).astype(np.int64)
            total = int(lengths.sum())
            if total <= 1 << 20:
                break
        off = np.zeros(n + 1, np.int64)
        np.cumsum(lengths, out=off[1:])
        start = int(r.integers(0, SERVE_POOL - total + 1))
        src = pool_sym if isinstance(engines[p], ProbNfaEngine) else pool
        return Request(p, src[start:start + total], off)

    qps, duration = args.serve_qps, args.serve_seconds
    arr_rng = np.random.default_rng(11)
    n_req = int(qps * duration)
    arrive = np.cumsum(arr_rng.exponential(1.0 / qps, size=n_req))
    reqs = [make_request(arr_rng) for _ in range(n_req)]
    # a second, independent Poisson stream of small requests, so that the
    # latency tail of requests <= 1 MiB rests on hundreds of samples
    small_qps = args.serve_small_qps
    sm_rng = np.random.default_rng(13)
    n_small = int(small_qps * duration)
    arrive_small = np.cumsum(sm_rng.exponential(1.0 / small_qps, size=n_small)) if n_small else np.zeros(0)
    small_reqs = [make_small_request(sm_rng) for _ in range(n_small)]
    schedule = sorted([(t, r) for t, r in zip(arrive, reqs)] + [(t, r) for t, r in zip(arrive_small, small_reqs)],
                      key=lambda x: x[0])
    server = Server(engines, hbm_budget=args.serve_hbm_gb << 30, mix_form=args.serve_mix_form)
    t_start = time.perf_counter()
    with server:
        for t_due, r in schedule:
            due = t_start + t_due
            while True:
                now = time.perf_counter()
                if now >= due:
                    break
                time.sleep(min(due - now, 0.0005))
            server.submit(r)
        unfinished = sum(not r.wait(max(0.0, t_start + duration + 120 - time.perf_counter()))
                         for _, r in schedule)
    wall = time.perf_counter() - t_start
    if unfinished:
        log(f"serve: {unfinished} requests unfinished after the time limit")
        reqs = [r for r in reqs if r.done]
        small_reqs = [r for r in small_reqs if r.done]
    lat_s = np.array([r.latency_ms for r in small_reqs])
    kind_s = np.array(["prob" if isinstance(engines[r.pattern], ProbNfaEngine) else "dfa_nfa" for r in small_reqs])
    small_stream = {"qps": small_qps, "requests": len(small_reqs),
                    "bytes": int(sum(r.nbytes for r in small_reqs)),
                    "p50_ms": pct(lat_s, 50), "p99_ms": pct(lat_s, 99), "max_ms": pct(lat_s, 100),
                    "dfa_nfa": {"requests": int((kind_s == "dfa_nfa").sum()),
                                "p50_ms": pct(lat_s[kind_s == "dfa_nfa"], 50),
                                "p99_ms": pct(lat_s[kind_s == "dfa_nfa"], 99)},
                    "prob": {"requests": int((kind_s == "prob").sum()), "p50_ms": pct(lat_s[kind_s == "prob"], 50),
                             "p99_ms": pct(lat_s[kind_s == "prob"], 99)}}
    for kind in ("dfa_nfa", "prob"):   # where the time goes: waiting for a slot, or running
        sel = kind_s == kind
        for name, a_, b_ in (("wait_slot", "admitted", "launched"), ("run", "launched", "done")):
            dt = np.array([(getattr(r, b_) - getattr(r, a_)) * 1e3 for r in small_reqs])[sel] if len(small_reqs) \
                else np.zeros(0)
            small_stream[kind][f"{name}_p50_ms"] = pct(dt, 50)
            small_stream[kind][f"{name}_p99_ms"] = pct(dt, 99)
    log(f"serve: small-request stream {small_stream}")
    lat = np.array([r.latency_ms for r in reqs])
    small = np.array([r.nbytes <= 1 << 20 for r in reqs])
    kinds = np.array(["prob" if isinstance(engines[r.pattern], ProbNfaEngine) else "dfa_nfa" for r in reqs])
    total_bytes = sum(r.nbytes for r in reqs)
    by_lane = {}
    for lane in ("small", "bulk", "prob_small", "prob_bulk"):
        sel = np.array([r.lane == lane for r in reqs])
        by_lane[lane] = {"requests": int(sel.sum()), "p50_ms": pct(lat[sel], 50), "p99_ms": pct(lat[sel], 99)}
    sel_dn = small & (kinds == "dfa_nfa")
    phases = {}
    for name, a_, b_ in (("admit", "arrival", "admitted"), ("wait_slot", "admitted", "launched"),
                         ("run", "launched", "done")):
        dt = np.array([(getattr(r, b_) - getattr(r, a_)) * 1e3 for r in reqs])
        phases[name] = {"p50_ms_le_1MiB": pct(dt[small], 50), "p99_ms_le_1MiB": pct(dt[small], 99),
                        "p50_ms_all": pct(dt, 50)}
    log(f"serve: latency phases {phases}")
    log(f"serve: {n_req} requests at {qps} QPS in {wall:.1f}s, {total_bytes / 1e9:.1f} GB; "
        f"p50 {pct(lat, 50)} ms p99 {pct(lat, 99)} ms; <= 1 MiB: {int(small.sum())} requests p50 "
        f"{pct(lat[small], 50)} p99 {pct(lat[small], 99)} ms (DFA/NFA only: p99 {pct(lat[sel_dn], 99)} ms); "
        f"ticks {dict(server.ticks)}")

    # saturated burst: served vs one launch per request (DFA / NFA requests)
    burst_rng = np.random.default_rng(12)
    burst = []
    while len(burst) < args.serve_burst:
        r = make_request(burst_rng)
        if not isinstance(engines[r.pattern], ProbNfaEngine):
            burst.append(r)
    bb = sum(r.nbytes for r in burst)
    srv = Server(engines, hbm_budget=args.serve_hbm_gb << 30, mix_form=args.serve_mix_form)
    t0 = time.perf_counter()
    with srv:
        for r in burst:
            srv.submit(r)
        for r in burst:
            if not r.wait(300):
                raise RuntimeError("burst request not served within 300 s")
    served_s = time.perf_counter() - t0
    from trex_amd import _lib
    from trex_amd.device import DeviceBuffer
    t0 = time.perf_counter()
    same = True
    for r in burst:   # plain: each request alone, in its pattern's best form, synchronously
        e = engines[r.pattern]
        n = len(r.offsets) - 1
        if isinstance(e, DfaEngine):
            d = DeviceBuffer.from_array(r.data if r.nbytes else np.zeros(16, np.uint8), stream)
            d_m, d_a = DeviceBuffer(n), DeviceBuffer(n)
            plan = ChunkPlan(r.offsets[:-1], r.offsets[1:],
                             lookback_chunk_len(r.nbytes, lookback_target_chunks(e.table)))
            e.match_plan(plan, d.ptr, d_m.ptr, d_a.ptr, stream)
            got = d_m.download(np.uint8, n, stream).astype(np.uint64)
            acc = d_a.download(np.uint8, n, stream)
            plan.close()
            for b in (d, d_m, d_a):
                b.free()
        else:   # strings over 1 MiB through the queue with this NFA alone, shorter ones walked whole
            got, acc, _, _ = e.match_arrays(r.data, r.offsets)
        same &= bool(np.array_equal(got, r.result[0]) and np.array_equal(acc, r.result[1]))
    plain_s = time.perf_counter() - t0
    log(f"serve burst: {len(burst)} DFA/NFA requests, {bb / 1e9:.2f} GB: served {served_s:.2f}s "
        f"({bb / served_s / 1e9:.1f} Gchars/s) vs plain {plain_s:.2f}s ({bb / plain_s / 1e9:.1f} Gchars/s), "
        f"same results {same}")
    return {"workload": "serving_128_patterns", "qps": qps, "requests": n_req, "unfinished": unfinished,
            "seconds": round(wall, 2), "prob_sequence_cap": args.serve_prob_cap,
            "bytes": total_bytes, "hbm_budget_gb": args.serve_hbm_gb,
            "latency_ms": {"all": {"p50": pct(lat, 50), "p99": pct(lat, 99)},
                           "le_1MiB": {"requests": int(small.sum()), "p50": pct(lat[small], 50),
                                       "p99": pct(lat[small], 99)},
                           "le_1MiB_dfa_nfa": {"requests": int(sel_dn.sum()), "p50": pct(lat[sel_dn], 50),
                                               "p99": pct(lat[sel_dn], 99)},
                           "by_lane": by_lane, "phases": phases},
            "served_gchars_s": round(total_bytes / wall / 1e9, 3),
            "burst": {"requests": len(burst), "bytes": bb, "served_s": round(served_s, 3),
                      "plain_s": round(plain_s, 3), "served_gchars_s": round(bb / served_s / 1e9, 2),
                      "plain_gchars_s": round(bb / plain_s / 1e9, 2), "speedup": round(plain_s / served_s, 2),
                      "same_results": same,
                      "plain": "each request alone, synchronously, in its pattern's best form: look-back plan "
                               "(DFA), persistent queue with the one NFA for strings over 1 MiB (NFA)"},
            "small_stream": small_stream}


def sweep(args, data, offsets, batch, engines, outs, stream) -> int:
    """Kernel variants x prefetch x layout."""
    from trex_amd.engine import DeviceBatch
    n = len(offsets) - 1
    layouts = {batch.layout: batch}
    for name in ("csr", "staggered", "interleaved"):
        if name not in layouts:
            layouts[name] = DeviceBatch(data, offsets, layout=name)
    ref = []
    for i, eng in enumerate(engines):
        eng.match_batch_device(batch, outs[i][0].ptr, outs[i][1].ptr, stream)
        ref.append(outs[i][0].download(np.uint8, n, stream))
    for name, b in layouts.items():
        for v, pf in [("k1_16", 16), ("k2_window", 8), ("k2_window", 16), ("k2_window_f", 8), ("k2_window_f", 16),
                      ("k2_window_fr", 8), ("k2_window_fr", 16)]:
            times, same = [], True
            for i, eng in enumerate(engines):
                eng.set_variant(v)
                eng.set_prefetch(pf)
                times.append(time_launches(lambda: eng.match_batch_device(b, outs[i][0].ptr, outs[i][1].ptr, stream),
                                           stream, 3))
                same &= bool(np.array_equal(outs[i][0].download(np.uint8, n, stream), ref[i]))
            ms = float(np.mean(times))
            log(f"sweep {name:11s} {v:12s} pf{pf:2d}: kernel avg {ms:.4f} ms ({min(times):.4f}-{max(times):.4f}) "
                f"= {(data.nbytes + 18 * n) / (ms * 1e-3) / 1e9:.0f} GB/s, same={same}")
    for name, b in layouts.items():
        if b is not batch:
            b.free()
    return 0


if __name__ == "__main__":
    sys.exit(main())
