"""ctypes binding of ``libtrexhip.so`` (declared in ``include/trex_hip.h``).

The library is loaded from this package directory only. If it is missing the
import fails with an error naming the build command: there is no CPU
fallback behind any matching entry point.
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_int64, c_size_t, c_uint8, c_uint32, c_uint64, c_void_p
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libtrexhip.so"
KERNEL_DIR = PKG / "kernels"

TREX_OK = 0
ERR_NAMES = {-1: "TREX_ERR_NO_DEVICE", -2: "TREX_ERR_ARCH", -3: "TREX_ERR_INVALID",
             -4: "TREX_ERR_CAPACITY", -5: "TREX_ERR_HIP", -6: "TREX_ERR_NOT_INIT",
             -7: "TREX_ERR_TIMEOUT"}

u8p = POINTER(c_uint8)
i64p = POINTER(c_int64)
u64p = POINTER(c_uint64)
f32p = POINTER(c_float)
f64p = POINTER(c_double)
intp = POINTER(c_int)
vpp = POINTER(c_void_p)

# name -> (restype, argtypes)
_SIGNATURES = {
    "trex_device_check": (c_int, []),
    "trex_last_error": (c_int, [c_char_p, c_int]),
    "trex_set_wait_bound_us": (c_int, [c_uint32]),
    "trex_device_info": (c_int, [c_char_p, c_int, intp, i64p]),
    "trex_malloc": (c_int, [vpp, c_size_t]),
    "trex_free": (c_int, [c_void_p]),
    "trex_host_alloc": (c_int, [vpp, c_size_t]),
    "trex_host_free": (c_int, [c_void_p]),
    "trex_memcpy_h2d": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p]),
    "trex_memcpy_d2h": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p]),
    "trex_memcpy_d2d": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p]),
    "trex_memcpy_kernel": (c_int, [c_void_p, c_void_p, c_int64, c_void_p]),
    "trex_stream_create": (c_int, [vpp]),
    "trex_stream_destroy": (c_int, [c_void_p]),
    "trex_stream_sync": (c_int, [c_void_p]),
    "trex_device_sync": (c_int, []),
    "trex_event_create": (c_int, [vpp]),
    "trex_event_destroy": (c_int, [c_void_p]),
    "trex_event_record": (c_int, [c_void_p, c_void_p]),
    "trex_event_elapsed_ms": (c_int, [c_void_p, c_void_p, f32p]),
    "trex_event_query": (c_int, [c_void_p]),
    "trex_event_sync": (c_int, [c_void_p]),
    "trex_stream_wait_event": (c_int, [c_void_p, c_void_p]),
    "trex_gen_symbols": (c_int, [c_void_p, c_int64, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, c_void_p]),
    # DFA
    "trex_dfa_create": (c_int, [u8p, u8p, c_int, c_int, vpp]),
    "trex_dfa_destroy": (c_int, [c_void_p]),
    "trex_dfa_info": (c_int, [c_void_p, intp, intp, intp]),
    "trex_dfa_set_variant": (c_int, [c_void_p, c_int]),
    "trex_dfa_set_prefetch": (c_int, [c_void_p, c_int]),
    "trex_dfa_set_rotation": (c_int, [c_void_p, c_int]),
    "trex_dfa_rotation": (c_int, [c_void_p, POINTER(c_int), POINTER(c_int)]),
    "trex_dfa_match_ranges": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p,
                                      c_void_p]),
    "trex_dfa_multi_create": (c_int, [c_void_p, c_int, vpp]),
    "trex_dfa_multi_destroy": (c_int, [c_void_p]),
    "trex_dfa_multi_set_form": (c_int, [c_void_p, c_int]),
    "trex_dfa_multi_set_group": (c_int, [c_void_p, c_int]),
    "trex_dfa_multi_info": (c_int, [c_void_p, intp, intp, intp]),
    "trex_dfa_multi_match_ranges": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p,
                                            c_void_p]),
    "trex_dfa_multi_match_interleaved": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_void_p,
                                                 c_void_p, c_void_p]),
    "trex_dfa_match_interleaved": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p,
                                      c_void_p]),
    # NFA
    "trex_nfa_create": (c_int, [c_void_p, c_int, c_void_p, c_uint64, c_uint64, vpp]),
    "trex_nfa_destroy": (c_int, [c_void_p]),
    "trex_nfa_info": (c_int, [c_void_p, intp, intp]),
    "trex_nfa_set_variant": (c_int, [c_void_p, c_int]),
    "trex_nfa_match": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p]),
    "trex_nfa_match_ranges": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p,
                                      c_void_p]),
    "trex_nfa_match_host": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p, f32p, f32p]),
    "nfa_tc_engine_init": (c_int, [f32p, f32p, f32p, c_int, c_int, c_int, c_int]),
    "nfa_tc_engine_dispatch": (c_int, [u8p, intp, intp, c_int, c_int, f32p, f32p]),
    "nfa_tc_engine_destroy": (None, []),
    # probabilistic NFA
    "trex_pnfa_create": (c_int, [c_void_p, c_int, c_void_p, c_void_p, vpp]),
    "trex_pnfa_destroy": (c_int, [c_void_p]),
    "trex_pnfa_info": (c_int, [c_void_p, intp, intp, intp, intp]),
    "trex_pnfa_set_waves": (c_int, [c_void_p, c_int]),
    "trex_pnfa_set_grid": (c_int, [c_void_p, c_int64]),
    "trex_pnfa_set_mode": (c_int, [c_void_p, c_int]),
    "trex_pnfa_rounded": (c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    "trex_pnfa_loglik": (c_int, [c_void_p, c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p]),
    "trex_pnfa_loglik_ranges": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p,
                                        c_void_p]),
    "trex_pnfa_loglik_chunked": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_void_p, c_void_p,
                                         c_void_p]),
    "trex_pnfa_loglik_host": (c_int, [c_void_p, c_void_p, c_int64, c_int64, c_void_p, c_void_p, f32p, f32p]),
    # mixed batches
    "trex_mix_create": (c_int, [c_void_p, c_void_p, c_int, vpp]),
    "trex_mix_destroy": (c_int, [c_void_p]),
    "trex_mix_plan": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_int64, c_int64, c_int64,
                              c_int]),
    "trex_mix_info": (c_int, [c_void_p, i64p, i64p, i64p, i64p]),
    "trex_mix_layout": (c_int, [c_void_p] + [c_void_p] * 9),
    "trex_mix_run": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_uint32, c_void_p]),
    "trex_mix_status": (c_int, [c_void_p, c_void_p, c_void_p]),
    "trex_mix_error": (c_int, [c_void_p, POINTER(ctypes.c_uint32), c_void_p]),
    "trex_mix_set_form": (c_int, [c_void_p, c_int]),
    "trex_mix_set_grid": (c_int, [c_void_p, c_int64]),
    "trex_mix_stats": (c_int, [c_void_p, POINTER(ctypes.c_uint32), c_void_p]),
    "trex_graph_begin": (c_int, [c_void_p]),
    "trex_graph_end": (c_int, [c_void_p, vpp]),
    "trex_graph_abort": (c_int, [c_void_p]),
    "trex_graph_launch": (c_int, [c_void_p, c_void_p]),
    "trex_graph_destroy": (c_int, [c_void_p]),
    "trex_mix_capture": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_uint32, c_void_p]),
    "trex_mix_replay": (c_int, [c_void_p, c_void_p]),
    # chunk plans
    "trex_plan_create": (c_int, [c_void_p, c_void_p, c_int64, c_int64, vpp]),
    "trex_plan_destroy": (c_int, [c_void_p]),
    "trex_plan_info": (c_int, [c_void_p, i64p, i64p]),
    "trex_dfa_match_plan": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_uint32, c_void_p]),
    "trex_plan_error": (c_int, [c_void_p, POINTER(ctypes.c_uint32), c_void_p]),
    "trex_plan_image_size": (c_int, [c_void_p, i64p]),
    "trex_plan_set_hold": (c_int, [c_void_p, ctypes.c_uint32]),
    "trex_plan_set_chains": (c_int, [c_void_p, c_int]),
    "trex_plan_set_image_gap": (c_int, [c_void_p, c_int64]),
    "trex_plan_build_image": (c_int, [c_void_p, c_void_p, ctypes.c_int64, c_void_p, ctypes.c_int64]),
    # the allocation-free look-back (static device scratch)
    "trex_dfa_match_chunked_static": (c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_int64, c_int64,
                                              c_void_p, c_void_p, c_void_p, ctypes.c_uint32, c_void_p]),
    "trex_dfa_match_plan_image": (c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_int64, c_void_p, c_void_p,
                                          ctypes.c_uint32, c_void_p]),
    "trex_copy": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p]),
    "trex_dfa_match": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p]),
    "trex_dfa_match_host": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p, f32p, f32p]),
    # reference-compatible
    "prefix_engine_device_check": (c_int, []),
    "prefix_engine_init": (c_int, [u8p, u8p, c_int, c_int, c_int, c_int, c_int]),
    "prefix_engine_dispatch": (c_int, [u8p, intp, intp, c_int, c_int, f32p, f32p]),
    "prefix_engine_destroy": (None, []),
    "monoid_batch_engine_device_check": (c_int, []),
    "monoid_batch_engine_init": (c_int, [c_int, c_int, c_int, u8p, u8p, u8p, u8p, c_int, c_int]),
    "monoid_batch_engine_set_kgram": (c_int, [u8p, u8p, c_int, c_int, c_int]),
    "monoid_batch_engine_dispatch": (c_int, [u8p, intp, intp, c_int, c_int, f32p, f32p]),
    "monoid_batch_engine_dispatch_prefix": (c_int, [u8p, intp, intp, c_int, c_int, f32p, f32p]),
    "monoid_batch_engine_destroy": (None, []),
}

_lib = None


class TrexError(RuntimeError):
    def __init__(self, code: int, what: str, message: str):
        super().__init__(f"{what} failed: {ERR_NAMES.get(code, code)}: {message}")
        self.code = code


def lib() -> ctypes.CDLL:
    """The loaded library (loaded on first use)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists() or not list(KERNEL_DIR.glob("*.hsaco")):
            raise ImportError(f"trex_amd native library not built ({LIB_PATH}, {KERNEL_DIR}/*.hsaco); "
                              "run `python -m trex_amd.build`")
        handle = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def exported_symbols() -> list[str]:
    return list(_SIGNATURES)


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    lib().trex_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc: int, what: str) -> None:
    if rc != TREX_OK:
        raise TrexError(rc, what, last_error())


def call(name: str, *args) -> int:
    """Call ``name`` and raise :class:`TrexError` on a non-zero return."""
    rc = getattr(lib(), name)(*args)
    check(rc, name)
    return rc


def device_info() -> dict:
    """Architecture, compute units and HBM bytes of the current device."""
    buf = ctypes.create_string_buffer(128)
    cus, hbm = c_int(), c_int64()
    call("trex_device_info", buf, len(buf), ctypes.byref(cus), ctypes.byref(hbm))
    return {"arch": buf.value.decode(errors="replace"), "cus": cus.value, "hbm_bytes": hbm.value}


def device_name() -> str:
    d = device_info()
    return f"{d['arch'].split(':')[0]}/{d['cus']}cu"


def device_available() -> bool:
    """True iff the library is built and a gfx950 device is visible."""
    try:
        return lib().trex_device_check() == TREX_OK
    except (ImportError, OSError):
        return False
