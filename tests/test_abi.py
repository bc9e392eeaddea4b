"""The C ABI library loads and exports every symbol include/trex_hip.h declares
(no compute calls: this runs without a GPU)."""

import ctypes
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "trex_hip.h"


def declared_functions() -> list[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void)\s+(\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ["trex_device_check", "trex_dfa_create", "trex_dfa_match", "prefix_engine_init",
                 "monoid_batch_engine_dispatch"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    from trex_amd import _lib
    assert _lib.LIB_PATH.exists(), "run python -m trex_amd.build"
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(declared_functions()) <= exported


def test_python_binding_covers_header():
    from trex_amd import _lib
    assert set(declared_functions()) <= set(_lib.exported_symbols())


def test_kernel_module_built_for_gfx950():
    from trex_amd import _lib
    objs = sorted(_lib.KERNEL_DIR.glob("*.hsaco"))
    assert objs, "run python -m trex_amd.build"
    blob = b""
    for o in objs:
        data = o.read_bytes()
        assert data[:4] == b"\x7fELF", o
        assert b"gfx950" in data, o
        blob += data
    for kernel in [b"dfa16_k2w_pf16", b"dfa16_k2f_pf16", b"dfa16_k1", b"dfa_k1", b"copy_u4", b"pnfa64_f16x2",
                   b"pnfa64_bf16x2", b"dfa16_multi_k2f_p4_il", b"dfa16_multi_k2w_p8", b"nfa64_window6",
                   b"nfa64_window5", b"mix_persistent_slice", b"pnfa64_chunk_f16x2", b"pnfa64_combine",
                   b"dfa16_chunks_window", b"dfa16_chunks_bytes", b"dfa16_chunks_window_img",
                   b"dfa16_chunks_bytes_img", b"dfa16_chunks_window_c2", b"dfa16_k2f_rot_pf8",
                   b"dfa16_multi_k2fp_p4_il", b"dfa16_multi_k2fc_f3_b5_il", b"dfa16_multi_k2fc_f0_b8",
                   b"dfa16_multi_k2fc_f8_b0_il", b"chunk_table_count", b"chunk_table_fill"]:
        assert kernel in blob


def test_header_is_plain_c():
    """The boundary header compiles as C99 and as C++ on its own: plain
    pointers and sizes, no HIP or torch types in the signatures."""
    for lang, std in (("c", "-std=c99"), ("c++", "-std=c++17")):
        r = subprocess.run(["gcc" if lang == "c" else "g++", "-x", lang, std, "-fsyntax-only", "-Wall", "-Werror",
                            "-I", str(ROOT / "include"), "-"], input='#include "trex_hip.h"\nint main(void) { return 0; }\n',
                           capture_output=True, text=True)
        assert r.returncode == 0, (lang, r.stderr)


def test_library_loads_without_device():
    """Loading and the error path work on a host without a GPU."""
    from trex_amd import _lib
    lib = _lib.lib()
    rc = lib.trex_device_check()
    assert rc in (0, -1, -2)
    if rc:
        assert _lib.last_error()
        assert not _lib.device_available()
