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

VARIANTS = {"auto": 0, "k1_generic": 1, "k1_16": 2, "k2_window": 3}
VARIANT_NAMES = {v: k for k, v in VARIANTS.items()}


def pack_strings(strings: Iterable) -> tuple[np.ndarray, np.ndarray]:
    """CSR-pack ``str`` (latin-1) / ``bytes`` items: returns (uint8 bytes, int64 offsets)."""
    items = [s.encode("latin-1") if isinstance(s, str) else bytes(s) for s in strings]
    lengths = np.fromiter((len(b) for b in items), dtype=np.int64, count=len(items))
    offsets = np.zeros(len(items) + 1, dtype=np.int64)
    np.cumsum(lengths, out=offsets[1:])
    data = np.frombuffer(b"".join(items), dtype=np.uint8) if items else np.zeros(0, np.uint8)
    return data, offsets


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

    # -- matching ------------------------------------------------------------
    def match_device(self, d_bytes: int, d_offsets: int, n_strings: int, d_final: int, d_accept: int,
                     stream: Stream | None = None) -> None:
        """Asynchronous dispatch on device pointers (inputs resident in HBM)."""
        _lib.call("trex_dfa_match", self.handle, d_bytes, d_offsets, int(n_strings), d_final, d_accept,
                  stream.handle if stream else None)

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


def upload_batch(data: np.ndarray, offsets: np.ndarray) -> tuple[DeviceBuffer, DeviceBuffer]:
    """Copy a CSR batch into device memory (bytes padded to a 16-byte multiple)."""
    data, offsets = _as_csr(data=data, offsets=offsets)
    d_bytes = DeviceBuffer(max(16, -(-len(data) // 16) * 16))
    d_bytes.upload(data)
    d_offsets = DeviceBuffer.from_array(offsets)
    return d_bytes, d_offsets
