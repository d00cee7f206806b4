"""ctypes binding of the HIP library ``libpyjuice_amd.so`` (C ABI declared in
``include/pyjuice_amd.h``).

The library is built in-tree by ``__graft_entry__.build()`` (``pyjuice_amd/_build.py``).
There is no fallback: every device-side operation of the package goes through this
library and raises if it is missing or if no HIP device is present.
"""
from __future__ import annotations

import ctypes
import os
import threading

import torch

__all__ = ["lib", "lib_path", "stream", "ptr", "check", "NativeError", "is_available"]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_NAME = "libpyjuice_amd.so"
_lock = threading.Lock()
_lib = None


class NativeError(RuntimeError):
    pass


def lib_path() -> str:
    return os.path.join(_HERE, _LIB_NAME)


_c_p = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_f32 = ctypes.c_float
_f64 = ctypes.c_double

# name -> argument types (return type is always int32 status, 0 = success)
_SIGNATURES = {
    "pjamd_abi_version": [],
    "pjamd_data_to_vb": [_c_p, _i32, _c_p, _i64, _i64, _c_p],
    "pjamd_input_forward": [_i32, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _i64, _i64, _i64, _i32, _c_p],
    "pjamd_input_backward": [_i32, _c_p, _c_p, _i32, _c_p, _c_p, _c_p, _c_p, _i64, _i64, _i64, _i32, _c_p],
    "pjamd_input_backward_tied": [_i32, _c_p, _c_p, _i32, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _i64, _i32, _i64, _i64, _i32,
                                  _c_p],
    "pjamd_input_em": [_i32, _c_p, _c_p, _i32, _c_p, _c_p, _c_p, _c_p, _i64, _f32, _f32, _i32, _c_p],
    "pjamd_input_backward_em": [_c_p, _c_p, _i32, _c_p, _c_p, _c_p, _c_p, _i64, _i64, _i64, _i32, _c_p, _c_p, _f32,
                                _f32, _i32, _c_p],
    "pjamd_input_sample": [_i32, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _i32, _i64, _i64, _i64, _c_p],
    "pjamd_prod_forward": [_c_p, _c_p, _c_p, _c_p, _i32, _i32, _i32, _i64, _i32, _c_p],
    "pjamd_prod_backward": [_c_p, _c_p, _c_p, _c_p, _i32, _i32, _i32, _i64, _c_p],
    "pjamd_data_mask": [_c_p, _c_p, _i32, _i64, _i64, _c_p],
    "pjamd_input_missing_flows": [_i32, _c_p, _i32, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _i64, _i64, _i64,
                                  _c_p],
    "pjamd_sum_forward": [_c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _i32, _i32, _i32, _i64, _i32, _i32, _c_p],
    "pjamd_sum_backward_ele_split": [_c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _i32, _i32, _i32, _i32,
                                     _i32, _i64, _i32, _i32, _c_p, _i32, _c_p],
    "pjamd_sum_forward_alg": [_c_p] * 7 + [_i32, _i32, _i32, _i64, _i32, _f32, _c_p],
    "pjamd_sum_backward_ele_alg": [_c_p] * 9 + [_i32, _i32, _i32, _i32, _i64, _i32, _f32, _i32, _c_p],
    "pjamd_sum_backward_par_alg": [_c_p, _i32] + [_c_p] * 9 + [_i32, _i32, _i32, _i64, _i32, _c_p],
    "pjamd_sum_forward_split": [_c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _i32, _i32, _i32, _i64, _i32, _i32,
                                _c_p, _i32, _c_p],
    "pjamd_flows_to_log": [_c_p, _i64, _c_p],
    "pjamd_sum_forward_queue": [_c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _i32, _i32, _i32, _i64, _i32, _c_p, _i32,
                                _c_p],
    "pjamd_sum_forward_split_noalloc": [_c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _i32, _i32, _i32, _i64, _i32,
                                        _i32, _i32, _c_p],
    "pjamd_sum_forward_prod": [_c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _i32, _i32, _i32, _i64, _i32, _c_p, _i32,
                               _i32, _c_p],
    "pjamd_sum_modify_flows": [_c_p, _c_p, _c_p, _i32, _i32, _i64, _c_p],
    "pjamd_sum_backward_ele": [_c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p,
                               _i32, _i32, _i32, _i32, _i64, _i32, _i32, _c_p],
    "pjamd_sum_backward_ele_to_children": [_c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p,
                                           _i32, _i32, _i32, _i32, _i32, _i64, _i32, _i32, _c_p],
    "pjamd_sum_backward_par": [_c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p,
                               _i32, _i32, _i32, _i64, _i32, _i32, _c_p],
    "pjamd_sum_backward_par_batched": [_c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p,
                                       _i32, _i32, _i64, _i32, _i32, _c_p],
    "pjamd_debug_mfma16_probe": [_i32, _c_p, _c_p, _c_p, _c_p],
    "pjamd_chain_pack": [_c_p, _c_p, _c_p, _c_p, _c_p, _i32, _i32, _c_p],
    "pjamd_chain_forward": [_c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _i32, _i64, _i32, _c_p],
    "pjamd_chain_backward": [_c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _i32, _i64, _i32, _c_p],
    "pjamd_sum_em": [_c_p, _c_p, _i32, _c_p, _i32, _f32, _f32, _i32, _i32, _c_p],
    "pjamd_pflow_fuse": [_c_p, _i32, _i64, _i64, _c_p, _i32, _c_p],
    "pjamd_sample_sum": [_c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _i32, _i32, _i32, _i64, _c_p],
    "pjamd_sample_prefix": [_c_p, _c_p, _c_p, _c_p, _c_p, _i64, _i32, _c_p],
    "pjamd_sample_prod": [_c_p, _c_p, _c_p, _c_p, _i32, _i32, _i32, _i64, _c_p],
    "pjamd_fill_f32": [_c_p, _f32, _i64, _c_p],
    "pjamd_debug_mfma_probe": [_i32, _c_p, _c_p, _c_p, _c_p],
}


def _load():
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        path = lib_path()
        if not os.path.exists(path):
            raise NativeError(f"pyjuice_amd native library missing at {path}; build it with "
                              f"`python -c 'import __graft_entry__ as g; g.build()'` from the repository root.")
        L = ctypes.CDLL(path)
        for name, args in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _i32
        L.pjamd_error_string.argtypes = [_i32]
        L.pjamd_error_string.restype = ctypes.c_char_p
        _lib = L
        return _lib


def lib():
    return _lib if _lib is not None else _load()


def is_available() -> bool:
    try:
        lib()
        return True
    except (NativeError, OSError):
        return False


def stream(device=None) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def ptr(t) -> ctypes.c_void_p:
    if t is None:
        return ctypes.c_void_p(0)
    return ctypes.c_void_p(t.data_ptr())


def check(rc: int, name: str):
    if rc != 0:
        msg = lib().pjamd_error_string(rc)
        raise NativeError(f"{name} failed with status {rc}: {msg.decode() if msg else '?'}")


# Optional per-launch timing: while a ``LaunchTimer`` is active, every ``call`` made with an
# ``algo=(bytes, flops)`` description is bracketed by HIP events on the current stream.
_timer = None


class LaunchTimer:
    """Collects ``(name, start_event, end_event, algo_bytes, algo_flops)`` per launch."""

    def __init__(self, names=None):
        self.names = None if names is None else set(names)
        self.records = []

    def __enter__(self):
        global _timer
        _timer = self
        return self

    def __exit__(self, *exc):
        global _timer
        _timer = None

    def summary(self):
        """Per entry point: launches, total ms, algorithmic bytes and flops, and the largest
        launch (most algorithmic bytes) as ``big = (bytes, ms)``."""
        torch.cuda.synchronize()
        out = dict()
        for name, a, b, nbytes, flops in self.records:
            d = out.setdefault(name, {"launches": 0, "ms": 0.0, "bytes": 0.0, "flops": 0.0, "big": (0.0, 0.0)})
            ms = a.elapsed_time(b)
            d["launches"] += 1
            d["ms"] += ms
            d["bytes"] += nbytes
            d["flops"] += flops
            if nbytes > d["big"][0]:
                d["big"] = (nbytes, ms)
        return out


def call(name: str, *args, algo=None):
    t = _timer
    if t is not None and algo is not None and (t.names is None or name in t.names):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        check(getattr(lib(), name)(*args), name)
        b.record()
        t.records.append((name, a, b, float(algo[0]), float(algo[1])))
        return
    check(getattr(lib(), name)(*args), name)
