"""Locates and loads the TorchScript operator library (python/python/hastypy/ffi/hasty_ffi.py).

The library is built in-tree (``python hastycompute_amd/build.py``) into
``hastycompute_amd/lib/_hasty_torch.so``; it links ``libhasty_hip.so`` beside it.
There is no CPU fallback: a missing library raises on import.
"""
from __future__ import annotations

import os
import threading

import torch

_LIB_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lib")
_lock = threading.Lock()
_loaded = False


def get_ffi_libfile(debug: bool = False, opsys: str = "Linux") -> str:
    """Path of the operator library (the arguments are accepted for interface parity)."""
    return os.path.join(_LIB_DIR, "_hasty_torch.so")


def load() -> None:
    global _loaded
    with _lock:
        if _loaded:
            return
        path = get_ffi_libfile()
        if not os.path.exists(path):
            raise ImportError(f"{path} is not built; run `python hastycompute_amd/build.py`")
        torch.classes.load_library(path)
        _loaded = True
