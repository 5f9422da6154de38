"""Loader of libchimera_hip.so.

The ctypes signatures are read from include/chimera_hip.h, so the header is the
single description of the C ABI.  There is no fallback: if the library is
missing or fails to load, every caller gets an ImportError-like RuntimeError.
"""
from __future__ import annotations

import ctypes
import os
import re
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libchimera_hip.so")
HEADER = os.path.join(ROOT, "include", "chimera_hip.h")

_DECL = re.compile(r"^(int32_t|const char\*|void\*)\s+(chimera_\w+)\((.*?)\);", re.S | re.M)


def _ctype(param: str):
    p = param.strip()
    if p in ("void", ""):
        return None
    if "*" in p:
        return ctypes.c_void_p
    typ = p.rsplit(" ", 1)[0].replace("const ", "").strip()
    return {"double": ctypes.c_double, "int32_t": ctypes.c_int32, "int64_t": ctypes.c_int64}[typ]


def parse_header(path: str = HEADER):
    """{name: (restype, [argtypes])} for every entry point the header declares."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for ret, name, params in _DECL.findall(text):
        args = [_ctype(p) for p in params.replace("\n", " ").split(",")]
        args = [a for a in args if a is not None]
        rt = {"int32_t": ctypes.c_int32, "const char*": ctypes.c_char_p, "void*": ctypes.c_void_p}[ret]
        out[name] = (rt, args)
    return out


_lib = None

# Held around every library call made from Python, and by a step capture for
# its whole length: HIP invalidates a stream capture when another thread
# allocates, frees or copies synchronously meanwhile, and any library call may
# grow scratch memory or upload a table.  Reentrant, so the calls inside a
# capture take it again.
DEVICE_LOCK = threading.RLock()


def lib():
    """The loaded library with argtypes set (loads on first use)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libchimera_hip is not built ({LIB_PATH} missing); run __graft_entry__.build() "
                           "or `python -m chimera_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    for name, (rt, args) in parse_header().items():
        fn = getattr(L, name)
        fn.restype = rt
        fn.argtypes = args
    _lib = L
    return L


class ChimeraError(RuntimeError):
    def __init__(self, name: str, status: int, msg: str):
        super().__init__(f"{name} returned {status}: {msg}")
        self.status = status


def check(name: str, status: int):
    if status != 0:
        raise ChimeraError(name, status, lib().chimera_last_error().decode(errors="replace"))
