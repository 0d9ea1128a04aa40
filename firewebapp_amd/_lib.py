"""ctypes binding of libfirewave.so.

The argument types of every entry point are read from the C header
(include/firewave.h), so the binding and the header cannot drift apart.
The library is required: there is no CPU fallback.  If it is missing, every
call raises ``FirewaveNotBuilt``.
"""
from __future__ import annotations

import ctypes as C
import os
import re
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libfirewave.so")
HEADER = os.path.join(ROOT, "include", "firewave.h")

FW_OK = 0
FW_E_CONFIG = -1
FW_E_NOCONV = -2
FW_E_NONFINITE = -3
FW_E_OVERFLOW = -4
FW_E_HIP = -5
FW_F32 = 0
FW_F64 = 1
FW_REDUCE_BLOCKS = 256
FW_STEP_SCALARS = 8       # fw_spread_fields.step_scalars doubles


class FirewaveNotBuilt(RuntimeError):
    pass


class FirewaveError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{msg} (code {code})")
        self.code = code


class FwGrid(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("dtype", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double),
        ("dz", C.c_void_p), ("d_above", C.c_void_p), ("d_below", C.c_void_p),
        ("z_face", C.c_void_p),
        ("n_z_bed", C.c_int32), ("reserved", C.c_int32),
        ("inv_dz", C.c_void_p), ("inv_d_above", C.c_void_p), ("inv_d_below", C.c_void_p),
    ]


_TYPES = {
    "int": C.c_int, "int32_t": C.c_int32, "int64_t": C.c_int64,
    "double": C.c_double, "float": C.c_float, "uint64_t": C.c_uint64,
}


def _ctype(decl: str):
    decl = decl.strip()
    if "*" in decl:
        return C.c_void_p
    toks = decl.replace("const ", "").split()
    base = toks[0]
    if base not in _TYPES:
        raise ValueError(f"unmapped C type in header: {decl!r}")
    return _TYPES[base]


def parse_header(path: str = HEADER):
    """Return {name: (restype, [argtypes])} for every declared entry point."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in re.finditer(r"^(int|const char\*|double|void)\s+(fw_\w+)\s*\(([^)]*)\)\s*;",
                         text, flags=re.M):
        ret, name, args = m.group(1), m.group(2), m.group(3)
        args = " ".join(args.split())
        argtypes = ([] if args in ("", "void")
                    else [_ctype(t) for t in _split_args(args)])
        restype = {"int": C.c_int, "const char*": C.c_char_p,
                   "double": C.c_double, "void": None}[ret]
        out[name] = (restype, argtypes)
    return out


def _split_args(args: str):
    parts = [a.strip() for a in args.split(",")]
    res = []
    for a in parts:
        # drop the parameter name
        m = re.match(r"^(.*?)([\w]+)$", a)
        typ = m.group(1).strip() if m else a
        res.append(typ)
    return res


_lock = threading.Lock()
_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise FirewaveNotBuilt(
                f"{LIB_PATH} is missing: build it with "
                "`python -m firewebapp_amd.build` (there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in parse_header().items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
        return L


def check(rc: int, what: str = ""):
    if rc != FW_OK:
        msg = lib().fw_last_error().decode(errors="replace")
        raise FirewaveError(rc, f"{what}: {msg}" if what else msg)
    return rc


def call(name: str, *args):
    """Call entry point `name`, raising FirewaveError on a non-zero code."""
    return check(getattr(lib(), name)(*args), name)
