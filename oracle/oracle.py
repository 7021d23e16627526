"""ctypes wrapper of oracle/_build/liboracle.so - the CPU oracle (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker or the timed CPU baseline; the product path
(dba_amd) never does.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

_lib = None


def _params_type():
    from dba_amd.abi import Params  # the struct layout is the shared interface
    return Params


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} missing: build it with __graft_entry__.build()")
        _lib = load_path(LIB_PATH)
    return _lib


def load_path(path):
    """The oracle built at `path` (tests load copies built with a planted mistake)."""
    lib = C.CDLL(path)
    P = C.POINTER(_params_type())
    lib.ora_create.restype = C.c_void_p
    lib.ora_create.argtypes = [P]
    lib.ora_destroy.argtypes = [C.c_void_p]
    for name in ("ora_set_field", "ora_get_field"):
        getattr(lib, name).argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        getattr(lib, name).restype = C.c_int
    lib.ora_step.argtypes = [C.c_void_p, C.c_int]
    for name in ("ora_get_awake", "ora_set_awake", "ora_get_next"):
        getattr(lib, name).argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        getattr(lib, name).restype = C.c_int
    lib.ora_ntiles.argtypes = [C.c_void_p]
    lib.ora_history.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
    lib.ora_time.argtypes = [C.c_void_p]
    lib.ora_time.restype = C.c_double
    lib.ora_last_dt.argtypes = [C.c_void_p]
    lib.ora_last_dt.restype = C.c_double
    lib.ora_dilate.argtypes = [C.c_void_p, C.c_void_p] + [C.c_int] * 7
    lib.ora_compact.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    lib.ora_compact.restype = C.c_int
    lib.ora_bench.argtypes = [C.c_void_p, C.c_int, C.c_int]
    lib.ora_bench.restype = C.c_double
    return lib


class OracleSim:
    """One oracle simulation with the dba_params of the GPU instance it checks."""

    def __init__(self, params, lib=None):
        self.lib = lib if lib is not None else load()
        self.params = params.copy()
        self.h = self.lib.ora_create(C.byref(self.params))
        if not self.h:
            raise ValueError("oracle does not support these parameters")
        self.n = params.nx * params.ny * params.nz
        self.shape = (params.nz, params.ny, params.nx) if params.nz > 1 else (params.ny, params.nx)

    def set_field(self, f, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        assert a.size == self.n
        assert self.lib.ora_set_field(self.h, f, a.ctypes.data) == 0

    def get_field(self, f):
        out = np.empty(self.n)
        assert self.lib.ora_get_field(self.h, f, out.ctypes.data) == 0
        return out.reshape(self.shape)

    def step(self, n=1):
        assert self.lib.ora_step(self.h, n) == 0

    @property
    def ntiles(self):
        return self.lib.ora_ntiles(self.h)

    def awake(self, s=0):
        out = np.zeros(self.ntiles, dtype=np.uint8)
        assert self.lib.ora_get_awake(self.h, s, out.ctypes.data) == 0
        return out

    def set_awake(self, s, mask):
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        assert self.lib.ora_set_awake(self.h, s, m.ctypes.data) == 0

    def next_flags(self, s=0):
        out = np.zeros(self.ntiles, dtype=np.uint8)
        assert self.lib.ora_get_next(self.h, s, out.ctypes.data) == 0
        return out

    def history(self):
        cap = max(1, self.params.history_cap)
        buf = np.zeros(2 * cap, dtype=np.int32)
        n = self.lib.ora_history(self.h, buf.ctypes.data, cap)
        return buf[: 2 * n].reshape(-1, 2).copy()

    def time(self):
        return self.lib.ora_time(self.h)

    def bench(self, nsteps, nthreads=1):
        return self.lib.ora_bench(self.h, nsteps, nthreads)

    def close(self):
        if self.h:
            self.lib.ora_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dilate(N, ntx, nty, ntz=1, kind=0, wrap=(1, 1, 1)):
    lib = load()
    N = np.ascontiguousarray(N, dtype=np.uint8).reshape(-1)
    A = np.zeros_like(N)
    lib.ora_dilate(N.ctypes.data, A.ctypes.data, ntx, nty, ntz, kind, wrap[0], wrap[1], wrap[2])
    return A


def compact(A):
    lib = load()
    A = np.ascontiguousarray(A, dtype=np.uint8).reshape(-1)
    out = np.zeros(max(1, A.size), dtype=np.int32)
    c = lib.ora_compact(A.ctypes.data, A.size, out.ctypes.data)
    return out[:c].copy()
