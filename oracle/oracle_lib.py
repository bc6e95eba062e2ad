"""ctypes loader for the FP64 CPU oracle (oracle/liboracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it, as the checker.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = C.CDLL(ORACLE_SO)
        L.orc_step_batch.restype = C.c_int64
        L.orc_step_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                     C.c_double, C.c_int, C.c_void_p]
        L.orc_update_activity_batch.restype = None
        L.orc_update_activity_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p,
                                                C.c_int64, C.c_int]
        L.orc_record_rel_residual.restype = C.c_double
        L.orc_record_rel_residual.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_int64, C.c_void_p]
        L.orc_sizeof_record.restype = C.c_size_t
        L.orc_lu_decomposition.restype = C.c_int
        L.orc_lu_decomposition.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        L.orc_lu_back_substitution.restype = None
        L.orc_lu_back_substitution.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        L.orc_rsolve.restype = C.c_int
        L.orc_rsolve.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                 C.c_int]
        _lib = L
    return _lib


class Records:
    """Array of orc_substep_record, one per cell."""

    def __init__(self, ncell):
        self.size = lib().orc_sizeof_record()
        self.buf = C.create_string_buffer(self.size * max(ncell, 1))
        self.ncell = ncell

    @property
    def ptr(self):
        return C.addressof(self.buf)

    def at(self, cell):
        return self.ptr + cell * self.size


def record_fields(records, name):
    """numpy view of one double field of every orc_substep_record"""
    import ctypes as C
    off = {"dt": 0, "r0_norm": None, "final_rel_residual": None}
    size = records.size
    # r0_norm and final_rel_residual are the last two doubles of the record
    idx = {"r0_norm": size - 16, "final_rel_residual": size - 8, "dt": 0}[name]
    raw = np.frombuffer(records.buf, dtype=np.uint8, count=size * records.ncell)
    return np.ndarray((records.ncell,), dtype=np.float64, buffer=raw, offset=idx, strides=(size,))


def step_batch(net, opt, state, dt, nthreads=0, records=None):
    """RStep over all cells of a host CellState (in place). Returns #failed."""
    n = net.to_abi()
    o = opt.to_abi()
    s = state.to_abi()
    return lib().orc_step_batch(C.byref(n), C.byref(o), C.byref(s), state.ncell, dt,
                                nthreads, records.ptr if records else None)


def update_activity_batch(net, opt, state, nthreads=0):
    n = net.to_abi()
    o = opt.to_abi()
    s = state.to_abi()
    lib().orc_update_activity_batch(C.byref(n), C.byref(o), C.byref(s), state.ncell,
                                    nthreads)


def rel_residuals(net, opt, state, records, cells):
    n = net.to_abi()
    o = opt.to_abi()
    s = state.to_abi()
    return np.array([lib().orc_record_rel_residual(C.byref(n), C.byref(o), C.byref(s),
                                                   int(c), records.at(int(c)))
                     for c in cells])


def lu_solve(a, b):
    """LUDecomposition + LUBackSubstitution on a dense matrix (row-major)."""
    a = np.ascontiguousarray(a, dtype=np.float64).copy()
    b = np.ascontiguousarray(b, dtype=np.float64).copy()
    n = a.shape[0]
    indx = np.zeros(n, dtype=np.int32)
    d = C.c_int(0)
    err = lib().orc_lu_decomposition(a.ctypes.data, n, indx.ctypes.data, C.byref(d))
    if err:
        return None
    lib().orc_lu_back_substitution(a.ctypes.data, n, indx.ctypes.data, b.ctypes.data)
    return b
