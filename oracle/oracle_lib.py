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
# the same sources built -O3 with FMA contraction: timing of the CPU
# baseline only, never the checker
ORACLE_FAST_SO = os.path.join(ORACLE_DIR, "liboracle_fast.so")
ORACLE_F90_SO = os.path.join(ORACLE_DIR, "liboracle_f90.so")
ORACLE_F90_IEEE_SO = os.path.join(ORACLE_DIR, "liboracle_f90_ieee.so")

_lib = None
_fast = None
_f90 = {}

MAXCOMP, MAXCPLX, MAXMNRL, MAXRXN, SOMDEC_MAX_DOWN = 96, 512, 32, 32, 4


class OrcCell(C.Structure):
    """ctypes mirror of orc_cell (oracle/pfchem_oracle.h)"""
    _fields_ = [
        ("pri_molal", C.c_double * MAXCOMP),
        ("total", C.c_double * MAXCOMP),
        ("immobile", C.c_double * MAXCOMP),
        ("sec_molal", C.c_double * MAXCPLX),
        ("pri_act_coef", C.c_double * MAXCOMP),
        ("sec_act_coef", C.c_double * MAXCPLX),
        ("ln_act_h2o", C.c_double),
        ("mnrl_volfrac", C.c_double * MAXMNRL),
        ("mnrl_area", C.c_double * MAXMNRL),
        ("mnrl_rate", C.c_double * MAXMNRL),
        ("porosity", C.c_double), ("sat", C.c_double), ("den_kg", C.c_double),
        ("volume", C.c_double), ("temp", C.c_double),
        ("f_t", C.c_double), ("f_w", C.c_double),
        ("has_f_t", C.c_int), ("has_f_w", C.c_int),
        ("dtotal", C.c_double * (MAXCOMP * MAXCOMP)),
        ("eqcplx_logK", C.c_double * MAXCPLX),
        ("kinmnrl_logK", C.c_double * MAXMNRL),
        ("sd_upstream_nc", C.c_double * MAXRXN),
        ("sd_downstream_nc", C.c_double * (MAXRXN * SOMDEC_MAX_DOWN)),
        ("sd_mineral_c_stoich", C.c_double * MAXRXN),
        ("sd_mineral_n_stoich", C.c_double * MAXRXN),
        ("activity_error", C.c_int),
    ]

    def arr(self, name, n):
        return np.ctypeslib.as_array(getattr(self, name))[:n]


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
        L.orc_sizeof_cell.restype = C.c_size_t
        vp = C.c_void_p
        for name, args in [("orc_cell_defaults", [vp, vp]),
                           ("orc_update_temp_dependent_coefs", [vp, vp, vp]),
                           ("orc_activity_coefficients", [vp, vp, vp]),
                           ("orc_total", [vp, vp]),
                           ("orc_accumulation", [vp, vp, vp]),
                           ("orc_accumulation_derivative", [vp, vp, C.c_double, vp]),
                           ("orc_reaction", [vp, vp, C.c_double, vp, vp, C.c_int]),
                           ("orc_reaction_derivative_numerical", [vp, vp, C.c_double, vp, vp]),
                           ("orc_rreact_system", [vp, vp, vp, vp, C.c_double, vp, vp, C.c_int]),
                           ("orc_update_kinetic_state", [vp, vp, C.c_double]),
                           ("orc_gather", [vp, vp, C.c_int64, vp]),
                           ("orc_scatter", [vp, vp, vp, C.c_int64])]:
            f = getattr(L, name)
            f.restype = None
            f.argtypes = args
        L.orc_rreact.restype = C.c_int
        L.orc_rreact.argtypes = [vp, vp, vp, vp, C.c_double, vp, vp, vp]
        L.orc_rstep.restype = C.c_int
        L.orc_rstep.argtypes = [vp, vp, vp, vp, C.c_double, vp, vp, vp, vp]
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


def fast_lib():
    """the -O3 build of the oracle (CPU-baseline timing only)"""
    global _fast
    if _fast is None:
        if not os.path.exists(ORACLE_FAST_SO):
            build()
        L = C.CDLL(ORACLE_FAST_SO)
        L.orc_step_batch.restype = C.c_int64
        L.orc_step_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                     C.c_double, C.c_int, C.c_void_p]
        _fast = L
    return _fast


def f90_lib(ieee=False):
    """the Fortran restatement of RStep (pfchem_oracle_rstep.F90): the -O3
    build for the CPU-baseline timing, or (ieee) the build without FMA
    contraction that the tests hold bit-equal to the checker"""
    path = ORACLE_F90_IEEE_SO if ieee else ORACLE_F90_SO
    if path not in _f90:
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_f90_step_batch.restype = C.c_int64
        L.orc_f90_step_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                         C.c_double, C.c_int]
        _f90[path] = L
    return _f90[path]


def step_batch_f90(net, opt, state, dt, nthreads=0, ieee=False):
    """RStep over all cells of a host CellState (in place) on the Fortran
    restatement.  Returns #failed; raises for networks outside its scope
    (aqueous complexes and TST kinetic minerals, activity coefficients fixed
    within the step, isothermal log K)."""
    n = net.to_abi()
    o = opt.to_abi()
    s = state.to_abi()
    r = f90_lib(ieee).orc_f90_step_batch(C.byref(n), C.byref(o), C.byref(s), state.ncell, dt,
                                     nthreads)
    if r < 0:
        raise ValueError("network or options outside the Fortran restatement's scope")
    return r


def step_batch(net, opt, state, dt, nthreads=0, records=None, fast=False):
    """RStep over all cells of a host CellState (in place). Returns #failed.
    fast: run the -O3 build (for timing the CPU baseline)."""
    n = net.to_abi()
    o = opt.to_abi()
    s = state.to_abi()
    L = fast_lib() if fast else lib()
    return L.orc_step_batch(C.byref(n), C.byref(o), C.byref(s), state.ncell, dt,
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


class CellCtx:
    """one oracle cell bound to a network and options (keeps the ABI structs
    and their arrays alive)"""

    def __init__(self, net, opt):
        self.net, self.opt = net, opt
        self.n = net.to_abi()
        self.o = opt.to_abi()
        self.c = OrcCell()
        assert C.sizeof(OrcCell) == lib().orc_sizeof_cell()
        lib().orc_cell_defaults(C.byref(self.n), C.byref(self.c))
        self.ncomp = net.naqcomp + net.nimmobile

    def gather(self, state, cell):
        s = state.to_abi()
        lib().orc_gather(C.byref(self.n), C.byref(s), cell, C.byref(self.c))
        return self

    def set(self, name, values):
        a = np.ctypeslib.as_array(getattr(self.c, name))
        v = np.asarray(values, dtype=np.float64)
        a[:v.size] = v

    def get(self, name, n):
        return np.ctypeslib.as_array(getattr(self.c, name))[:n].copy()

    def activity(self):
        lib().orc_activity_coefficients(C.byref(self.n), C.byref(self.o), C.byref(self.c))

    def total(self):
        lib().orc_total(C.byref(self.n), C.byref(self.c))

    def dtotal(self):
        naq = self.net.naqcomp
        return self.get("dtotal", naq * naq).reshape(naq, naq)

    def reaction(self, dt, derivative=True):
        n = self.ncomp
        res = np.zeros(n)
        jac = np.zeros((n, n))
        lib().orc_reaction(C.byref(self.n), C.byref(self.c), dt, res.ctypes.data,
                           jac.ctypes.data, 1 if derivative else 0)
        return res, jac

    def reaction_numerical(self, dt):
        n = self.ncomp
        res = np.zeros(n)
        jac = np.zeros((n, n))
        lib().orc_reaction_derivative_numerical(C.byref(self.n), C.byref(self.c), dt,
                                                res.ctypes.data, jac.ctypes.data)
        return res, jac

    def rreact_system(self, fixed, dt, derivative=True):
        n = self.ncomp
        res = np.zeros(n)
        jac = np.zeros((n, n))
        f = np.ascontiguousarray(fixed, dtype=np.float64)
        lib().orc_rreact_system(C.byref(self.n), C.byref(self.o), C.byref(self.c),
                                f.ctypes.data, dt, res.ctypes.data, jac.ctypes.data,
                                1 if derivative else 0)
        return res, jac

    def rstep(self, dt, guess=None):
        n = self.ncomp
        g = np.zeros(MAXCOMP)
        naq = self.net.naqcomp
        g[:naq] = self.get("pri_molal", naq) if guess is None else guess[:naq]
        g[naq:n] = self.get("immobile", n - naq)
        ns, ni, nk = C.c_int(0), C.c_int(0), C.c_int(0)
        rec = Records(1)
        err = lib().orc_rstep(C.byref(self.n), C.byref(self.o), C.byref(self.c),
                              g.ctypes.data, dt, C.byref(ns), C.byref(ni), C.byref(nk),
                              rec.ptr)
        return err, ns.value, ni.value

    def rreact(self, dt, guess=None):
        """one RReact solve from the cell's totals (no RStep sub-stepping and
        no small-total clamp); returns (ierror, iterations, fail_reason)"""
        n = self.ncomp
        naq = self.net.naqcomp
        g = np.zeros(MAXCOMP)
        g[:naq] = self.get("pri_molal", naq) if guess is None else guess[:naq]
        g[naq:n] = self.get("immobile", n - naq)
        ni, why = C.c_int(0), C.c_int(0)
        rec = Records(1)
        err = lib().orc_rreact(C.byref(self.n), C.byref(self.o), C.byref(self.c),
                               g.ctypes.data, dt, C.byref(ni), rec.ptr, C.byref(why))
        return err, ni.value, why.value

    def update_kinetic_state(self, dt):
        lib().orc_update_kinetic_state(C.byref(self.n), C.byref(self.c), dt)


def _sens_lib():
    L = lib()
    if not getattr(L, "_sens_bound", False):
        vp = C.c_void_p
        L.orc_adjoint_batch.restype = C.c_int64
        L.orc_adjoint_batch.argtypes = [vp, vp, vp, C.c_int64, C.c_double, vp, vp, C.c_int,
                                        C.c_int, vp, vp, vp]
        L.orc_tangent_batch.restype = C.c_int64
        L.orc_tangent_batch.argtypes = [vp, vp, vp, C.c_int64, C.c_double, vp, vp, vp, C.c_int,
                                        C.c_int, vp, vp, vp]
        L._sens_bound = True
    return L


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def adjoint_batch(net, opt, state, dt, wv, ws, exact=True, nthreads=0):
    """(sum over cells of d f / d p, f per cell, status per cell) with
    p = (log10 k of the kinetic minerals, log10 K of the complexes)"""
    n, o, s = net.to_abi(), opt.to_abi(), state.to_abi()
    wv, ws = _f64(wv), _f64(ws)
    grad = np.zeros(max(net.nkinmnrl + net.neqcplx, 1))
    f = np.zeros(state.ncell)
    st = np.zeros(state.ncell, dtype=np.int32)
    _sens_lib().orc_adjoint_batch(C.byref(n), C.byref(o), C.byref(s), state.ncell, dt,
                                  wv.ctypes.data, ws.ctypes.data, 1 if exact else 0, nthreads,
                                  grad.ctypes.data, f.ctypes.data, st.ctypes.data)
    return grad[:net.nkinmnrl + net.neqcplx], f, st


def tangent_batch(net, opt, state, dt, wv, ws, dp, exact=True, nthreads=0):
    """(d f per cell along dp, f per cell, status per cell)"""
    n, o, s = net.to_abi(), opt.to_abi(), state.to_abi()
    wv, ws, dp = _f64(wv), _f64(ws), _f64(dp)
    df = np.zeros(state.ncell)
    f = np.zeros(state.ncell)
    st = np.zeros(state.ncell, dtype=np.int32)
    _sens_lib().orc_tangent_batch(C.byref(n), C.byref(o), C.byref(s), state.ncell, dt,
                                  wv.ctypes.data, ws.ctypes.data, dp.ctypes.data,
                                  1 if exact else 0, nthreads, df.ctypes.data, f.ctypes.data,
                                  st.ctypes.data)
    return df, f, st


def equilibrate_batch(net, opt, con, state, conc, nthreads=0):
    """ReactionEquilibrateConstraint over all cells of a host CellState (in
    place); conc: (naqcomp, ld) constraint values.  Returns #failed."""
    L = lib()
    if not getattr(L, "_con_set", False):
        L.orc_equilibrate_batch.restype = C.c_int64
        L.orc_equilibrate_batch.argtypes = [C.c_void_p] * 5 + [C.c_int64, C.c_int]
        L._con_set = True
    n = net.to_abi()
    o = opt.to_abi()
    c = con.to_abi(net)
    s = state.to_abi()
    cv = np.ascontiguousarray(conc, dtype=np.float64)
    assert cv.shape[0] == net.naqcomp and cv.shape[1] >= state.ncell
    assert cv.shape[1] == state.ld
    return L.orc_equilibrate_batch(C.byref(n), C.byref(o), C.byref(c), C.byref(s),
                                   cv.ctypes.data, state.ncell, nthreads)
