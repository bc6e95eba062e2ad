"""ctypes binding of libpfchem.so (the HIP chemistry step, include/pfchem.h).

The library must be built (``__graft_entry__.build()`` or
``make -C pflotran-elm-interface_amd/csrc``); loading fails loudly
otherwise -- there is no CPU fallback.  torch is used only for device memory
and streams: it must be imported before the library is loaded so both share
one HIP runtime.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Optional

import numpy as np

from . import abi
from .network import CellState, ReactionNetwork, SolverOptions, state_fields

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "libpfchem.so")

_lib = None


class PfchemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("pfchem error %d (%s): %s" % (code, abi_strerror(code), msg))
        self.code = code


def load(path: str = LIB_PATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libpfchem.so not built at %s (run __graft_entry__.build())" % path)
    import torch  # noqa: F401  (shares torch's HIP runtime)
    L = C.CDLL(path)
    vp = C.c_void_p
    L.pfchem_create.argtypes = [vp, vp, C.c_int, C.POINTER(vp)]
    L.pfchem_destroy.argtypes = [vp]
    L.pfchem_strerror.restype = C.c_char_p
    L.pfchem_strerror.argtypes = [C.c_int]
    L.pfchem_last_error.restype = C.c_char_p
    L.pfchem_last_error.argtypes = [vp]
    L.pfchem_step.argtypes = [vp, C.c_int64, vp, C.c_double, C.c_int, vp]
    L.pfchem_set_pass_timing.argtypes = [vp, C.c_int]
    L.pfchem_pass_time.argtypes = [vp, C.POINTER(C.c_float)]
    L.pfchem_step_host.argtypes = [vp, C.c_int64, vp, C.c_double, C.c_int]
    L.pfchem_update_activity.argtypes = [vp, C.c_int64, vp, vp]
    L.pfchem_residual.argtypes = [vp, C.c_int64, vp, vp, vp, C.c_double, vp, vp]
    L.pfchem_residual_jacobian.argtypes = [vp, C.c_int64, vp, vp, vp, C.c_double, vp, vp, vp]
    L.pfchem_equilibrate.argtypes = [vp, C.c_int64, vp, vp, vp, vp]
    L.pfchem_equilibrate_static.argtypes = [vp, C.c_int64, vp, vp, vp, vp]
    L.pfchem_set_launch.argtypes = [vp, C.c_int, vp]
    L.pfchem_get_launch.argtypes = [vp, C.c_int, vp]
    L.pfchem_tangent.argtypes = [vp, C.c_int64, vp, C.c_double, vp, vp, vp, C.c_int, vp, vp, vp,
                                 vp]
    L.pfchem_adjoint.argtypes = [vp, C.c_int64, vp, C.c_double, vp, vp, C.c_int, vp, vp, vp, vp,
                                 vp]
    L.pfchem_prepare_kernels.argtypes = [vp, vp, C.c_char_p]
    L.pfchem_autotune.argtypes = [vp, C.c_int, C.c_int64, vp, C.c_double, C.c_char_p,
                                  C.c_char_p, C.c_int, vp, vp]
    L.pfchem_tune_key.argtypes = [vp, C.c_char_p, C.c_int, C.c_char_p, C.c_int]
    L.pfchem_tune_cache_query.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int,
                                          C.c_int64, vp]
    L.pfchem_transport_setup.argtypes = [vp, C.c_int64, vp]
    L.pfchem_transport_step.argtypes = [vp, C.c_int64, vp, vp, C.c_double, C.c_double, C.c_int,
                                        vp, vp]
    _lib = L
    return L


def tune_cache_query(cache_path: str, key: str, bucket: str, variant: int,
                     ncell: int) -> Optional[dict]:
    """pfchem_tune_cache_query: the launch shape a tuner cache file holds
    for (key, else bucket; variant; nearest cell-count bucket), or None"""
    L = load()
    c = abi.Launch()
    rc = L.pfchem_tune_cache_query(cache_path.encode(), key.encode(), bucket.encode(), variant,
                                   ncell, C.byref(c))
    if rc < 0:
        raise PfchemError(-rc, "pfchem_tune_cache_query")
    if rc == 0:
        return None
    return dict(lanes_per_cell=c.lanes_per_cell, rows_per_lane=c.rows_per_lane,
                block_size=c.block_size, blocks_per_cu=c.blocks_per_cu)


def prepare_kernels(net: ReactionNetwork, opt: SolverOptions,
                    source_path: Optional[str] = None) -> int:
    """pfchem_prepare_kernels: generate and compile (no device needed) the
    network-specialized one-lane kernels of (net, opt).  Returns PFCHEM_OK,
    or PFCHEM_ERR_UNSUPPORTED when the one-lane kernel does not take the
    network; raises on a compilation failure."""
    L = load()
    n, o = net.to_abi(), opt.to_abi()
    rc = L.pfchem_prepare_kernels(C.byref(n), C.byref(o),
                                  source_path.encode() if source_path else None)
    if rc not in (abi.PFCHEM_OK, abi.PFCHEM_ERR_UNSUPPORTED):
        raise PfchemError(rc, "pfchem_prepare_kernels")
    return rc


def _stream_ptr(stream):
    """hipStream_t of a torch stream (torch's current stream by default)"""
    import torch
    return (stream or torch.cuda.current_stream()).cuda_stream


def abi_strerror(code):
    try:
        return load().pfchem_strerror(code).decode()
    except Exception:  # pragma: no cover
        return "?"


class DeviceState:
    """SoA state in device memory (torch tensors shaped (rows, ld))."""

    def __init__(self, host: CellState, device="cuda"):
        import torch
        self.net = host.net
        self.ncell = host.ncell
        self.ld = host.ld
        self.fields = host.fields
        self.t: Dict[str, "torch.Tensor"] = {}
        for name, a in host.arrays.items():
            self.t[name] = torch.from_numpy(np.ascontiguousarray(a)).to(device)

    def clone(self) -> "DeviceState":
        """device-side copy (same shapes, fresh allocations)"""
        d = DeviceState.__new__(DeviceState)
        d.net, d.ncell, d.ld, d.fields = self.net, self.ncell, self.ld, self.fields
        d.t = {name: t.clone() for name, t in self.t.items()}
        return d

    def to_host(self) -> CellState:
        out = CellState.__new__(CellState)
        out.net, out.ncell, out.ld, out.fields = self.net, self.ncell, self.ld, self.fields
        out.arrays = {k: v.cpu().numpy() for k, v in self.t.items()}
        return out

    def to_abi(self) -> abi.State:
        s = abi.State()
        s.ld = self.ld
        for name, _ in abi.State._fields_:
            if name == "ld":
                continue
            tt = self.t.get(name)
            setattr(s, name, None if tt is None or tt.shape[0] == 0 else tt.data_ptr())
        return s


class Context:
    """pfchem_ctx: network + options uploaded to one device."""

    def __init__(self, net: ReactionNetwork, opt: SolverOptions, device: int = 0):
        self.L = load()
        self.net = net
        self.opt = opt
        self._n = net.to_abi()
        self._o = opt.to_abi()
        h = C.c_void_p()
        rc = self.L.pfchem_create(C.byref(self._n), C.byref(self._o), device, C.byref(h))
        if rc:
            raise PfchemError(rc, "pfchem_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.pfchem_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc:
            raise PfchemError(rc, "%s: %s" % (what, self.L.pfchem_last_error(self.h).decode()))

    def step(self, st: DeviceState, dt: float, variant: int = abi.KERNEL_BATCHED_MIXED,
             stream=None, ncell: Optional[int] = None):
        s = st.to_abi()
        sp = None
        if stream is not None:
            sp = stream.cuda_stream
        else:
            import torch
            sp = torch.cuda.current_stream().cuda_stream
        n = st.ncell if ncell is None else ncell
        self._check(self.L.pfchem_step(self.h, n, C.byref(s), dt, variant, sp), "pfchem_step")

    def set_pass_timing(self, on: bool = True):
        """pfchem_set_pass_timing: HIP events around the one-lane pass"""
        self._check(self.L.pfchem_set_pass_timing(self.h, 1 if on else 0), "pfchem_set_pass_timing")

    def pass_time(self) -> float:
        """pfchem_pass_time: the last timed one-lane pass, in ms"""
        ms = C.c_float()
        self._check(self.L.pfchem_pass_time(self.h, C.byref(ms)), "pfchem_pass_time")
        return float(ms.value)

    def step_host(self, st: CellState, dt: float, variant: int = abi.KERNEL_BATCHED_MIXED):
        s = st.to_abi()
        rc = self.L.pfchem_step_host(self.h, st.ncell, C.byref(s), dt, variant)
        if rc and rc != abi.PFCHEM_ERR_CELLS:
            self._check(rc, "pfchem_step_host")
        return rc

    def update_activity(self, st: DeviceState, stream=None):
        import torch
        s = st.to_abi()
        sp = (stream or torch.cuda.current_stream()).cuda_stream
        self._check(self.L.pfchem_update_activity(self.h, st.ncell, C.byref(s), sp),
                    "pfchem_update_activity")

    def residual(self, st: DeviceState, fixed_total, fixed_immobile, dt: float,
                 jacobian: bool = True, stream=None):
        """pfchem_residual_jacobian: the GIRT reaction block of one iterate at
        the state's free molalities.  fixed_total / fixed_immobile: device
        tensors (naqcomp, ld) / (nimmobile, ld) or None.  Returns (res
        (ncomp, ld), jac (ncomp * ncomp, ld) or None) as device tensors; the
        call also rewrites total, sec_molal and mnrl_rate in `st`."""
        import torch
        s = st.to_abi()
        sp = (stream or torch.cuda.current_stream()).cuda_stream
        ncomp = self.net.naqcomp + self.net.nimmobile
        res = torch.zeros((ncomp, st.ld), dtype=torch.float64, device="cuda")
        jac = (torch.zeros((ncomp * ncomp, st.ld), dtype=torch.float64, device="cuda")
               if jacobian else None)

        def ptr(t):
            return None if t is None or t.numel() == 0 else t.data_ptr()
        self._check(self.L.pfchem_residual_jacobian(self.h, st.ncell, C.byref(s), ptr(fixed_total),
                                                    ptr(fixed_immobile), dt, res.data_ptr(),
                                                    ptr(jac), sp), "pfchem_residual_jacobian")
        return res, jac

    def equilibrate(self, st: DeviceState, con, conc, stream=None, check: bool = True):
        """pfchem_equilibrate: constraint speciation of every cell of `st`.
        con: constraint.Constraint; conc: device tensor (naqcomp, ld) of
        constraint values.  Blocks until done.  With check, a cell that
        failed raises PfchemError (its code is in st's status array);
        otherwise the return code is returned."""
        import torch
        if conc.dtype != torch.float64 or not conc.is_cuda or not conc.is_contiguous():
            raise ValueError("conc must be a contiguous float64 device tensor")
        if tuple(conc.shape) != (self.net.naqcomp, st.ld):
            raise ValueError("conc must have shape (naqcomp, ld)")
        s = st.to_abi()
        k = con.to_abi(self.net)
        sp = (stream or torch.cuda.current_stream()).cuda_stream
        rc = self.L.pfchem_equilibrate(self.h, st.ncell, C.byref(s), C.byref(k),
                                       conc.data_ptr(), sp)
        if check:
            self._check(rc, "pfchem_equilibrate")
        return rc

    def set_launch(self, variant, lanes_per_cell, block_size=256, blocks_per_cu=4,
                   rows_per_lane=1):
        c = abi.Launch(lanes_per_cell, block_size, blocks_per_cu, rows_per_lane)
        self._check(self.L.pfchem_set_launch(self.h, variant, C.byref(c)), "pfchem_set_launch")

    def try_set_launch(self, variant, lanes_per_cell, block_size=256, blocks_per_cu=4,
                       rows_per_lane=1):
        """set_launch that returns False instead of raising when the shape does
        not fit this network"""
        c = abi.Launch(lanes_per_cell, block_size, blocks_per_cu, rows_per_lane)
        return self.L.pfchem_set_launch(self.h, variant, C.byref(c)) == abi.PFCHEM_OK

    def _weights(self, w_vf, w_sec):
        nm, nc = self.net.nkinmnrl, self.net.neqcplx
        wv = np.ascontiguousarray(np.zeros(nm) if w_vf is None else w_vf, dtype=np.float64)
        ws = np.ascontiguousarray(np.zeros(nc) if w_sec is None else w_sec, dtype=np.float64)
        if wv.shape != (nm,) or ws.shape != (nc,):
            raise ValueError("weights must have nkinmnrl / neqcplx entries")
        return wv, ws

    def tangent(self, st: DeviceState, dt: float, dp, w_vf=None, w_sec=None, exact=True,
                stream=None):
        """pfchem_tangent: (df per cell, f per cell, status per cell) as device
        tensors, for the parameter direction dp (log10 k then log10 K)"""
        import torch
        wv, ws = self._weights(w_vf, w_sec)
        dpv = np.ascontiguousarray(dp, dtype=np.float64)
        if dpv.shape != (wv.size + ws.size,):
            raise ValueError("dp must have nkinmnrl + neqcplx entries")
        n = st.ncell
        df = torch.zeros(n, dtype=torch.float64, device="cuda")
        f = torch.zeros(n, dtype=torch.float64, device="cuda")
        status = torch.zeros(n, dtype=torch.int32, device="cuda")
        s = st.to_abi()
        self._check(self.L.pfchem_tangent(self.h, n, C.byref(s), dt, wv.ctypes.data,
                                          ws.ctypes.data, dpv.ctypes.data, 1 if exact else 0,
                                          df.data_ptr(), f.data_ptr(), status.data_ptr(),
                                          _stream_ptr(stream)), "pfchem_tangent")
        return df, f, status

    def adjoint(self, st: DeviceState, dt: float, w_vf=None, w_sec=None, exact=True,
                per_cell=False, stream=None):
        """pfchem_adjoint: (gradient summed over cells (numpy), f per cell,
        status per cell, per-cell gradients or None)"""
        import torch
        wv, ws = self._weights(w_vf, w_sec)
        n = st.ncell
        np_ = wv.size + ws.size
        grad = np.zeros(max(np_, 1))
        f = torch.zeros(n, dtype=torch.float64, device="cuda")
        status = torch.zeros(n, dtype=torch.int32, device="cuda")
        gc = torch.zeros((np_, n), dtype=torch.float64, device="cuda") if per_cell else None
        s = st.to_abi()
        self._check(self.L.pfchem_adjoint(self.h, n, C.byref(s), dt, wv.ctypes.data,
                                          ws.ctypes.data, 1 if exact else 0, grad.ctypes.data,
                                          gc.data_ptr() if gc is not None else None,
                                          f.data_ptr(), status.data_ptr(),
                                          _stream_ptr(stream)), "pfchem_adjoint")
        return grad[:np_], f, status, gc

    def autotune(self, st: DeviceState, dt: float, variant: int = abi.KERNEL_BATCHED_MIXED,
                 cache_path: Optional[str] = None, table_path: Optional[str] = None,
                 force: bool = False, quick: bool = False) -> dict:
        """pfchem_autotune: time the step over the launch shapes the context
        accepts on these cells, keep the fastest that agrees with the current
        shape; returns the chosen shape (the state is left as given)"""
        import torch
        s = st.to_abi()
        c = abi.Launch()
        flags = (1 if force else 0) | (2 if quick else 0)
        sp = torch.cuda.current_stream().cuda_stream
        self._check(self.L.pfchem_autotune(self.h, variant, st.ncell, C.byref(s), dt,
                                           cache_path.encode() if cache_path else None,
                                           table_path.encode() if table_path else None,
                                           flags, C.byref(c), sp), "pfchem_autotune")
        return dict(lanes_per_cell=c.lanes_per_cell, rows_per_lane=c.rows_per_lane,
                    block_size=c.block_size, blocks_per_cu=c.blocks_per_cu)

    def transport_setup(self, ncell, id_up, id_dn, bc_cell):
        """pfchem_transport_setup: the connection topology (host arrays)"""
        up = np.ascontiguousarray(id_up, dtype=np.int64)
        dn = np.ascontiguousarray(id_dn, dtype=np.int64)
        bc = np.ascontiguousarray(bc_cell, dtype=np.int64)
        c = abi.Connections(up.size, up.ctypes.data if up.size else None,
                            dn.ctypes.data if dn.size else None, bc.size,
                            bc.ctypes.data if bc.size else None)
        self._check(self.L.pfchem_transport_setup(self.h, ncell, C.byref(c)),
                    "pfchem_transport_setup")

    def transport_step(self, st: DeviceState, coefs: dict, dt: float, rtol: float = 1e-12,
                       max_iterations: int = 1000, stream=None, check: bool = True):
        """pfchem_transport_step: one implicit transport step of the state's
        totals; coefs: device (torch) arrays area, velocity, coef_over_dist,
        bc_area, bc_velocity, bc_coef_over_dist, bc_total [naq, nbc].
        Returns (iterations, return code)"""
        s = st.to_abi()
        ptr = lambda k: coefs[k].data_ptr() if coefs.get(k) is not None and coefs[k].numel() else None
        tc = abi.TransportCoefs(*[ptr(k) for k in ("area", "velocity", "coef_over_dist", "bc_area",
                                                   "bc_velocity", "bc_coef_over_dist",
                                                   "bc_total")])
        it = C.c_int32(0)
        rc = self.L.pfchem_transport_step(self.h, st.ncell, C.byref(s), C.byref(tc), dt, rtol,
                                          max_iterations, C.byref(it), _stream_ptr(stream))
        if check:
            self._check(rc, "pfchem_transport_step")
        return it.value, rc

    def tune_key(self):
        """(key, bucket) of this context in a tuner cache"""
        k = C.create_string_buffer(256)
        b = C.create_string_buffer(256)
        self._check(self.L.pfchem_tune_key(self.h, k, 256, b, 256), "pfchem_tune_key")
        return k.value.decode(), b.value.decode()

    def get_launch(self, variant):
        c = abi.Launch()
        self._check(self.L.pfchem_get_launch(self.h, variant, C.byref(c)), "pfchem_get_launch")
        return dict(lanes_per_cell=c.lanes_per_cell, rows_per_lane=c.rows_per_lane,
                    block_size=c.block_size, blocks_per_cu=c.blocks_per_cu)
