"""pfchem_equilibrate_static (the allocation-free constraint speciation,
staging in the library's fixed device array) against pfchem_equilibrate on
the same inputs, through the C ABI: two calls in a row, the second with a
smaller input (fewer cells and a smaller reaction table), each equal to the
allocating entry point's answer bit for bit (the same kernel on the same
constraint image); a constraint whose staging exceeds the array is refused
before any launch."""
import ctypes as C

import numpy as np
import pytest

import test_oracle_constraint as T
from test_gpu_constraint import _case

from pfchem import abi

pytestmark = pytest.mark.gpu

FIELDS = ("pri_molal", "total", "sec_molal", "pri_act_coef", "sec_act_coef", "num_iterations",
          "status")


def _both(ctx, net, con, v, n):
    """(allocating result, static result, return codes) on fresh copies of
    the same n-cell state"""
    import torch
    from pfchem.lib import DeviceState
    out, rcs = [], []
    for fn in (ctx.L.pfchem_equilibrate, ctx.L.pfchem_equilibrate_static):
        ds = DeviceState(T._state(net, n))
        conc = torch.from_numpy(np.ascontiguousarray(v[:, :n])).cuda()
        s = ds.to_abi()
        k = con.to_abi(net)
        sp = torch.cuda.current_stream().cuda_stream
        rcs.append(fn(ctx.h, n, C.byref(s), C.byref(k), conc.data_ptr(), sp))
        torch.cuda.synchronize()
        out.append(ds.to_host())
    return out[0], out[1], rcs


def test_static_equals_allocating_twice():
    from pfchem.lib import Context
    print("case equilibrate_static two calls", flush=True)
    rng = np.random.default_rng(5)
    # first call: gas + mineral constraints (two reaction rows), 2048 cells
    net, opt, con, v = _case("free-gas-mineral", rng)
    ctx = Context(net, opt)
    ref, got, rcs = _both(ctx, net, con, v, 2048)
    assert rcs == [abi.PFCHEM_OK, abi.PFCHEM_OK], rcs
    for f in FIELDS:
        np.testing.assert_array_equal(got[f][:, :2048], ref[f][:, :2048], err_msg=f)
    assert np.all(got["status"][0, :2048] == 0)
    # second call, same context: totals only (no reaction rows), 300 cells;
    # the array keeps the first call's contents, which must not leak in
    net2, opt2, con2, v2 = _case("ca-carbonate-totals", rng)
    ctx2 = Context(net2, opt2)
    ref2, got2, rcs2 = _both(ctx2, net2, con2, v2, 300)
    assert rcs2 == [abi.PFCHEM_OK, abi.PFCHEM_OK], rcs2
    for f in FIELDS:
        np.testing.assert_array_equal(got2[f][:, :300], ref2[f][:, :300], err_msg=f)
    # a failing call reports its own failures only (the failure count is
    # cleared per call): 2 iterations leave every cell unconverged
    con2.max_iterations = 2
    ref3, got3, rcs3 = _both(ctx2, net2, con2, v2, 64)
    assert rcs3 == [abi.PFCHEM_ERR_CELLS, abi.PFCHEM_ERR_CELLS], rcs3
    np.testing.assert_array_equal(got3["status"][:, :64], ref3["status"][:, :64])
    msg = ctx2.L.pfchem_last_error(ctx2.h).decode()
    assert msg.startswith("64 cell(s)"), msg


def test_static_refuses_oversized_table():
    """a reaction table whose staging exceeds 32 KiB is refused with
    PFCHEM_ERR_UNSUPPORTED before any launch"""
    import torch
    from pfchem.lib import Context, DeviceState
    print("case equilibrate_static oversized", flush=True)
    rng = np.random.default_rng(9)
    net, opt, con, v = _case("free-gas-mineral", rng)
    ctx = Context(net, opt)
    k = con.to_abi(net)
    nr, ms = 400, 3
    ids = np.zeros((nr, ms + 1), dtype=np.int32)
    ids[:, 0] = 1
    st = np.ones((nr, ms))
    lk = np.zeros(nr)
    k.nrxn = nr
    k.rxn_maxsp = ms
    k.rxnspecid = ids.ctypes.data_as(C.POINTER(C.c_int32))
    k.rxnstoich = st.ctypes.data_as(C.POINTER(C.c_double))
    k.rxn_logK = lk.ctypes.data_as(C.POINTER(C.c_double))
    k.rxnh2ostoich = C.cast(None, C.POINTER(C.c_double))
    k.rxn_logKcoef = C.cast(None, C.POINTER(C.c_double))
    ds = DeviceState(T._state(net, 16))
    conc = torch.from_numpy(np.ascontiguousarray(v[:, :16])).cuda()
    s = ds.to_abi()
    rc = ctx.L.pfchem_equilibrate_static(ctx.h, 16, C.byref(s), C.byref(k), conc.data_ptr(), None)
    assert rc == abi.PFCHEM_ERR_UNSUPPORTED
    assert "32768" in ctx.L.pfchem_last_error(ctx.h).decode()
