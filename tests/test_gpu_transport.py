"""GPU OSRT transport step (pfchem_transport_setup / pfchem_transport_step)
against the transport oracle (oracle/oracle_transport.py: the reference's
RTCalculateTransportMatrix / RTCalculateRHS_t1 / TFluxCoef system solved
directly) and against closed forms.  Bar: 1e-10 relative to the largest
total of each component (the device solves iteratively to a relative
residual of 1e-13)."""
import numpy as np
import pytest

from pfchem import abi, workloads
from pfchem.lib import Context, DeviceState, PfchemError

import transport_util as TU

pytestmark = pytest.mark.gpu


def _run(c, naq_net, rtol=1e-13):
    import torch
    wl = naq_net
    st = wl.state.__class__(wl.net, c["ncell"])
    n = c["ncell"]
    st["total"][:, :n] = c["total"]
    st["porosity"][0, :n] = c["por"]
    st["saturation"][0, :n] = c["sat"]
    st["volume"][0, :n] = c["vol"]
    ctx = Context(wl.net, wl.opt)
    ctx.transport_setup(n, c["id_up"], c["id_dn"], c["bc_cell"])
    ds = DeviceState(st)
    it, _ = ctx.transport_step(ds, TU.device_coefs(c), c["dt"], rtol=rtol)
    torch.cuda.synchronize()
    return ds.to_host()["total"][:, :n], it


def _net(naq):
    wl = workloads.random_network_workload(ncell=8)
    assert wl.net.naqcomp == naq
    return wl


@pytest.mark.parametrize("shape,dry", [((20, 16, 12), 0.0), ((40, 8, 4), 0.02)])
def test_transport_matches_oracle(shape, dry):
    wl = _net(15)
    c = TU.box_case(*shape, 15, seed=7, dry_frac=dry)
    print("case transport", shape, dry, flush=True)
    ref = TU.oracle_step(c)
    got, it = _run(c, wl)
    err = (np.abs(got - ref) / np.abs(ref).max(axis=1, keepdims=True)).max()
    print("transport", shape, "iterations", it, "max err", err, flush=True)
    assert err <= 1e-10


def test_transport_advection_closed_form():
    """one implicit upwind step through a column from c = 0 with inflow c_b:
    c_i = c_b (Cr/(1+Cr))^(i+1) for every component"""
    wl = _net(15)
    n, q = 64, 2e-5
    c = TU.box_case(n, 1, 1, 15, seed=1)
    c["vel"][:] = q
    c["cod"][:] = 0.0
    c["bc_cod"][:] = 0.0
    c["bc_vel"][:] = [q, -q]
    c["bc_total"][:, 0] = 1e-3
    c["por"][:] = 0.25
    c["sat"][:] = 1.0
    c["total"][:] = 0.0
    got, _ = _run(c, wl)
    cr = q * c["dt"] / 0.25
    want = 1e-3 * (cr / (1.0 + cr)) ** np.arange(1, n + 1)
    # iterative solution: error relative to the largest value (the tail of
    # the geometric profile reaches 1e-42)
    np.testing.assert_allclose(got, np.broadcast_to(want, got.shape), rtol=0,
                               atol=1e-10 * want.max())
    np.testing.assert_allclose(got[:, :8], np.broadcast_to(want[:8], (got.shape[0], 8)),
                               rtol=1e-9)


def test_transport_conserves_mass_without_boundaries():
    wl = _net(15)
    c = TU.box_case(12, 10, 8, 15, seed=5, dry_frac=0.01)
    c["bc_cell"] = c["bc_cell"][:0]
    for k in ("bc_area", "bc_vel", "bc_cod"):
        c[k] = c[k][:0]
    c["bc_total"] = c["bc_total"][:, :0]
    got, _ = _run(c, wl)
    psv = c["por"] * c["sat"] * 1000.0 * c["vol"]
    np.testing.assert_allclose((psv * got).sum(axis=1), (psv * c["total"]).sum(axis=1),
                               rtol=1e-11)


def test_transport_argument_checks():
    wl = _net(15)
    ctx = Context(wl.net, wl.opt)
    ds = DeviceState(wl.state)
    c = TU.box_case(2, 2, 2, 15)
    with pytest.raises(PfchemError) as e:   # no set-up yet
        ctx.transport_step(ds, TU.device_coefs(c), 3600.0)
    assert e.value.code == abi.PFCHEM_ERR_INVALID
    with pytest.raises(PfchemError) as e:   # a cell id out of range
        ctx.transport_setup(8, np.array([0, 9]), np.array([1, 2]), np.array([], np.int64))
    assert e.value.code == abi.PFCHEM_ERR_INVALID
