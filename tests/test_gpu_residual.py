"""GPU reaction block of the global implicit residual and Jacobian
(pfchem_residual_jacobian: RTResidualNonFlux / RTJacobianNonFlux per cell,
reactive_transport.F90:2393-2430, 2599-2642, 3172-3183, 3270-3300, in the
RReact form of reaction.F90:3896-3916) against the FP64 oracle's
orc_total + orc_rreact_system on the same cells.

Networks: calcite, the random 15-primary network, and the ELM-CN SomDec
network (5 aqueous + 10 immobile species, the sandbox's reactions and N2O
emission).  Inputs: the workload's free molalities / immobile
concentrations, and fixed (previous-time) totals drawn around the cells'
current ones, so that the accumulation term does not cancel to zero; a few
cells are dry.

Bars: residual relative error <= 1e-9, with a floor of 1e-12 of the
magnitudes of the terms it sums (accumulation rate psv*total/dt and the
reaction term of that row: the residual is their difference); Jacobian entries <= 1e-11 of the row's largest entry;
totals and secondary molalities written back <= 1e-13 relative.
"""
import numpy as np
import pytest

import oracle_lib as O
from pfchem import abi, workloads

pytestmark = pytest.mark.gpu

DT = 3600.0


def _case(which, ncell, seed=7):
    if which == "calcite":
        wl = workloads.calcite_workload(ncell)
    elif which == "somdec":
        wl = workloads.somdec_workload(ncell=ncell)
    elif which == "somdec-oxygen":
        wl = workloads.somdec_workload(ncell=ncell, oxygen=1.0e-5)
    elif which == "somdec-complexes":
        wl = workloads.somdec_workload(ncell=ncell, complexes=True)
    elif which == "prefactor":
        import pin_util as PU
        wl = PU.prefactor_workload(ncell)
    else:
        wl = workloads.random_network_workload(ncell=ncell)
    st = wl.state.copy()
    rng = np.random.default_rng(seed)
    n = st.ncell
    st["saturation"][0, rng.choice(n, size=max(1, n // 100), replace=False)] = 0.0
    naq = wl.net.naqcomp
    fixed = st["total"][:naq].copy()
    fixed[:, :n] *= 1.0 + 1e-3 * rng.standard_normal(fixed[:, :n].shape)
    fimm = None
    if wl.net.nimmobile:
        fimm = st["immobile"][:wl.net.nimmobile].copy()
        fimm[:, :n] *= 1.0 + 1e-3 * rng.standard_normal(fimm[:, :n].shape)
    return wl, st, fixed, fimm


def _oracle(wl, st, fixed, fimm, cells, dt):
    net = wl.net
    ncomp = net.naqcomp + net.nimmobile
    res = np.zeros((ncomp, cells.size))
    jac = np.zeros((ncomp, ncomp, cells.size))
    tot = np.zeros((net.naqcomp, cells.size))
    sec = np.zeros((net.neqcplx, cells.size))
    acc = np.zeros((ncomp, cells.size))
    cx = O.CellCtx(net, wl.opt)
    for q, c in enumerate(cells):
        cc = cx.gather(st, int(c))
        cc.total()
        por, sat = st["porosity"][0, c], st["saturation"][0, c]
        vol = st["volume"][0, c]
        fa = np.zeros(ncomp)
        if sat >= 1e-40:
            psv = por * sat * 1000.0 * vol
            fa[:net.naqcomp] = psv * fixed[:, c]
            acc[:net.naqcomp, q] = np.abs(psv * cc.get("total", net.naqcomp)) / dt
            if net.nimmobile:
                fa[net.naqcomp:] = vol * fimm[:, c]
                acc[net.naqcomp:, q] = np.abs(vol * cc.get("immobile", net.nimmobile)) / dt
        r, J = cc.rreact_system(fa, dt)
        rx, _ = cc.reaction(dt, derivative=False)
        acc[:, q] += np.abs(rx)
        res[:, q] = r
        jac[:, :, q] = J
        tot[:, q] = cc.get("total", net.naqcomp)
        sec[:, q] = cc.get("sec_molal", net.neqcplx)
    return res, jac, tot, sec, acc


@pytest.mark.parametrize("which,ncell", [("calcite", 1000), ("random15", 2048), ("somdec", 1024),
                                         ("somdec-oxygen", 512), ("somdec-complexes", 512),
                                         ("prefactor", 512)])
def test_residual_jacobian_matches_oracle(which, ncell):
    import torch
    from pfchem.lib import Context, DeviceState
    wl, st, fixed, fimm = _case(which, ncell)
    dt = wl.dt if which.startswith("somdec") else DT
    print("case residual", which, ncell, flush=True)
    n = st.ncell
    dry_all = np.where(st["saturation"][0, :n] < 1e-40)[0]
    cells = np.arange(n) if n <= 1000 else np.unique(np.r_[np.arange(0, n, 3), n - 1, dry_all])
    r_ref, j_ref, t_ref, s_ref, acc = _oracle(wl, st, fixed, fimm, cells, dt)
    ctx = Context(wl.net, wl.opt)
    ds = DeviceState(st)
    ft = torch.from_numpy(np.ascontiguousarray(fixed)).cuda()
    fi = None if fimm is None else torch.from_numpy(np.ascontiguousarray(fimm)).cuda()
    res, jac = ctx.residual(ds, ft, fi, dt)
    torch.cuda.synchronize()
    res = res.cpu().numpy()[:, cells]
    ncomp = wl.net.naqcomp + wl.net.nimmobile
    jac = jac.cpu().numpy().reshape(ncomp, ncomp, -1)[:, :, cells]
    got = ds.to_host()
    d = np.abs(res - r_ref)
    floor = 1e-12 * acc
    rel = np.where(d == 0, 0.0, d / np.maximum(np.abs(r_ref), np.maximum(floor, 1e-300)))
    print(which, "residual max rel err %.3e" % rel.max(), flush=True)
    assert rel.max() <= 1e-9
    scale = np.abs(j_ref).max(axis=1, keepdims=True)
    scale = np.where(scale > 0, scale, 1.0)
    je = (np.abs(jac - j_ref) / scale).max()
    print(which, "jacobian max err / row scale %.3e" % je, flush=True)
    assert je <= 1e-11
    te = (np.abs(got["total"][:, cells] - t_ref) / np.abs(t_ref)).max()
    assert te <= 1e-13
    if wl.net.neqcplx:
        se = (np.abs(got["sec_molal"][:, cells] - s_ref) / np.abs(s_ref)).max()
        print(which, "totals %.3e sec %.3e" % (te, se), flush=True)
        assert se <= 1e-13
    # dry cells: zero residual, identity Jacobian
    dry = np.where(st["saturation"][0, cells] < 1e-40)[0]
    assert dry.size > 0
    assert np.all(res[:, dry] == 0.0)
    eye = np.eye(ncomp)[:, :, None]
    assert np.all(jac[:, :, dry] == eye)


def test_residual_vanishes_at_a_converged_step():
    """pfchem_step's FP64 answer, fed back with the step's initial totals as
    the fixed accumulation, leaves a residual at the Newton tolerance for the
    cells that took one sub-step (the two kernels share nothing but the
    network): largest entry <= 1e-10 of the cell's largest accumulation
    rate."""
    import torch
    from pfchem.lib import Context, DeviceState
    wl = workloads.random_network_workload(ncell=4096)
    print("case residual at converged step", flush=True)
    n = wl.state.ncell
    naq = wl.net.naqcomp
    ctx = Context(wl.net, wl.opt)
    ds = DeviceState(wl.state)
    # RStep solves against totals clamped at 1e-40 (reaction.F90:3617-3631)
    fixed = ds.t["total"][:naq].clone().clamp(min=1e-40)
    ctx.step(ds, wl.dt, abi.KERNEL_PLAIN_FP64)
    res, _ = ctx.residual(ds, fixed, None, wl.dt, jacobian=False)
    torch.cuda.synchronize()
    got = ds.to_host()
    one = (got["num_sub_steps"][0, :n] == 1) & (got["status"][0, :n] == 0)
    # a mineral that dissolved away in the step has no rate at the new state
    gone = np.any((got["mnrl_volfrac"][:, :n] == 0.0) & (wl.state["mnrl_volfrac"][:, :n] > 0.0),
                  axis=0)
    one &= ~gone
    assert one.mean() > 0.9
    psv = (wl.state["porosity"][0, :n] * wl.state["saturation"][0, :n] * 1000.0 *
           wl.state["volume"][0, :n])
    # per cell: largest residual entry over the largest accumulation rate
    # (single totals can cancel to ~0, e.g. the proton's)
    acc = (psv[None, :] * np.abs(got["total"][:naq, :n]) / wl.dt).max(axis=0)
    r = np.abs(res.cpu().numpy()[:naq, :n]).max(axis=0)
    rel = (r / np.maximum(acc, 1e-300))[one]
    print("max |res| / accumulation rate %.3e" % rel.max(), flush=True)
    assert rel.max() <= 1e-10


def test_residual_argument_checks():
    import torch
    from pfchem.lib import Context, DeviceState, PfchemError
    wl = workloads.calcite_workload(64)
    ctx = Context(wl.net, wl.opt)
    ds = DeviceState(wl.state)
    with pytest.raises(PfchemError):
        ctx.residual(ds, None, None, 1.0)
    ft = ds.t["total"][:wl.net.naqcomp].clone()
    with pytest.raises(PfchemError):
        ctx.residual(ds, ft, None, 0.0)
    sd = workloads.somdec_workload(ncell=64)
    ctx2 = Context(sd.net, sd.opt)
    ds2 = DeviceState(sd.state)
    with pytest.raises(PfchemError) as e:
        ctx2.residual(ds2, ds2.t["total"][:sd.net.naqcomp].clone(), None, 1800.0)
    assert e.value.code == abi.PFCHEM_ERR_INVALID
    torch.cuda.synchronize()


def test_residual_variable_size_cells():
    """Per-cell active primaries (64-primary network, nact ~ U{8..64}): each
    cell's block equals the oracle's on the network restricted to its
    active primaries (variable_util.sub_network); rows and columns beyond
    nact are zero."""
    import torch
    import variable_util as VU
    from pfchem.lib import Context, DeviceState
    from pfchem.network import CellState
    wl = workloads.variable_network_workload(ncell=256)
    print("case residual variable 256", flush=True)
    st = wl.state
    n = st.ncell
    naq = wl.net.naqcomp
    rng = np.random.default_rng(11)
    fixed = st["total"][:naq].copy()
    fixed[:, :n] *= 1.0 + 1e-3 * rng.standard_normal(fixed[:, :n].shape)
    ctx = Context(wl.net, wl.opt)
    ds = DeviceState(st)
    res, jac = ctx.residual(ds, torch.from_numpy(np.ascontiguousarray(fixed)).cuda(), None, wl.dt)
    torch.cuda.synchronize()
    res = res.cpu().numpy()
    jac = jac.cpu().numpy().reshape(naq, naq, -1)
    got = ds.to_host()
    nact_all = np.clip(st["ncomp_active"][0, :n], 1, naq)
    worst_r = worst_j = 0.0
    for c in range(0, n, 4):
        nact = int(nact_all[c])
        sub, kc, km = VU.sub_network(wl.net, nact)
        ss = CellState(sub, 1)
        for name, a in ss.arrays.items():
            if name not in st.arrays:
                continue
            kind = VU.ROWS.get(name)
            src = st.arrays[name]
            if kind == "aq":
                a[:, 0] = src[:nact, c]
            elif kind == "c":
                a[:, 0] = src[kc, c]
            elif kind == "m":
                a[:, 0] = src[km, c]
            elif a.shape[0]:
                a[:, 0] = src[: a.shape[0], c]
        opt = wl.opt.__class__(**{**wl.opt.__dict__, "variable_ncomp": False})
        cc = O.CellCtx(sub, opt).gather(ss, 0)
        cc.total()
        psv = (st["porosity"][0, c] * st["saturation"][0, c] * 1000.0 * st["volume"][0, c])
        r_ref, j_ref = cc.rreact_system(psv * fixed[:nact, c], wl.dt)
        rx, _ = cc.reaction(wl.dt, derivative=False)
        floor = 1e-12 * (np.abs(psv * cc.get("total", nact)) / wl.dt + np.abs(rx))
        d = np.abs(res[:nact, c] - r_ref)
        worst_r = max(worst_r, float((d / np.maximum(np.abs(r_ref), floor)).max()))
        scale = np.abs(j_ref).max(axis=1, keepdims=True)
        worst_j = max(worst_j, float((np.abs(jac[:nact, :nact, c] - j_ref) / scale).max()))
        assert np.all(res[nact:, c] == 0.0)
        assert np.all(jac[nact:, :, c] == 0.0) and np.all(jac[:, nact:, c] == 0.0)
        t_ref = cc.get("total", nact)  # totals can cancel: floor at 1e-13 of the largest
        np.testing.assert_allclose(got["total"][:nact, c], t_ref, rtol=1e-13,
                                   atol=1e-13 * np.abs(t_ref).max())
    print("variable: residual max rel err %.3e, jacobian %.3e" % (worst_r, worst_j), flush=True)
    assert worst_r <= 1e-9
    assert worst_j <= 1e-11
