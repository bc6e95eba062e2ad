"""GPU tangent and adjoint of the reaction step (pfchem_tangent /
pfchem_adjoint) against the FP64 oracle's sensitivities and central finite
differences of the oracle step.

Objective per cell: end-of-step mineral volume fractions plus the
secondary molality of CO2(aq) (complex 2 of the calcite network; complex 0
stands in on the random 15-primary network).  Parameters: log10 rate
constants of the kinetic minerals, log10 K of the complexes.
Bars: GPU vs oracle adjoint 1e-8 relative (same mathematics, FP64); adjoint
vs tangent 1e-9; adjoint vs central FD (h = 1e-6) 1e-6 relative for every
component above the FD noise floor.
"""
import dataclasses

import numpy as np
import pytest

import oracle_lib as O
from pfchem import abi, workloads

pytestmark = pytest.mark.gpu


def _setup(which, ncell):
    if which == "calcite":
        wl = workloads.calcite_workload(ncell)
        co2 = 2
    else:
        wl = workloads.random_network_workload(ncell=ncell)
        co2 = 0
    net = wl.net
    wv = np.ones(net.nkinmnrl)
    ws = np.zeros(net.neqcplx)
    ws[co2] = 1.0
    return wl, wv, ws


@pytest.mark.parametrize("which,ncell", [("calcite", 2048), ("random15", 1024)])
def test_adjoint_matches_oracle(which, ncell):
    from pfchem.lib import Context, DeviceState
    wl, wv, ws = _setup(which, ncell)
    print("case sens", which, ncell, flush=True)
    g_ref, f_ref, st_ref = O.adjoint_batch(wl.net, wl.opt, wl.state, wl.dt, wv, ws)
    ok = st_ref == 0
    assert ok.mean() > 0.95
    ctx = Context(wl.net, wl.opt)
    ds = DeviceState(wl.state)
    g, f, st, gc = ctx.adjoint(ds, wl.dt, wv, ws, per_cell=True)
    st = st.cpu().numpy()
    f = f.cpu().numpy()
    np.testing.assert_array_equal(st == 0, ok)
    np.testing.assert_allclose(f[ok], f_ref[ok], rtol=1e-8)
    scale = np.abs(g_ref).max()
    np.testing.assert_allclose(g / scale, g_ref / scale, atol=1e-8)
    # the per-cell gradients sum to the reduced one
    np.testing.assert_allclose(gc.cpu().numpy()[:, ok].sum(axis=1) / scale, g / scale, atol=1e-12)
    # tangent along a random direction: df = grad . dp, summed over cells
    dp = np.random.default_rng(3).standard_normal(g.size)
    df, f2, st2 = ctx.tangent(ds, wl.dt, dp, wv, ws)
    df_sum = df.cpu().numpy()[ok].sum()
    assert df_sum == pytest.approx(float(g @ dp), rel=1e-9, abs=1e-12 * scale)


def test_adjoint_vs_central_differences():
    """finite differences of the oracle step, h = 1e-6 in log10 p; cells whose
    sub-step path changes under the perturbation are left out (not
    differentiable there)"""
    from pfchem.lib import Context, DeviceState
    wl, wv, ws = _setup("calcite", 512)
    net, opt = wl.net, wl.opt
    opt = dataclasses.replace(opt, max_rel_residual_tolerance=1e-13,
                              max_relative_change_tolerance=1e-14, maximum_reaction_iterations=100)
    nm, nc = net.nkinmnrl, net.neqcplx

    def run(net2):
        s = wl.state.copy()
        O.step_batch(net2, opt, s, wl.dt, 0)
        n = s.ncell
        # VF increments from the rates (VF itself is ~0.1, so the change
        # would be lost to rounding), clipped minerals at -VF0
        dv = s.arrays["mnrl_rate"][:, :n] * net2.kinmnrl_molar_vol[:, None] * wl.dt
        dv = np.where(s.arrays["mnrl_volfrac"][:, :n] == 0.0,
                      -wl.state.arrays["mnrl_volfrac"][:, :n], dv)
        f = (wv[:, None] * dv).sum(0) + (ws[:, None] * s.arrays["sec_molal"][:, :n]).sum(0)
        return f, s.arrays["num_sub_steps"][0, :n]

    _, ns0 = run(net)
    keep = ns0 == 1
    h = 1e-6
    fds = []
    for q in range(nm + nc):
        r = []
        for sg in (1, -1):
            if q < nm:
                n2 = dataclasses.replace(net, kinmnrl_rate_constant=net.kinmnrl_rate_constant *
                                         10 ** (sg * h * (np.arange(nm) == q)))
            else:
                n2 = dataclasses.replace(net, eqcplx_logK=net.eqcplx_logK +
                                         sg * h * (np.arange(nc) == q - nm))
            r.append(run(n2))
        keep &= (r[0][1] == 1) & (r[1][1] == 1)
        fds.append((r[0][0] - r[1][0]) / (2 * h))
    idx = np.where(keep)[0]
    assert idx.size > 0.9 * wl.state.ncell
    sub = wl.state.subset(idx)
    ctx = Context(net, opt)
    g, f, st, _ = ctx.adjoint(DeviceState(sub), wl.dt, wv, ws)
    assert int((st.cpu().numpy() != 0).sum()) == 0
    fd = np.array([x[idx].sum() for x in fds])
    big = np.abs(fd) > 1e-6 * np.abs(fd).max()
    rel = np.abs(g - fd) / np.abs(fd)
    print("adjoint vs FD: max rel %.3e over %d components" % (rel[big].max(), big.sum()))
    assert np.all(rel[big] < 1e-6)
    # the reference's analytic Jacobian (mineral block x rho/1000) moves the
    # sensitivities by a visible amount only through the mineral terms
    g_ref, _, _, _ = ctx.adjoint(DeviceState(sub), wl.dt, wv, ws, exact=False)
    assert np.all(np.abs(g_ref - g) <= 1e-2 * np.abs(g) + 1e-30)


def test_adjoint_hand_over_to_the_general_kernel():
    """cells whose secondary activity coefficients do not follow the
    charge / ion-size classes (here every 5th cell, one complex's
    coefficient changed) are handed from the one-lane adjoint to the general
    kernel; every cell still matches the oracle's adjoint"""
    from pfchem.lib import Context, DeviceState
    wl, wv, ws = _setup("random15", 1024)
    st0 = wl.state.copy()
    n = st0.ncell
    net = wl.net
    k = int(np.flatnonzero(np.abs(net.eqcplx_Z) > 0)[0])
    st0.arrays["sec_act_coef"][k, 0:n:5] *= 1.01
    print("case sens hand-over", flush=True)
    g_ref, f_ref, st_ref = O.adjoint_batch(net, wl.opt, st0, wl.dt, wv, ws)
    ok = st_ref == 0
    ctx = Context(net, wl.opt)
    g, f, st, gc = ctx.adjoint(DeviceState(st0), wl.dt, wv, ws, per_cell=True)
    st = st.cpu().numpy()
    np.testing.assert_array_equal(st == 0, ok)
    np.testing.assert_allclose(f.cpu().numpy()[ok], f_ref[ok], rtol=1e-8)
    scale = np.abs(g_ref).max()
    np.testing.assert_allclose(g / scale, g_ref / scale, atol=1e-8)
