"""The oracle's step sensitivities (oracle/pfchem_oracle_adjoint.c) against
central finite differences of the oracle step itself (h = 1e-6 in log10 p)
and against each other (adjoint vs tangent): the checks the GPU kernels are
then held to."""
import dataclasses

import numpy as np

import oracle_lib as O
from pfchem import workloads


def _fd(wl, opt, wv, ws, h=1e-6):
    net = wl.net
    nm, nc = net.nkinmnrl, net.neqcplx

    def run(net2):
        s = wl.state.copy()
        O.step_batch(net2, opt, s, wl.dt, 0)
        n = s.ncell
        dv = s.arrays["mnrl_rate"][:, :n] * net2.kinmnrl_molar_vol[:, None] * wl.dt
        dv = np.where(s.arrays["mnrl_volfrac"][:, :n] == 0.0,
                      -wl.state.arrays["mnrl_volfrac"][:, :n], dv)
        return ((wv[:, None] * dv).sum(0) + (ws[:, None] * s.arrays["sec_molal"][:, :n]).sum(0),
                s.arrays["num_sub_steps"][0, :n])

    _, ns0 = run(net)
    keep = ns0 == 1
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
    return idx, np.array([x[idx].sum() for x in fds])


def test_oracle_adjoint_vs_finite_differences():
    wl = workloads.calcite_workload(256)
    opt = dataclasses.replace(wl.opt, max_rel_residual_tolerance=1e-13,
                              max_relative_change_tolerance=1e-14,
                              maximum_reaction_iterations=100)
    wv = np.ones(wl.net.nkinmnrl)
    ws = np.zeros(wl.net.neqcplx)
    ws[2] = 1.0                                   # CO2(aq)
    idx, fd = _fd(wl, opt, wv, ws)
    assert idx.size > 0.9 * wl.state.ncell
    g, f, st = O.adjoint_batch(wl.net, opt, wl.state.subset(idx), wl.dt, wv, ws)
    assert np.all(st == 0)
    big = np.abs(fd) > 1e-6 * np.abs(fd).max()
    assert np.all(np.abs(g - fd)[big] < 1e-6 * np.abs(fd)[big])


def test_oracle_adjoint_equals_tangent_and_catches_a_dropped_term():
    wl = workloads.random_network_workload(ncell=128)
    net = wl.net
    wv = np.ones(net.nkinmnrl)
    ws = np.zeros(net.neqcplx)
    ws[0] = 1.0
    g, f, st = O.adjoint_batch(net, wl.opt, wl.state, wl.dt, wv, ws)
    ok = st == 0
    assert ok.mean() > 0.95
    rng = np.random.default_rng(5)
    for _ in range(3):
        dp = rng.standard_normal(g.size)
        df, f2, st2 = O.tangent_batch(net, wl.opt, wl.state, wl.dt, wv, ws, dp)
        np.testing.assert_allclose(df[ok].sum(), g @ dp, rtol=1e-9)
        np.testing.assert_array_equal(f2, f)
    # the reference's Jacobian convention (mineral block x rho/1000) is a
    # different, slightly wrong, linearization: the check above must see it
    dp = np.zeros(g.size)
    dp[:net.nkinmnrl] = 1.0
    d_exact, _, _ = O.tangent_batch(net, wl.opt, wl.state, wl.dt, wv, ws, dp, exact=True)
    d_ref, _, _ = O.tangent_batch(net, wl.opt, wl.state, wl.dt, wv, ws, dp, exact=False)
    assert not np.allclose(d_exact[ok], d_ref[ok], rtol=1e-12, atol=0)
