"""The network-specialized one-lane kernel's generated source, compiled for the
host (one thread plays a one-lane wavefront, tests/lane_host), against the FP64
oracle on the same seeded cells.  This checks the generated arithmetic, the
RStep / RReact state machine (reaction.F90:3564-4055), the FP32 hand-over to
the FP64 pass and the write-out without a GPU; the GPU tests check the device
build of the same source.

Bar: relative concentration error <= 1e-8 against the oracle (measured:
calcite 7e-15 FP64 / 2e-11 mixed; 15-primary network 2e-12 / 5e-11), the
oracle's FP64 relative residual of the answer <= 1e-10 (or the oracle's own),
and the FP64 pass takes the oracle's Newton iteration counts (the mixed pass
converges linearly near the root, so its counts may differ by one).
"""
import tempfile

import numpy as np
import pytest

from pfchem import abi, workloads, lib
from pfchem.network import SolverOptions

import gpu_util as G
import lane_util as LU

CONC_TOL = 1e-8


def _check(wl, ref, rec, got, label, iters_equal):
    n = wl.state.ncell
    assert np.all(got["status"][0, :n] == 0), label
    errs = G.compare(ref, got, n, net=wl.net)
    print(label, errs)
    for f, e in errs.items():
        assert e <= CONC_TOL, (label, f, e)
    rr, bar = G.gate_rel_residual(wl, got, rec, range(n), ref)
    assert np.all(rr <= bar), label
    same = np.mean(got["num_iterations"][0, :n] == ref["num_iterations"][0, :n])
    assert same >= iters_equal, (label, same)


@pytest.fixture(scope="module")
def calcite_host():
    wl = workloads.calcite_workload(1000)
    ref, rec, nfail = G.run_oracle(wl)
    assert nfail == 0
    exe = LU.build_host(wl.net, wl.opt, tempfile.mkdtemp())
    return wl, ref, rec, exe


def test_calcite_fp64(calcite_host):
    wl, ref, rec, exe = calcite_host
    got = LU.run_host(exe, wl.state, wl.dt, f32=False)
    _check(wl, ref, rec, got, "calcite/lane-fp64", 0.999)


def test_calcite_mixed(calcite_host):
    wl, ref, rec, exe = calcite_host
    got = LU.run_host(exe, wl.state, wl.dt, f32=True)
    # the FP32-preconditioned iteration converges linearly: counts may differ
    _check(wl, ref, rec, got, "calcite/lane-mixed", 0.9)


def test_random15_both(tmp_path):
    wl = workloads.random_network_workload(ncell=2048)
    ref, rec, nfail = G.run_oracle(wl)
    assert nfail == 0
    exe = LU.build_host(wl.net, wl.opt, str(tmp_path))
    _check(wl, ref, rec, LU.run_host(exe, wl.state, wl.dt, f32=False), "random15/lane-fp64",
           0.995)
    _check(wl, ref, rec, LU.run_host(exe, wl.state, wl.dt, f32=True), "random15/lane-mixed",
           0.9)


def test_hand_over_without_later_pass(calcite_host):
    """with no later pass (escalate = 0), a cell whose FP32 Newton does not
    converge within the iteration limit is cut like the reference's and
    finishes: a 1-iteration limit forces every cell through the cuts"""
    import dataclasses
    wl, ref, rec, _ = calcite_host
    opt = dataclasses.replace(wl.opt, maximum_reaction_iterations=1, maximum_reaction_cuts=2)
    wl2 = dataclasses.replace(wl, opt=opt)
    st = wl.state.subset(np.arange(64))
    ref2 = st.copy()
    import oracle_lib as O
    O.step_batch(wl.net, opt, ref2, wl.dt, 0)
    exe = LU.build_host(wl.net, opt, tempfile.mkdtemp())
    got = LU.run_host(exe, st, wl.dt, f32=False, escalate=False)
    np.testing.assert_array_equal(got["status"][0], ref2["status"][0])
    np.testing.assert_array_equal(got["num_iterations"][0], ref2["num_iterations"][0])
    for f in ("pri_molal", "total", "sec_molal", "mnrl_volfrac"):
        np.testing.assert_allclose(got[f], ref2[f], rtol=1e-12, atol=0)


def test_rreact_call_bound(tmp_path):
    """the generated source stops RStep after options.maximum_rreact_calls
    RReact calls as the oracle does: the same cells fail, the others end as
    the oracle's (the FP64 pass; no later pass)"""
    import dataclasses
    import oracle_lib as O
    wl = workloads.random_network_workload(ncell=256)
    n = wl.state.ncell
    opt = dataclasses.replace(wl.opt, maximum_reaction_iterations=5, maximum_rreact_calls=3)
    ref = wl.state.copy()
    O.step_batch(wl.net, opt, ref, wl.dt, 0)
    ok = ref["status"][0, :n] == 0
    assert 20 < ok.sum() < n
    exe = LU.build_host(wl.net, opt, str(tmp_path))
    got = LU.run_host(exe, wl.state, wl.dt, f32=False, escalate=False)
    np.testing.assert_array_equal(got["status"][0, :n] != 0, ~ok)
    np.testing.assert_array_equal(got["num_sub_steps"][0, ok], ref["num_sub_steps"][0, ok])
    for f in ("pri_molal", "total", "sec_molal", "mnrl_volfrac"):
        a, b = ref[f][:, :n][:, ok], got[f][:, :n][:, ok]
        assert np.max(np.abs(b - a) / np.maximum(np.abs(a), 1e-300)) <= CONC_TOL, f


def test_prepare_refuses_other_networks():
    """the one-lane kernel takes aqueous networks of <= 16 primaries without
    the sandbox, per-cell active counts or rate prefactors (whose Jacobian
    terms are not symmetric); the others report UNSUPPORTED"""
    wl = workloads.variable_network_workload(ncell=8)
    assert lib.prepare_kernels(wl.net, wl.opt) == abi.PFCHEM_ERR_UNSUPPORTED
    wl = workloads.somdec_workload(ncell=8)
    assert lib.prepare_kernels(wl.net, wl.opt) == abi.PFCHEM_ERR_UNSUPPORTED
    import pin_util as PU
    wl = PU.prefactor_workload(8)
    assert lib.prepare_kernels(wl.net, wl.opt) == abi.PFCHEM_ERR_UNSUPPORTED


# the paths the pin tests cover: decay, temperature-dependent log K with
# Arrhenius, and the kinetic-mineral rate options
import pin_util as PU


def _run_pin_case(wl, label, iters_equal=0.99):
    ref, rec, nfail = G.run_oracle(wl)
    assert nfail == 0
    exe = LU.build_host(wl.net, wl.opt, tempfile.mkdtemp())
    for f32 in (False, True):
        got = LU.run_host(exe, wl.state, wl.dt, f32=f32)
        _check(wl, ref, rec, got, label + ("/mixed" if f32 else "/fp64"),
               iters_equal if not f32 else 0.9)
        for f in ("immobile",):
            if wl.net.nimmobile:
                e = float(G.rel_err(got[f], ref[f]).max())
                assert e <= CONC_TOL, (label, f, e)


def test_decay_network():
    wl, _ = PU.decay_workload(256)
    _run_pin_case(wl, "decay")


def test_nonisothermal_network():
    _run_pin_case(PU.nonisothermal_workload(256), "non-isothermal")


@pytest.mark.parametrize("opts", [
    dict(kinmnrl_Temkin_const=2.0, kinmnrl_min_scale_factor=1.5),
    dict(kinmnrl_affinity_power=1.7),
    dict(kinmnrl_rate_limiter=3.0, kinmnrl_affinity_threshold=2.0, kinmnrl_irreversible=1),
], ids=["temkin+scale", "power", "limiter+threshold+irreversible"])
def test_mineral_option_networks(opts):
    _run_pin_case(PU.mineral_options_workload(256, **opts), "options")


# the network-specialized adjoint (PFLANE_ADJOINT module), host build, against
# the oracle's adjoint (implicit-function theorem through the recorded
# sub-steps); bar 1e-10 of the largest gradient entry (measured ~1e-15)
@pytest.mark.parametrize("which", ["calcite", "random15"])
@pytest.mark.parametrize("exact", [True, False], ids=["exact", "reference-jacobian"])
def test_adjoint_host(which, exact, tmp_path):
    import oracle_lib as O
    if which == "calcite":
        wl, co2 = workloads.calcite_workload(1000), 2
    else:
        wl, co2 = workloads.random_network_workload(ncell=512), 0
    net = wl.net
    wv = np.ones(net.nkinmnrl)
    ws = np.zeros(net.neqcplx)
    ws[co2] = 1.0
    g_ref, f_ref, st_ref = O.adjoint_batch(net, wl.opt, wl.state, wl.dt, wv, ws, exact)
    exe = LU.build_host_adjoint(net, wl.opt, str(tmp_path))
    g, f, st = LU.run_host_adjoint(exe, net, wl.state, wl.dt, wv, ws, exact)
    np.testing.assert_array_equal(st, st_ref)
    ok = st == 0
    assert ok.mean() > 0.95
    np.testing.assert_allclose(f[ok], f_ref[ok], rtol=1e-12)
    scale = np.abs(g_ref).max()
    err = np.abs(g[:, ok].sum(axis=1) - g_ref).max() / scale
    print(which, exact, "adjoint host vs oracle", err)
    assert err < 1e-10

