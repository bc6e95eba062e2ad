"""GPU parity on the paths the oracle pin tests cover (tests/test_oracle_pins.py):
immobile species with first-order decay (RImmobileDecay,
reaction_immobile.F90:244-294), temperature-dependent log K with an Arrhenius
rate constant (ReactionInterpolateLogK, reaction_aux.F90:1285-1312), the
kinetic-mineral rate options (reaction_mineral.F90:729-908) and rate
prefactors (:820-870, 964-1033; the one-lane kernel does not take them, so
the batched variants run the team kernel), and
RActivityCoefficients LAG through pfchem_update_activity
(reaction.F90:4555-4611), and the bound on RReact calls per RStep
(options.maximum_rreact_calls).  Every kernel variant and both batched kernels (the
one-lane kernel, the default, and a 16x1 team) run through the C ABI against
the FP64 oracle on the same seeded cells.

Bar: relative concentration error <= 1e-8 against the oracle, FP64 relative
residual of the answer <= 1e-10 (or twice the oracle's own where rounding
keeps it above), activity coefficients <= 1e-12.
"""
import numpy as np
import pytest

import oracle_lib as O
from pfchem import abi, workloads
from pfchem.lib import Context, DeviceState

import gpu_util as G
import pin_util as PU

pytestmark = pytest.mark.gpu

CONC_TOL = 1e-8
# (variant, lanes per cell: None = the context's default)
KERNELS = [(abi.KERNEL_PLAIN_FP64, None), (abi.KERNEL_BATCHED_FP64, None),
           (abi.KERNEL_BATCHED_MIXED, None), (abi.KERNEL_BATCHED_FP64, 16),
           (abi.KERNEL_BATCHED_MIXED, 16)]
IDS = ["plain", "lane-fp64", "lane-mixed", "team16-fp64", "team16-mixed"]


def _parity(wl, variant, lanes, label, residual=True):
    ref, rec, nfail = G.run_oracle(wl)
    assert nfail == 0
    print("case", label, flush=True)
    got, _ = G.run_gpu(wl, variant, lanes=lanes)
    n = wl.state.ncell
    assert np.all(got["status"][0, :n] == 0), label
    fields = G.FIELDS_F64 if wl.net.nimmobile else tuple(f for f in G.FIELDS_F64
                                                         if f != "immobile")
    errs = G.compare(ref, got, n, fields, net=wl.net)
    print(label, errs, flush=True)
    for f, e in errs.items():
        assert e <= CONC_TOL, (label, f, e)
    if residual:
        rr, bar = G.gate_rel_residual(wl, got, rec, range(n), ref)
        assert rr.size > 0.8 * n and np.all(rr <= bar), label


@pytest.mark.parametrize("variant,lanes", KERNELS, ids=IDS)
def test_decay(variant, lanes):
    wl, _ = PU.decay_workload(512)
    _parity(wl, variant, lanes, "decay/" + IDS[KERNELS.index((variant, lanes))])


@pytest.mark.parametrize("variant,lanes", KERNELS, ids=IDS)
def test_nonisothermal(variant, lanes):
    wl = PU.nonisothermal_workload(512)
    _parity(wl, variant, lanes, "nonisothermal/" + IDS[KERNELS.index((variant, lanes))])


@pytest.mark.parametrize("variant,lanes", KERNELS, ids=IDS)
def test_prefactors(variant, lanes):
    wl = PU.prefactor_workload(512)
    _parity(wl, variant, lanes, "prefactor/" + IDS[KERNELS.index((variant, lanes))])


OPTION_SETS = [dict(kinmnrl_Temkin_const=2.0, kinmnrl_min_scale_factor=1.5),
               dict(kinmnrl_affinity_power=1.7),
               dict(kinmnrl_rate_limiter=3.0, kinmnrl_affinity_threshold=2.0,
                    kinmnrl_irreversible=1)]


@pytest.mark.parametrize("opts", OPTION_SETS, ids=["temkin+scale", "power", "limiter+cutoffs"])
@pytest.mark.parametrize("variant,lanes", KERNELS, ids=IDS)
def test_mineral_options(variant, lanes, opts):
    wl = PU.mineral_options_workload(512, **opts)
    _parity(wl, variant, lanes, "options/" + IDS[KERNELS.index((variant, lanes))])


def test_update_activity_lag():
    """pfchem_update_activity (LAG) against the oracle's
    RActivityCoefficients on the 15-primary network's cells"""
    import torch
    wl = workloads.random_network_workload(ncell=4096)
    st = wl.state.copy()
    # move the state away from the coefficients it was built with
    st["pri_molal"][:, :] *= 1.3
    ref = st.copy()
    O.update_activity_batch(wl.net, wl.opt, ref)
    ctx = Context(wl.net, wl.opt)
    ds = DeviceState(st)
    ctx.update_activity(ds)
    torch.cuda.synchronize()
    got = ds.to_host()
    for f in ("pri_act_coef", "sec_act_coef"):
        e = float(G.rel_err(got[f], ref[f]).max())
        print("update_activity", f, e, flush=True)
        assert e <= 1e-12, (f, e)


def test_update_activity_newton():
    """pfchem_update_activity with the NEWTON algorithm (reaction.F90:4403-
    4551) against the oracle: coefficients and the secondary molalities the
    fixed point rewrites"""
    import dataclasses
    import torch
    wl = workloads.random_network_workload(ncell=4096)
    opt = dataclasses.replace(wl.opt, act_coef_update_algorithm=abi.ACT_COEF_ALGORITHM_NEWTON)
    st = wl.state.copy()
    st["pri_molal"][:, :] *= 1.3
    ref = st.copy()
    O.update_activity_batch(wl.net, opt, ref)
    ctx = Context(wl.net, opt)
    ds = DeviceState(st)
    ctx.update_activity(ds)
    torch.cuda.synchronize()
    got = ds.to_host()
    # cells where the fixed point fails (the reference's NaN outputs,
    # reaction.F90:4420-4431) must be the same cells
    fail_ref = np.isnan(ref["pri_act_coef"]).any(axis=0)
    fail_gpu = np.isnan(got["pri_act_coef"]).any(axis=0)
    print("update_activity NEWTON failed cells", int(fail_ref.sum()), int(fail_gpu.sum()), flush=True)
    np.testing.assert_array_equal(fail_gpu, fail_ref)
    ok = ~fail_ref
    for f in ("pri_act_coef", "sec_act_coef", "sec_molal"):
        e = float(G.rel_err(got[f][:, ok], ref[f][:, ok]).max())
        print("update_activity NEWTON", f, e, flush=True)
        assert e <= 1e-12, (f, e)


def test_step_newton_activity_per_iteration():
    """RStep with RActivityCoefficients NEWTON in every Newton iteration: the
    plain kernel against the oracle; the batched kernels refuse the option"""
    import dataclasses
    wl = workloads.calcite_workload(1000)
    opt = dataclasses.replace(wl.opt, act_coef_update_algorithm=abi.ACT_COEF_ALGORITHM_NEWTON,
                              act_coef_update_frequency=abi.ACT_COEF_FREQUENCY_NEWTON_ITER)
    wl = dataclasses.replace(wl, opt=opt)
    # with the coefficients recomputed inside every Newton iteration the
    # recorded sub-step system is not a fixed function of the free
    # molalities, so only the concentrations are compared
    _parity(wl, abi.KERNEL_PLAIN_FP64, None, "newton-activity/plain", residual=False)
    from pfchem.lib import PfchemError
    ctx = Context(wl.net, wl.opt)
    with pytest.raises(PfchemError) as e:
        ctx.step(DeviceState(wl.state), wl.dt, abi.KERNEL_BATCHED_FP64)
    assert e.value.code == abi.PFCHEM_ERR_UNSUPPORTED


@pytest.mark.parametrize("variant,lanes", KERNELS, ids=IDS)
def test_rreact_call_bound(variant, lanes):
    """options.maximum_rreact_calls: with 5 Newton iterations allowed many
    cells cut dt, and a bound of 3 RReact calls fails the cells that need
    more.  The bound only truncates the sub-step loop, so against the same
    kernel's unbounded run every cell fails or ends equal to it, cells that
    took more than 3 sub-steps fail and one-sub-step cells do not (the mixed
    variants hand a failing cell to the FP64 pass, which starts it again
    with its own count).  Against the oracle under the same bound: the same
    failing cells (a cell whose fifth Newton iteration sits at the
    convergence test can fall either way under device rounding, so 1% may
    differ) and the oracle's concentrations where both end with the same
    sub-steps."""
    import dataclasses
    wl = workloads.random_network_workload(ncell=512)
    opt = dataclasses.replace(wl.opt, maximum_reaction_iterations=5)
    free_wl = dataclasses.replace(wl, opt=opt)
    wl = dataclasses.replace(wl, opt=dataclasses.replace(opt, maximum_rreact_calls=3))
    n = wl.state.ncell
    label = "call-bound/" + IDS[KERNELS.index((variant, lanes))]
    print("case", label, flush=True)
    free, _ = G.run_gpu(free_wl, variant, lanes=lanes)
    got, _ = G.run_gpu(wl, variant, lanes=lanes)
    f_ok = free["status"][0, :n] == 0
    g_ok = got["status"][0, :n] == 0
    f_sub = free["num_sub_steps"][0, :n]
    assert np.all(f_ok[g_ok]), label
    assert not np.any(g_ok[f_ok & (f_sub > 3)]), label
    assert np.all(g_ok[f_ok & (f_sub == 1)]), label
    assert (f_ok & (f_sub > 3)).sum() > 10 and (f_ok & (f_sub == 1)).sum() > 50, label
    np.testing.assert_array_equal(got["num_sub_steps"][0, g_ok], f_sub[g_ok])
    for f in ("pri_molal", "total", "sec_molal", "mnrl_volfrac"):
        a, b = free[f][:, :n][:, g_ok], got[f][:, :n][:, g_ok]
        assert np.max(np.abs(b - a) / np.maximum(np.abs(a), 1e-300)) <= 1e-12, (label, f)
    ref, _, _ = G.run_oracle(wl)
    r_ok = ref["status"][0, :n] == 0
    agree = float(np.mean(r_ok == g_ok))
    print(label, "status agreement with the oracle", agree, flush=True)
    assert agree >= 0.99, (label, agree)
    both = r_ok & g_ok & (ref["num_sub_steps"][0, :n] == got["num_sub_steps"][0, :n])
    assert both.sum() >= 0.95 * r_ok.sum()
    for f in ("pri_molal", "total", "sec_molal", "mnrl_volfrac"):
        a, b = ref[f][:, :n][:, both], got[f][:, :n][:, both]
        e = float(np.max(np.abs(b - a) / np.maximum(np.abs(a), 1e-300)))
        print(label, f, e, flush=True)
        assert e <= CONC_TOL, (label, f, e)
