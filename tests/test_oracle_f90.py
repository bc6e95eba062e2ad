"""The Fortran restatement of RStep (oracle/pfchem_oracle_rstep.F90), the CPU
baseline bench.py times, against the C checker (oracle/pfchem_oracle.c).

Its IEEE build (no FMA contraction) must be bit-equal to the checker in
every output field: concentrations, secondary molalities, mineral volume
fractions and rates, iteration and sub-step counts and status.  That holds
only if it performs the reference's operations (RStep reaction.F90:3564-3738,
RReact :3742-4055, RTotal, RKineticMineral, RSolve, LU) in the same order,
so it pins the restatement to the checker, which the golden vectors and
closed forms pin in turn.  The -O3 build (FMA contraction) must agree to
rounding.
"""
import dataclasses

import numpy as np
import pytest

import oracle_lib as O
from pfchem import workloads

FIELDS = ("pri_molal", "total", "sec_molal", "mnrl_volfrac", "mnrl_rate", "num_iterations",
          "num_sub_steps", "status")


def _both(wl, state=None, opt=None, ieee=True):
    opt = opt or wl.opt
    st = state or wl.state
    a, b = st.copy(), st.copy()
    fa = O.step_batch(wl.net, opt, a, wl.dt, 0)
    fb = O.step_batch_f90(wl.net, opt, b, wl.dt, 0, ieee=ieee)
    return a, b, fa, fb


def _bit_equal(a, b, n, label):
    for f in FIELDS:
        np.testing.assert_array_equal(b[f][..., :n], a[f][..., :n], err_msg="%s %s" % (label, f))


@pytest.mark.parametrize("name", ["calcite", "random15"])
def test_ieee_build_is_bit_equal_to_the_checker(name):
    wl = (workloads.calcite_workload(ncell=1000) if name == "calcite"
          else workloads.random_network_workload(ncell=4096))
    a, b, fa, fb = _both(wl)
    assert fa == fb == 0
    _bit_equal(a, b, wl.state.ncell, name)


def test_failures_cuts_and_the_call_bound_bit_equal():
    """5 Newton iterations: cells cut dt, some fail at the cut limit with
    their totals restored, others at a 3-call bound -- the same cells, with
    the same partial results"""
    wl = workloads.random_network_workload(ncell=1024)
    for bound in (0, 3):
        opt = dataclasses.replace(wl.opt, maximum_reaction_iterations=5,
                                  maximum_rreact_calls=bound)
        a, b, fa, fb = _both(wl, opt=opt)
        assert fa == fb and 0 < fa < wl.state.ncell
        _bit_equal(a, b, wl.state.ncell, "bound %d" % bound)


def test_small_totals_dry_cells_and_linear_form_bit_equal():
    """totals <= 1e-40 solved at 1e-40 and restored (calcite's negative H+
    totals), dry cells (saturation below 1e-40: no reaction, identity
    Jacobian) and the non-log formulation's damped update"""
    wl = workloads.calcite_workload(ncell=256)
    st = wl.state.copy()
    st["saturation"][0, :32] = 0.0
    assert np.any(st["total"][0] <= 1e-40)
    a, b, fa, fb = _both(wl, state=st)
    assert fa == fb
    _bit_equal(a, b, st.ncell, "dry")
    lin = dataclasses.replace(wl.opt, use_log_formulation=False)
    a, b, fa, fb = _both(wl, opt=lin)
    assert fa == fb
    _bit_equal(a, b, st.ncell, "linear")


def test_o3_build_agrees_to_rounding():
    wl = workloads.random_network_workload(ncell=4096)
    a, b, fa, fb = _both(wl, ieee=False)
    n = wl.state.ncell
    assert fa == fb == 0
    np.testing.assert_array_equal(a["status"][0, :n], b["status"][0, :n])
    assert np.mean(a["num_iterations"][0, :n] == b["num_iterations"][0, :n]) >= 0.99
    for f in ("pri_molal", "total", "sec_molal"):
        x, y = a[f][:, :n], b[f][:, :n]
        assert np.max(np.abs(y - x) / np.abs(x)) <= 1e-10, f


def test_refuses_networks_outside_its_scope():
    """the restatement covers the cells/s workload's reactions only"""
    wl = workloads.somdec_workload(ncell=8)
    with pytest.raises(ValueError):
        O.step_batch_f90(wl.net, wl.opt, wl.state.copy(), wl.dt, 1)
    wl = workloads.variable_network_workload(ncell=8)
    with pytest.raises(ValueError):
        O.step_batch_f90(wl.net, wl.opt, wl.state.copy(), wl.dt, 1)
