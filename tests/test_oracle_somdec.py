"""The oracle's SomDec restatement (oracle/pfchem_oracle_somdec.c) against the
reference's clm_cn3 regression deck and against conservation of C and N.

clm_cn3 (regression_tests/ngee/CLMCNplus) is a one-cell GIRT run of the
CLM-CN litter/SOM cascade with Monod NH4+/NO3- limitation; its observation
file records every step (7 significant digits).  The first steps, while the
litter immobilizes N from both NH4+ and NO3-, are replayed with the same
step sizes through RReact and matched to the recorded precision.  (Later
steps are not a parity target: the deck's Newton runs with the sandbox's
approximate Jacobian to the GIRT tolerances, and the replay drifts from it
once N limits the litter reactions.)
"""
import numpy as np
import pytest

import golden_util as GU
import oracle_lib as O
import somdec_util as SU
from pfchem import abi, workloads
from pfchem.network import SolverOptions
from pfchem.somdec import CN_RATIO_MASS_TO_MOL

NICE_DT = np.array([225.0 * 2 ** k for k in range(13)] + [432000.0, 864000.0])


def _clm_cn3_cell(net=None):
    net = net or workloads.somdec_network()
    opt = SolverOptions(act_coef_update_algorithm=abi.ACT_COEF_ALGORITHM_LAG,
                        max_residual_tolerance=1e-30, max_rel_residual_tolerance=1e-14,
                        max_relative_change_tolerance=1e-14, maximum_reaction_iterations=100)
    cc = O.CellCtx(net, opt)
    cc.c.den_kg = 1000.0          # REFERENCE_LIQUID_DENSITY 1.d3
    cc.c.porosity = 0.25
    cc.c.sat = 1.0
    cc.c.volume = 1.0
    cc.c.temp = 25.0
    T0 = np.array([1e-40, 2e-6, 2e-6, 1e-20, 1e-6])   # CO2, NH4+, NO3-, N2O, H+ (pH 6)
    cc.set("total", T0)
    cc.set("pri_molal", T0)
    cc.set("pri_act_coef", np.ones(5))
    cc.set("immobile", [1e-40, 1e-40, 1e-40, 1e-40, 0.1852, 0.4578, 0.2662,
                        0.00508954, 0.01258096, 0.00731553])
    return net, cc


def test_clm_cn3_first_steps_match_the_gold():
    g = GU.golds()["clm_cn3"]
    rows = np.array(g["rows"])
    t = rows[:, 0] * 86400.0
    # totals of CO2, NH4+, NO3-, N2O (M) then the ten immobile species
    cols = [2, 3, 4, 5] + list(range(7, 17))
    net, cc = _clm_cn3_cell()
    for k in range(1, 21):
        dt = t[k] - t[k - 1]
        dt = NICE_DT[np.argmin(np.abs(NICE_DT - dt))]    # the times carry 7 digits
        err, nit, why = cc.rreact(dt)
        assert err == 0
        got = np.concatenate([cc.get("total", 5)[:4], cc.get("immobile", 10)])
        ref = rows[k, cols]
        big = np.abs(ref) > 1e-15
        rel = np.abs(got - ref)[big] / np.abs(ref)[big]
        assert rel.max() < 1e-5, (k, rel.max())


def test_clm_cn3_gold_catches_a_wrong_stoichiometry():
    """the same replay with the SOM1 C:N ratio taken as molar instead of mass
    (a plausible unit slip) misses the first steps"""
    g = GU.golds()["clm_cn3"]
    rows = np.array(g["rows"])
    net = workloads.somdec_network(som_cn=(12.0 * CN_RATIO_MASS_TO_MOL, 12.0, 10.0, 10.0))
    _, cc = _clm_cn3_cell(net)
    err, nit, why = cc.rreact(3600.0)
    got = cc.get("total", 5)[1]
    assert abs(got - rows[1, 3]) / rows[1, 3] > 1e-4


def test_somdec_step_conserves_carbon_and_nitrogen():
    wl = workloads.somdec_workload(4000)
    st = wl.state.copy()
    assert O.step_batch(wl.net, wl.opt, st, wl.dt, 0) == 0
    C0, N0 = SU.carbon_nitrogen(wl.net, wl.state.arrays, st.ncell)
    C1, N1 = SU.carbon_nitrogen(wl.net, st.arrays, st.ncell)
    np.testing.assert_allclose(C1, C0, rtol=1e-11)
    np.testing.assert_allclose(N1, N0, rtol=1e-9)
    # and something happened
    assert np.abs(st.arrays["immobile"][:, :st.ncell] -
                  wl.state.arrays["immobile"][:, :st.ncell]).max() > 1e-6


def test_somdec_moisture_zero_stops_decomposition():
    wl = workloads.somdec_workload(64)
    st = wl.state.copy()
    st.arrays["somdec_f_w"][0, :] = 0.0
    assert O.step_batch(wl.net, wl.opt, st, wl.dt, 0) == 0
    np.testing.assert_allclose(st.arrays["immobile"][:, :64], wl.state.arrays["immobile"][:, :64],
                               rtol=1e-12)
