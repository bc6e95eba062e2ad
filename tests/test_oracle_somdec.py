"""The oracle's SomDec restatement (oracle/pfchem_oracle_somdec.c) against the
reference's clm_cn3, clm_som1 and clm_nmin regression decks, the closed-form
Q10 response, and conservation of C and N.

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


# One-reaction SOM decks with N2O emission (regression_tests/ngee/CLMCNplus).
# Both are one-cell GIRT runs without flow, so every time step solves the
# RReact system of the cell: the replay takes the decks' step sizes and
# compares the end-of-run regression values (14 digits).  The ratio of N2O to
# the mineralized N is f_T(25 C) f_w(s=1) f_pH x N2O_FRAC_MINERALIZATION
# (SomDecNemission, reaction_sandbox_somdec.F90:3477-3640), so these pin the
# emission factors; NH4+ pins the mineralization stoichiometry (SomDecReact1,
# :1914-2160) and the tracking species Nmin / HRimm / NGASmin.
from pfchem import somdec as SD
from pfchem.network import ReactionNetwork

DAY = 86400.0


def _som_cell(prim, imm, pools, reactions, total0, immobile0, **build):
    sd = SD.build_somdec(pools, reactions, prim, imm, **build)
    net = ReactionNetwork(primary_species_names=list(prim), primary_spec_Z=np.zeros(len(prim)),
                          primary_spec_a0=np.zeros(len(prim)), immobile_names=list(imm),
                          somdec=sd)
    opt = SolverOptions(act_coef_update_algorithm=abi.ACT_COEF_ALGORITHM_LAG,
                        max_residual_tolerance=1e-30, max_rel_residual_tolerance=1e-14,
                        max_relative_change_tolerance=1e-14, maximum_reaction_iterations=100)
    cc = O.CellCtx(net, opt)
    cc.c.den_kg = 1000.0          # REFERENCE_LIQUID_DENSITY 1.d3
    cc.c.porosity = 0.25
    cc.c.sat = 1.0
    cc.c.volume = 1.0
    cc.c.temp = 25.0
    cc.set("total", total0)
    cc.set("pri_molal", total0)   # no complexes, unit activity
    cc.set("pri_act_coef", np.ones(len(prim)))
    cc.set("immobile", immobile0)
    return cc


def _clm_som1(n2o_frac=0.02, x0eps=1.0e-20):
    """clm_som1.in: SOM1 -> 0.72 SOM2 (both C:N 12), 0.7 /d, pH 6"""
    prim = ["CO2(aq)", "N2O(aq)", "NH4+", "NO3-", "H+"]
    imm = ["SOM1", "SOM2"]
    cc = _som_cell(prim, imm, [SD.Pool("SOM1", 12.0), SD.Pool("SOM2", 12.0)],
                   [SD.Reaction("SOM1", [("SOM2", 0.72)], rate_constant=0.7 / DAY)],
                   np.array([1e-20, 1e-40, 1e-8, 1e-20, 1e-6]), [1.0, 1e-20],
                   n2o_frac_mineralization=n2o_frac, x0eps=x0eps)
    times = np.array(GU.golds()["clm_som1"]["obs_times_d"]) * DAY
    # the observation series records every step, times to 7 digits; the
    # deck's steps (1 h, 2 h, then the 0.1 d maximum, the last one cut to
    # land on 200 d) are multiples of 360 s
    dts = np.round(np.diff(times) / 360.0) * 360.0
    for dt in dts:
        err, nit, why = cc.rreact(float(dt))
        assert err == 0
    tot = cc.get("total", 5)
    im = cc.get("immobile", 2)
    return dts, {"Total CO2(aq)": tot[0], "Total N2O(aq)": tot[1], "Total NH4+": tot[2],
                 "SOM1": im[0], "SOM2": im[1]}


def _gold_rel(case, got):
    fin = GU.golds()[case]["final"]
    return {k: abs(v - fin["CONCENTRATION: " + k]) / abs(fin["CONCENTRATION: " + k])
            for k, v in got.items()}


def test_clm_som1_matches_the_gold():
    dts, got = _clm_som1()
    fin = GU.golds()["clm_som1"]["final"]
    assert len(dts) == fin["SOLUTION: Transport: Time Steps"]
    assert dts[0] == 3600.0 and dts[1] == 7200.0 and dts[2] == 8640.0
    rel = _gold_rel("clm_som1", got)
    print("clm_som1", rel)
    # the deck's Newton stops at PETSc's default relative tolerance (1e-8)
    for k in ("Total CO2(aq)", "Total N2O(aq)", "Total NH4+", "SOM2"):
        assert rel[k] < 5e-8, (k, rel[k])
    # SOM1 ends on the plateau that the x0eps smoothing of the upstream pool
    # leaves (1.26e-20 mol/m^3), where each step moves it by ~1e-6
    assert rel["SOM1"] < 1e-4, rel["SOM1"]


def test_clm_som1_gold_catches_wrong_emission_and_smoothing():
    """a 0.5% slip in the N2O fraction, or the smoothing floor a decade off,
    misses the gold"""
    _, got = _clm_som1(n2o_frac=0.0201)
    assert _gold_rel("clm_som1", got)["Total N2O(aq)"] > 1e-3
    _, got = _clm_som1(x0eps=1.0e-21)
    assert _gold_rel("clm_som1", got)["SOM1"] > 0.5


def _nmin_steps():
    """clm_nmin.in: 1 h initial step doubling to the 1 d maximum, the last
    step cut to land on 365 d"""
    h = 3600.0
    dts = [h * 2 ** k for k in range(5)] + [DAY] * 363
    dts.append(365.0 * DAY - sum(dts))
    return dts


def test_clm_nmin_matches_the_gold():
    """SOIL4 (C:N 10, turnover 27.4 y of 365 d) respired to CO2 with N
    mineralized to NH4+ and emitted as N2O, tracked in Nmin / HRimm /
    NGASmin, pH 6.5"""
    prim = ["CO2(aq)", "NH4+", "N2O(aq)", "H+"]
    imm = ["SOIL4", "Nmin", "HRimm", "NGASmin"]
    cc = _som_cell(prim, imm, [SD.Pool("SOIL4", 10.0)],
                   [SD.Reaction("SOIL4", [], turnover_time=27.4 * 365.0 * DAY)],
                   np.array([1e-40, 1e-40, 1e-40, 10.0 ** -6.5]), [10.0, 1e-40, 1e-40, 1e-40])
    dts = _nmin_steps()
    fin = GU.golds()["clm_nmin"]["final"]
    assert len(dts) == fin["SOLUTION: Transport: Time Steps"]
    for dt in dts:
        err, nit, why = cc.rreact(dt)
        assert err == 0
    tot = cc.get("total", 4)
    im = cc.get("immobile", 4)
    got = {"Total CO2(aq)": tot[0], "Total NH4+": tot[1], "Total N2O(aq)": tot[2],
           "SOIL4": im[0], "Nmin": im[1], "HRimm": im[2], "NGASmin": im[3]}
    rel = _gold_rel("clm_nmin", got)
    print("clm_nmin", rel)
    for k, e in rel.items():
        assert e < 1e-12, (k, e)


@pytest.mark.parametrize("temp,factor", [(25.0, 1.0), (35.0, 2.0), (15.0, 0.5), (5.0, 0.25),
                                         (0.0, 2.0 ** -2.5)])
def test_q10_response_closed_form(temp, factor):
    """first-order decay of one fixed-C:N pool under the Q10 temperature
    response (Q10 = 2, reference 25 C; elm_rspfuncs.F90): one backward-Euler
    step leaves C0 / (1 + k f_T dt), with f_T = 1 at 25 C, Q10 per 10 C and
    continuous at 0 C where the branch for frozen soil starts"""
    k, dt, c0 = 1.0e-6, 1800.0, 10.0
    prim = ["CO2(aq)", "NH4+", "H+"]
    cc = _som_cell(prim, ["SOIL4"], [SD.Pool("SOIL4", 10.0)],
                   [SD.Reaction("SOIL4", [], rate_constant=k,
                                temperature_response_function=SD.TEMPERATURE_Q10, Q10=2.0)],
                   np.array([1e-6, 1e-6, 1e-7]), [c0])
    cc.c.temp = temp
    err, nit, why = cc.rreact(dt)
    assert err == 0
    got = cc.get("immobile", 1)[0]
    want = c0 / (1.0 + k * factor * dt)
    assert abs(got - want) / want < 1e-13, (temp, got, want)


def _ox_cell(ox_code_species, half_sat, ox_value, track=None, k=1.0e-5, c0=10.0):
    """one fixed-C:N pool (C:N 10) decaying to CO2 + NH4+ under the Monod
    oxygen response, oxidizer either the aqueous O2(aq) or an immobile OXI
    pool that nothing consumes"""
    prim = ["CO2(aq)", "NH4+", "H+", "O2(aq)"]
    imm = ["SOIL4", "OXI"]
    rx = SD.Reaction("SOIL4", [], rate_constant=k, Ox_response_function=SD.OX_MONOD,
                     Ox_half_saturation=half_sat, Ox_species=ox_code_species)
    o2 = ox_value if ox_code_species == "O2(aq)" else 2.5e-4
    oxi = ox_value if ox_code_species == "OXI" else 1.0
    cc = _som_cell(prim, imm, [SD.Pool("SOIL4", 10.0)], [rx],
                   np.array([1e-6, 1e-6, 1e-7, o2]), [c0, oxi], o2_species=track)
    return cc


@pytest.mark.parametrize("oxi,half_sat", [(1.0, 1.0), (0.1, 1.0), (3.0, 0.5), (1e-3, 2.0)])
def test_oxygen_monod_closed_form(oxi, half_sat):
    """GetAerobicCondition's Monod response (elm_rspfuncs.F90:282-316) with an
    oxidizer that the reaction does not consume scales the first-order rate
    by f = Ox / (K + Ox): one backward-Euler step leaves C0 / (1 + k f dt)"""
    k, dt, c0 = 1.0e-5, 3600.0, 10.0
    cc = _ox_cell("OXI", half_sat, oxi, k=k, c0=c0)
    err, nit, why = cc.rreact(dt)
    assert err == 0
    got = cc.get("immobile", 2)
    f = oxi / (half_sat + oxi)
    want = c0 / (1.0 + k * f * dt)
    assert abs(got[0] - want) / want < 1e-13, (got[0], want)
    assert got[1] == oxi  # the oxidizer is untouched


def test_oxygen_monod_closed_form_catches_a_wrong_half_saturation():
    """the same step with the half saturation 10% off misses the closed form
    by far more than the bar above"""
    k, dt, c0, oxi = 1.0e-5, 3600.0, 10.0, 1.0
    cc = _ox_cell("OXI", 1.1, oxi, k=k, c0=c0)
    assert cc.rreact(dt)[0] == 0
    want = c0 / (1.0 + k * oxi / (1.0 + oxi) * dt)
    assert abs(cc.get("immobile", 2)[0] - want) / want > 1e-4


def test_oxygen_tracking_balances_co2():
    """with O2_SPECIES_NAME the reaction consumes O2 as it releases CO2
    (reaction_sandbox_somdec.F90:2176-2186): total O2 + CO2 is unchanged and
    the CO2 released is the pool's C loss over the pore water volume, with
    O2 both tracked and the Monod oxidizer.  O2 stays at levels where the
    reference's iteration converges: its Jacobian adds the response's
    derivative in O2 (per mol/L) to the pool's column, and below ~1e-4 mol/L
    with K = 1e-5 that term stalls the Newton iteration of one RReact"""
    k, dt, c0 = 1.0e-5, 3600.0, 0.01
    psv = 0.25 * 1.0 * 1000.0 * 1.0  # porosity sat 1000 volume (L)
    for o2 in (2.5e-4, 1.0e-3):
        cc = _ox_cell("O2(aq)", 1.0e-5, o2, track="O2(aq)", k=k, c0=c0)
        t0 = cc.get("total", 4).copy()
        assert cc.rreact(dt)[0] == 0
        t1 = cc.get("total", 4)
        im = cc.get("immobile", 2)
        np.testing.assert_allclose(t1[0] + t1[3], t0[0] + t0[3], rtol=1e-13)
        np.testing.assert_allclose((t1[0] - t0[0]) * psv, (c0 - im[0]) * 1.0, rtol=1e-11)
        assert 0.0 < t1[3] < t0[3]


def test_somdec_with_complexes_conserves_and_reduces_without_them():
    """the sandbox on a network with aqueous complexes (NH4+ / NO3- / CO2
    ion pairs) reads component totals and dtotal with the complexes' share:
    C and N stay conserved over a step, and with the complexes' log K raised
    until they vanish the step reproduces the network without them"""
    import dataclasses
    import gpu_util as G
    n = 256
    wl = workloads.somdec_workload(n, complexes=True)
    ref, rec, nfail = G.run_oracle(wl)
    assert nfail == 0
    share = (ref["sec_molal"][0, :n] + ref["sec_molal"][1, :n]) / ref["total"][1, :n]
    assert np.median(share) > 0.05
    C0, N0 = SU.carbon_nitrogen(wl.net, wl.state.arrays, n)
    C1, N1 = SU.carbon_nitrogen(wl.net, ref, n)
    np.testing.assert_allclose(C1, C0, rtol=1e-12)
    np.testing.assert_allclose(N1, N0, rtol=1e-11)
    gone = dataclasses.replace(wl.net, eqcplx_logK=np.array([40.0, 40.0]))
    a, _, nf_a = G.run_oracle(dataclasses.replace(wl, net=gone))
    plain = workloads.somdec_workload(n)
    b, _, nf_b = G.run_oracle(plain)
    assert nf_a == 0 and nf_b == 0
    for f in ("total", "immobile"):
        np.testing.assert_allclose(a[f][:plain.net.naqcomp if f == "total" else None, :n],
                                   b[f][:, :n], rtol=1e-12, atol=1e-300)
