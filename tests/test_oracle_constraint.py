"""Constraint speciation in the oracle (orc_equilibrate, the restatement of
ReactionEquilibrateConstraint, reaction.F90:1328-2117), pinned by

* the reference's golds whose state is the constraint's output: the
  ca-carbonate batch decks (three primaries, total constraints) with unit
  activity (deck criterion 1e-12 absolute) and Debye-Hueckel activity, and
  the two-primary carbonate Debye-Hueckel deck;
* what each constraint type fixes, checked with an independent numpy
  evaluation of mass action and the Debye-Hueckel law at the answer: pH
  (-log10 of the H+ activity), free molality, total, mineral saturation
  (ln Q/K = 0), gas partial pressure, charge balance.

The Debye-Hueckel ca-carbonate gold is matched to 2e-9 in pH rather than the
deck's 1e-12: its three activity coefficients all correspond to an ionic
strength 5.6e-10 (1.9e-7 relative) below the one of the converged free-ion
molalities, which agree with the gold's (the pH gap is the gamma gap).
"""
import numpy as np
import pytest

import golden_util as GU
import oracle_lib as O
from pfchem import abi, constraint, database
from pfchem.network import CellState

DEN = 997.1604062673
LN10 = 2.30258509299
CA_PRI = ["HCO3-", "Ca++", "H+"]
CA_SEC = ["OH-", "CO2(aq)", "CO3--", "CaOH+", "CaHCO3+", "CaCO3(aq)"]
CAL_PRI = ["H+", "HCO3-", "Ca++"]
CAL_SEC = ["OH-", "CO3--", "CO2(aq)", "CaOH+", "CaHCO3+", "CaCO3(aq)"]


def _state(net, ncell=1):
    st = CellState(net, ncell)
    st["den_kg"][0, :] = DEN
    st["porosity"][0, :] = 0.25
    st["saturation"][0, :] = 1.0
    st["volume"][0, :] = 1.0
    if net.nkinmnrl:
        st["mnrl_area"][:, :] = 1.0
    return st


def _run(net, opt, con, values):
    """equilibrate one cell per column of values (naqcomp, ncell)"""
    values = np.asarray(values, dtype=np.float64)
    st = _state(net, values.shape[1])
    nfail = O.equilibrate_batch(net, opt, con, st, values, 1)
    return st, nfail


def _ca_carbonate():
    db = GU.db("ca-carbonate.dat")
    return db, database.build_network(db, CA_PRI, CA_SEC, temperature=25.0, isothermal=True)


def _gold(case):
    g = GU.golds()[case]["values"]
    return (np.array([g["CONCENTRATION: Total %s" % n] for n in CA_PRI]),
            np.array([g["GENERIC: Gamma %s" % n] for n in CA_PRI]), g["GENERIC: pH"])


def test_ca_carbonate_unit_activity_gold():
    db, net = _ca_carbonate()
    tot, gam, ph = _gold("ca-carbonate-unit-activity")
    con = constraint.from_deck(db, net, {}, 25.0, True)
    st, nfail = _run(net, GU.lag_options(), con, [[3e-3], [1e-3], [1e-3]])
    assert nfail == 0
    m, g = st["pri_molal"][:, 0], st["pri_act_coef"][:, 0]
    assert abs(-np.log10(m[2] * g[2]) - ph) < 1e-12
    assert np.all(np.abs(st["total"][:, 0] - tot) < 1e-12)
    assert np.all(g == gam)


def test_ca_carbonate_debye_huckel_gold():
    db, net = _ca_carbonate()
    tot, gam, ph = _gold("ca-carbonate-debye-huckel-activity")
    con = constraint.from_deck(db, net, {}, 25.0, True)
    opt = GU.lag_options(act_coef_update_frequency=abi.ACT_COEF_FREQUENCY_TIMESTEP)
    st, nfail = _run(net, opt, con, [[3e-3], [1e-3], [1e-3]])
    assert nfail == 0
    m, g = st["pri_molal"][:, 0], st["pri_act_coef"][:, 0]
    assert abs(-np.log10(m[2] * g[2]) - ph) < 5e-9
    assert np.all(np.abs(g / gam - 1.0) < 5e-8)
    assert np.all(np.abs(st["total"][:, 0] - tot) < 2e-11)
    # a planted mistake the gold catches: the ion-size parameters of Ca++
    # and HCO3- swapped
    net.primary_spec_a0 = np.array(net.primary_spec_a0)[[1, 0, 2]].copy()
    st, _ = _run(net, opt, con, [[3e-3], [1e-3], [1e-3]])
    assert abs(st["pri_act_coef"][1, 0] / gam[1] - 1.0) > 1e-3


def test_carbonate_debye_huckel_gold():
    g = GU.golds()["carbonate-debye-huckel-activity"]["values"]
    db = GU.db("carbonate.dat")
    net = GU.carbonate_network()
    con = constraint.from_deck(db, net, {}, 25.0, True)
    opt = GU.lag_options(act_coef_update_frequency=abi.ACT_COEF_FREQUENCY_TIMESTEP)
    st, nfail = _run(net, opt, con, [[g["CONCENTRATION: Total H+"]],
                                     [g["CONCENTRATION: Total HCO3-"]]])
    assert nfail == 0
    m, gam = st["pri_molal"][:, 0], st["pri_act_coef"][:, 0]
    assert abs(-np.log10(m[0] * gam[0]) - g["GENERIC: pH"]) < 1e-12
    assert abs(gam[0] - g["GENERIC: Gamma H+"]) < 1e-12
    assert abs(gam[1] - g["GENERIC: Gamma HCO3-"]) < 1e-12


# ---- what each constraint fixes, evaluated independently ----

def _db_with_co2_gas():
    """calcite.dat plus a CO2(g) = -H2O + H+ + HCO3- passive gas (log K as a
    function of temperature of the same form as the database's reactions)"""
    db = GU.db("calcite.dat")
    rx = database.DbReaction(name="CO2(g)", species=["H2O", "H+", "HCO3-"],
                             stoich=[-1.0, 1.0, 1.0],
                             logK=np.array([-7.6765, -7.8136, -8.0527, -8.3574, -8.7692,
                                            -9.2392, -9.7687, -10.3912]))
    db.gases["CO2(g)"] = rx
    return db


def _calcite_net(db):
    return database.build_network(db, CAL_PRI, CAL_SEC, [database.KineticMineralSpec(
        "Calcite", rate_constant=1e-9)], temperature=25.0, isothermal=True)


def _numpy_speciation(net, m, den, unit=False):
    """gamma (LAG Debye-Hueckel at these molalities and the secondaries
    they imply, iterated to a fixed point; 1 with unit), secondaries and
    totals"""
    Zp, a0p = np.asarray(net.primary_spec_Z), np.asarray(net.primary_spec_a0)
    Zs, a0s = np.asarray(net.eqcplx_Z), np.asarray(net.eqcplx_a0)
    A, B, Bd = net.debyeA, net.debyeB, net.debyeBdot
    gp, gs = np.ones_like(m), np.ones(len(Zs))
    for _ in range(200):
        sec = np.array([np.exp(-lk * LN10 + sum(nu * np.log(m[j] * gp[j])
                                                 for j, nu in zip(sp, st))) / gs[k]
                        for k, (sp, st, lk) in enumerate(zip(net.eqcplx_species,
                                                             net.eqcplx_stoich,
                                                             net.eqcplx_logK))])
        I = 0.5 * (np.sum(m * Zp ** 2) + np.sum(sec * Zs ** 2))
        sI = np.sqrt(I)

        def dh(z, a0):
            return np.where(np.abs(z) > 1e-10,
                            np.exp((-z * z * sI * A / (1 + a0 * B * sI) + Bd * I) * LN10), 1.0)
        gp2, gs2 = (gp, gs) if unit else (dh(Zp, a0p), dh(Zs, a0s))
        done = np.max(np.abs(gp2 - gp)) < 1e-15 and np.max(np.abs(gs2 - gs)) < 1e-15
        gp, gs = gp2, gs2
        if done:
            break
    tot = m.copy()
    for k, (sp, st) in enumerate(zip(net.eqcplx_species, net.eqcplx_stoich)):
        for j, nu in zip(sp, st):
            tot[j] += nu * sec[k]
    return gp, sec, tot * den / 1000.0


def _lnqk(sp, st, logK, m, gp):
    return -logK * LN10 + sum(nu * np.log(m[j] * gp[j]) for j, nu in zip(sp, st))


@pytest.mark.parametrize("activity", [False, True])
def test_ph_total_and_charge_balance_constraint(activity):
    """H+ pH 5, HCO3- total 1e-3 mol/L, Ca++ by charge balance (the 'west'
    constraint of the 1d-calcite deck)"""
    db = _db_with_co2_gas()
    net = _calcite_net(db)
    con = constraint.from_deck(db, net, {"H+": ("P",), "HCO3-": ("T",), "Ca++": ("Z",)},
                               25.0, True)
    opt = GU.lag_options(act_coef_update_frequency=(abi.ACT_COEF_FREQUENCY_TIMESTEP if activity
                                                    else abi.ACT_COEF_FREQUENCY_OFF))
    st, nfail = _run(net, opt, con, [[5.0, 6.5], [1e-3, 2e-3], [1e-6, 1e-6]])
    assert nfail == 0
    for c, ph in enumerate((5.0, 6.5)):
        m = st["pri_molal"][:, c]
        gp, sec, tot = _numpy_speciation(net, m, DEN, unit=not activity)
        assert np.allclose(st["pri_act_coef"][:, c], gp, rtol=1e-9, atol=0)
        assert abs(-np.log10(m[0] * gp[0]) - ph) < 1e-10
        assert abs(tot[1] - [1e-3, 2e-3][c]) < 1e-12
        charge = np.dot(net.primary_spec_Z, tot)
        assert abs(charge) < 1e-12
        assert m[2] > 0


def test_free_mineral_and_gas_constraint():
    """H+ free 1e-8, HCO3- from CO2(g) at 10^-3.5 bar, Ca++ in equilibrium with
    calcite (the 'initial' constraint of the 1d-calcite deck)"""
    db = _db_with_co2_gas()
    net = _calcite_net(db)
    con = constraint.from_deck(db, net, {"H+": ("F",), "HCO3-": ("G", "CO2(g)"),
                                         "Ca++": ("M", "Calcite")}, 25.0, True)
    assert con.rxn_names == ["CO2(g)", "Calcite"]
    opt = GU.lag_options(act_coef_update_frequency=abi.ACT_COEF_FREQUENCY_TIMESTEP)
    # the gas value as log10 partial pressure (<= 0) and as a pressure
    st, nfail = _run(net, opt, con, [[1e-8, 1e-8], [-3.5, 10 ** -3.5], [5e-4, 5e-4]])
    assert nfail == 0
    for c in range(2):
        m = st["pri_molal"][:, c]
        gp, _, _ = _numpy_speciation(net, m, DEN)
        assert abs(m[0] / (1e-8 * 1000.0 / DEN) - 1.0) < 4e-16
        gas = db.gases["CO2(g)"]
        assert abs(_lnqk([0, 1], [1.0, 1.0], -7.8136, m, gp) - np.log(10 ** -3.5)) < 1e-10
        assert abs(_lnqk(net.kinmnrl_species[0], net.kinmnrl_stoich[0], net.kinmnrl_logK[0],
                         m, gp)) < 1e-10
        assert gas.name == "CO2(g)"
    assert np.allclose(st["pri_molal"][:, 0], st["pri_molal"][:, 1], rtol=1e-12, atol=0)


def test_log_constraint_and_previous_solution_guess():
    db = _db_with_co2_gas()
    net = _calcite_net(db)
    opt = GU.lag_options(act_coef_update_frequency=abi.ACT_COEF_FREQUENCY_TIMESTEP)
    con = constraint.from_deck(db, net, {"H+": ("L",)}, 25.0, True)
    st, nfail = _run(net, opt, con, [[-7.0], [2e-3], [1e-3]])
    assert nfail == 0 and abs(st["pri_molal"][0, 0] / (10 ** -7.0 * 1000.0 / DEN) - 1) < 4e-16
    # restarting from the answer converges at once to the same state
    con.use_prev_soln_as_guess = True
    st2 = st.copy()
    assert O.equilibrate_batch(net, opt, con, st2, np.array([[-7.0], [2e-3], [1e-3]]), 1) == 0
    assert st2["num_iterations"][0, 0] < st["num_iterations"][0, 0]
    assert np.allclose(st2["pri_molal"], st["pri_molal"], rtol=1e-10, atol=0)


def test_failures_are_reported():
    db = _db_with_co2_gas()
    net = _calcite_net(db)
    opt = GU.lag_options()
    con = constraint.from_deck(db, net, {}, 25.0, True, max_iterations=2)
    st, nfail = _run(net, opt, con, [[1e-3], [2e-3], [1e-3]])
    assert nfail == 1 and st["status"][0, 0] == abi.PFCHEM_CELL_CONSTRAINT_MAX_ITERATIONS
    with pytest.raises(NotImplementedError):
        constraint.from_deck(db, net, {"H+": ("E",)}, 25.0, True).validate(net)
    with pytest.raises(KeyError):
        constraint.from_deck(db, net, {"Na+": ("T",)}, 25.0, True)
