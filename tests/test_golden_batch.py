"""The oracle against the reference's own regression golds
(regression_tests/ascem/batch, values collected into
tests/golden/reference_golds.json by tests/golden/make_golden.py).

* carbonate-unit-activity fixes the water density of these transport-only
  runs (its pH depends on nothing else once log K is the table's);
* carbonate-debye-huckel-activity then pins Debye-Hueckel (RActivity-
  Coefficients, reaction.F90:4368-4614), the totals (RTotalAqueous,
  :4665-4759) and the database reader: pH and both activity coefficients;
* calcite-kinetics pins the kinetic path end to end (RReact, RKineticMineral,
  RUpdateKineticState): 500 steps of 0.5 s of calcite dissolution.

The gold runs are GIRT (global implicit); one cell without transport solves
the same backward-Euler equations as RReact, to the GIRT Newton's own
convergence (its component totals differ from exact conservation by
~2e-13 mol/L, which bounds how close any exact solver can come).
"""
import numpy as np
import pytest

import golden_util as GU
from pfchem import abi, database

pytestmark = pytest.mark.filterwarnings("ignore::RuntimeWarning")


@pytest.fixture(scope="module")
def den():
    d = GU.density_from_unit_activity()
    # liquid water at 25 C and 1 atm
    assert 996.9 < d < 997.3
    return d


def test_logk_table_and_fit():
    db = GU.db("calcite.dat")
    assert list(db.temperatures) == [0.0, 25.0, 60.0, 100.0, 150.0, 200.0, 250.0, 300.0]
    assert set(db.secondary) >= {"OH-", "CO3--", "CO2(aq)", "CaOH+", "CaHCO3+", "CaCO3(aq)"}
    cal = db.minerals["Calcite"]
    assert cal.molar_volume == pytest.approx(36.934)
    # a five-term fit through exactly generated data recovers it
    coef = np.array([-3.0, 25.0, -0.01, 1500.0, -2.0e5])
    logk = [database.interpolate_logk(coef, t) for t in db.temperatures]
    fit = database.fit_logk_coef(db.temperatures, logk)
    for t in (0.0, 25.0, 100.0, 300.0):
        assert database.interpolate_logk(fit, t) == pytest.approx(
            database.interpolate_logk(coef, t), abs=1e-6)
    # the fit over the real table stays close to it at 25 C
    for name in ("OH-", "CO3--", "CO2(aq)", "CaHCO3+", "CaCO3(aq)"):
        rx = db.secondary[name]
        lk, _ = database.logk_at(db.temperatures, rx.logK, 25.0, isothermal=False)
        assert abs(lk - rx.logK[1]) < 0.02
        lk_iso, _ = database.logk_at(db.temperatures, rx.logK, 25.0, isothermal=True)
        assert lk_iso == rx.logK[1]


def test_network_from_database():
    net = GU.carbonate_network()
    assert net.primary_species_names == ["H+", "HCO3-"]
    assert list(net.primary_spec_Z) == [1.0, -1.0]
    assert list(net.primary_spec_a0) == [9.0, 4.0]
    # OH- = -H+ + H2O ; CO3-- = -H+ + HCO3- ; CO2(aq) = -H2O + H+ + HCO3-
    assert net.eqcplx_species == [[0], [0, 1], [0, 1]]
    assert net.eqcplx_stoich == [[-1.0], [-1.0, 1.0], [1.0, 1.0]]
    assert list(net.eqcplxh2ostoich) == [1.0, 0.0, -1.0]
    assert list(net.eqcplx_logK) == [13.9951, 10.3288, -6.3447]
    assert list(net.eqcplx_Z) == [-1.0, -2.0, 0.0]


def test_carbonate_debye_huckel_gold(den):
    g = GU.golds()["carbonate-debye-huckel-activity"]["values"]
    net = GU.carbonate_network()
    cc = GU.equilibrate(net, GU.lag_options(),
                        [g["CONCENTRATION: Total H+"], g["CONCENTRATION: Total HCO3-"]], den, True)
    gam = cc.get("pri_act_coef", 2)
    # regression criterion of the deck: 1e-12 absolute (batch.cfg)
    assert abs(GU.ph(cc) - g["GENERIC: pH"]) < 1e-12
    assert abs(gam[0] - g["GENERIC: Gamma H+"]) < 1e-12
    assert abs(gam[1] - g["GENERIC: Gamma HCO3-"]) < 1e-12


def test_carbonate_gold_fails_with_a_planted_mistake(den):
    """the Debye-Hueckel gold distinguishes the ion-size parameters: the same
    equilibrium with the two a0 swapped misses it"""
    g = GU.golds()["carbonate-debye-huckel-activity"]["values"]
    net = GU.carbonate_network()
    net.primary_spec_a0 = net.primary_spec_a0[::-1].copy()
    cc = GU.equilibrate(net, GU.lag_options(),
                        [g["CONCENTRATION: Total H+"], g["CONCENTRATION: Total HCO3-"]], den, True)
    assert abs(cc.get("pri_act_coef", 1)[0] - g["GENERIC: Gamma H+"]) > 1e-6


def calcite_kinetics_network():
    db = GU.db("calcite.dat")
    return database.build_network(
        db, ["H+", "HCO3-", "Ca++"],
        ["OH-", "CO3--", "CO2(aq)", "CaOH+", "CaHCO3+", "CaCO3(aq)"],
        # RATE_CONSTANT 1.d-13 mol/cm^2-sec
        [database.KineticMineralSpec("Calcite", rate_constant=1e-13 * 1e4)],
        temperature=25.0, isothermal=True)


def test_calcite_kinetics_gold(den):
    g = GU.golds()["calcite-kinetics"]["values"]
    net = calcite_kinetics_network()
    opt = GU.lag_options(act_coef_update_frequency=abi.ACT_COEF_FREQUENCY_TIMESTEP,
                         max_residual_tolerance=1e-30, max_rel_residual_tolerance=1e-12,
                         max_relative_change_tolerance=1e-13, maximum_reaction_iterations=50)
    cc = GU.equilibrate(net, opt, [-1e-5, 1e-5, 1e-5], den, True)
    cc.c.porosity = 0.5
    cc.c.sat = 1.0
    cc.c.volume = 1.0
    cc.set("mnrl_volfrac", [0.2])
    cc.set("mnrl_area", [100.0])            # 1 cm^2/cm^3
    t0 = cc.get("total", 3)
    for _ in range(500):                    # 250 s in steps of 0.5 s
        cc.activity()                       # ACTIVITY_COEFFICIENTS TIMESTEP
        err, nit, why = cc.rreact(0.5)
        assert err == 0
        cc.update_kinetic_state(0.5)
    tot = cc.get("total", 3)
    # exact conservation: each component moved by the calcite dissolved
    d = tot - t0
    assert d[1] == pytest.approx(d[2], rel=1e-9) and d[0] == pytest.approx(-d[2], rel=1e-9)
    # the deck's criteria: 1e-12 absolute for volume fractions, rates and
    # concentrations
    assert abs(cc.get("mnrl_volfrac", 1)[0] - g["VOLUME_FRACTION: Calcite VF"]) < 1e-12
    assert abs(cc.get("mnrl_rate", 1)[0] - g["RATE: Calcite Rate"]) < 1e-12
    assert cc.get("mnrl_rate", 1)[0] == pytest.approx(g["RATE: Calcite Rate"], rel=1e-9)
    for i, name in enumerate(["H+", "HCO3-", "Ca++"]):
        assert abs(tot[i] - g["CONCENTRATION: Total %s" % name]) < 1e-12
    # pH inherits the gold run's own H+ total truncation (2.4e-13 mol/L)
    assert abs(GU.ph(cc) - g["GENERIC: pH"]) < 2e-8
