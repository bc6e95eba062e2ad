"""Replays of the reference's batch regression decks (ascem/batch) with the
oracle, for the golden-vector tests.

The decks are single-cell GIRT runs with no transport, so each time step
solves the same backward-Euler cell equations as RReact.  Their initial
state comes from a constraint speciation (ReactionEquilibrateConstraint,
reaction.F90:1328-2117) with total ("T") constraints; `equilibrate` solves
the same fixed point: free molalities whose totals (mol/L) match the
constraint, with activity coefficients consistent with those molalities.
"""
import json
import os

import numpy as np
from scipy.optimize import brentq, fsolve

import oracle_lib as O
from pfchem import abi, database
from pfchem.network import SolverOptions

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def golds():
    with open(os.path.join(GOLDEN, "reference_golds.json")) as f:
        return json.load(f)


def db(name):
    return database.read_database(os.path.join(GOLDEN, name))


def carbonate_network():
    # transport-only runs take log K at 25 C from the table
    return database.build_network(db("carbonate.dat"), ["H+", "HCO3-"],
                                  ["OH-", "CO3--", "CO2(aq)"], temperature=25.0,
                                  isothermal=True)


def lag_options(**kw):
    return SolverOptions(act_coef_update_algorithm=abi.ACT_COEF_ALGORITHM_LAG, **kw)


def equilibrate(net, opt, totals, den, activity, cc=None):
    """oracle cell whose totals equal `totals` (mol/L) at density `den`"""
    cc = cc or O.CellCtx(net, opt)
    cc.c.den_kg = den
    n = net.naqcomp
    T = np.asarray(totals, dtype=np.float64)
    lnm = np.log(np.maximum(np.abs(T), 1e-14) / (den * 1e-3))

    def f(x):
        cc.set("pri_molal", np.exp(x))
        cc.total()
        return cc.get("total", n) - T

    def jac(x):
        cc.set("pri_molal", np.exp(x))
        cc.total()
        return cc.dtotal() * np.exp(x)[None, :]

    for _ in range(200):
        lnm = fsolve(f, lnm, fprime=jac, xtol=1e-14)
        f(lnm)
        if not activity:
            break
        g0 = cc.get("pri_act_coef", n)
        s0 = cc.get("sec_act_coef", net.neqcplx)
        cc.activity()
        if (np.max(np.abs(cc.get("pri_act_coef", n) - g0)) < 1e-16 and
                np.max(np.abs(cc.get("sec_act_coef", net.neqcplx) - s0), initial=0) < 1e-16):
            break
    f(lnm)
    return cc


def ph(cc):
    return -np.log10(cc.get("pri_molal", 1)[0] * cc.get("pri_act_coef", 1)[0])


def density_from_unit_activity():
    """Water density of the transport-only decks, from the unit-activity
    carbonate gold (pH there depends on the density alone: the totals are in
    mol/L, the speciation in mol/kg)."""
    g = golds()["carbonate-unit-activity"]["values"]
    net = carbonate_network()
    opt = lag_options()
    T = [g["CONCENTRATION: Total H+"], g["CONCENTRATION: Total HCO3-"]]
    return brentq(lambda d: ph(equilibrate(net, opt, T, d, False)) - g["GENERIC: pH"],
                  990.0, 1000.0, xtol=1e-13, rtol=1e-15)
