"""Workloads of the paths the pin tests cover: immobile species with
first-order decay, temperature-dependent log K, kinetic-mineral rate options
(test infrastructure)."""
import dataclasses
import os

import numpy as np

from pfchem import database, workloads
from pfchem.network import CellState

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def decay_workload(ncell=200, k=(2.0e-5, 3.0e-5)):
    """the calcite workload plus two immobile species B1, B2; B1 decays
    through two first-order reactions (rate constants k)"""
    base = workloads.calcite_workload(ncell)
    net = dataclasses.replace(base.net, immobile_names=["B1", "B2"],
                              decayspecid=np.array([0, 0], dtype=np.int32),
                              decay_rate_constant=np.array(k))
    st = CellState(net, ncell)
    for name, a in base.state.arrays.items():
        if a.shape[0]:
            st.arrays[name][...] = a
    rng = np.random.default_rng(7)
    st["immobile"][0, :] = 10.0 ** rng.uniform(-4, 0, ncell)
    st["immobile"][1, :] = 10.0 ** rng.uniform(-4, 0, ncell)
    return dataclasses.replace(base, net=net, state=st), base


def calcite_db():
    return database.read_database(os.path.join(GOLDEN, "calcite.dat"))


CALCITE_COMPLEXES = ["OH-", "CO3--", "CO2(aq)", "CaHCO3+", "CaCO3(aq)"]


def logk_coefficients():
    """ReactionInterpolateLogK coefficients of the calcite workload's
    complexes and of calcite, fitted to calcite.dat's tables"""
    db = calcite_db()
    temps = np.asarray(db.temperatures, dtype=np.float64)
    ccoef = np.array([database.fit_logk_coef(temps, database._species(db, n).logK)
                      for n in CALCITE_COMPLEXES])
    mcoef = np.array([database.fit_logk_coef(temps, db.minerals["Calcite"].logK)])
    return ccoef, mcoef


def nonisothermal_workload(ncell=512, seed=11):
    """the calcite workload with calcite.dat's temperature-dependent log K
    (fitted coefficients), cells at T ~ U[5, 95] C, non-isothermal options"""
    wl = workloads.calcite_workload(ncell)
    ccoef, mcoef = logk_coefficients()
    net = dataclasses.replace(wl.net, eqcplx_logKcoef=ccoef, kinmnrl_logKcoef=mcoef,
                              kinmnrl_activation_energy=np.array([4.0e4]))
    opt = dataclasses.replace(wl.opt, isothermal=0)
    st = wl.state.copy()
    st["temperature"][0, :] = np.random.default_rng(seed).uniform(5.0, 95.0, ncell)
    return dataclasses.replace(wl, net=net, opt=opt, state=st)


def mineral_options_workload(ncell=512, **opts):
    """the calcite workload with calcite's rate options set (one value each)"""
    wl = workloads.calcite_workload(ncell)
    arrays = {k: (np.array([v], dtype=np.int32) if k == "kinmnrl_irreversible"
                  else np.array([float(v)])) for k, v in opts.items()}
    return dataclasses.replace(wl, net=dataclasses.replace(wl.net, **arrays))


def prefactor_workload(ncell=512, seed=5):
    """the calcite workload with calcite's rate given by two parallel rate
    prefactors (PREFACTOR blocks, reaction_mineral.F90:820-870):
    1e-5 a(H+)^0.37 / (1 + 1e9 a(H+)) a(Ca++)^0.31 and, with an activation
    energy of 20 kJ/mol, 3e-8 a(HCO3-)^0.5; per-cell temperatures U[15, 45] C
    (log K stays at its 25 C value: isothermal options)"""
    from pfchem.network import Prefactor
    wl = workloads.calcite_workload(ncell)
    net = wl.net
    ids = {n: i for i, n in enumerate(net.primary_species_names)}
    pf = [[Prefactor(1.0e-5, [(ids["H+"], 0.37, 1.0, 1.0e9), (ids["Ca++"], 0.31, 0.0, 0.0)]),
           Prefactor(3.0e-8, [(ids["HCO3-"], 0.5, 0.0, 0.0)], activation_energy=2.0e4)]]
    st = wl.state.copy()
    rng = np.random.default_rng(seed)
    st.arrays["temperature"][0, :ncell] = rng.uniform(15.0, 45.0, ncell)
    return dataclasses.replace(wl, net=dataclasses.replace(net, kinmnrl_prefactors=pf), state=st)
