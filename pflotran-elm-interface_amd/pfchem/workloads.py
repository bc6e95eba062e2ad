"""Seeded synthetic workloads for the chemistry step.

Networks and per-cell states are drawn from numpy's Philox bit generator,
so the same (seed, sizes) gives bit-identical inputs on every host.

calcite_workload (1000-cell grid, 3 primaries):
    H+ / HCO3- / Ca++ with the five complexes of calcite.dat (OH-, CO3--,
    CO2(aq), CaHCO3+, CaCO3(aq); stoichiometry and Z/a0 from
    regression_tests/ascem/batch/calcite.dat:4-20) and one kinetic calcite
    (-H+ + Ca++ + HCO3-, molar volume 36.934 cm^3/mol, calcite.dat:26),
    k = 1e-9 mol/m^2/s, area 1 m^2/m^3, T = 25 C, dt = 3600 s, porosity 0.3,
    saturation 1, V = 1 m^3, water density 997.05 kg/m^3.
random_network_workload (15 primaries, 60 complexes, 5 minerals):
    random networks with integer stoichiometry.

Construction shared by both (stated here because the sizes alone do not
fix it):
  * log K values are drawn uniformly from the stated interval, and a draw is
    kept only if the complex's molality at a reference composition lies in a
    sane window (rejection sampling).  Without this filter random log K over
    [-20, 15] with coefficients up to 3 gives complexes of 1e+30 mol/kg.
  * each cell's previous free-ion molalities m0 are drawn; its previous
    totals T0 follow from mass action with Debye-Hueckel activity
    coefficients (LAG form, three fixed-point passes); the totals after
    "transport" are T = (1-w) T0 + w T0[neighbour], w ~ U[0, wmax].  Totals
    that are convex mixtures of attainable totals are attainable, so every
    cell has a solution; the guess m0 is not it, so Newton has work to do.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import dataclasses
import os

import numpy as np

from .network import CellState, ReactionNetwork, SolverOptions, DEBYE_A, DEBYE_B, DEBYE_BDOT
from . import abi

LOG_TO_LN = 2.30258509299  # pflotran_constants.F90:84
WATER_DENSITY_25C = 997.05  # kg/m^3


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(key=int(seed)))


@dataclass
class Workload:
    name: str
    net: ReactionNetwork
    opt: SolverOptions
    state: CellState
    dt: float
    seed: int
    notes: str = ""


def bench_options(**kw) -> SolverOptions:
    """Solver controls used for the cells/s metric: the reference defaults
    except (a) the convergence tests, tightened so that accepted iterates have
    relative residual well below 1e-10 (the reference accepts at 1e-8 or on a
    1e-6 relative change, reaction_aux.F90:533-537) and lie within about
    1e-10 of the converged root: the relative-residual exit bounds the
    residual, not the error, and at 1e-12 it leaves up to 2e-8 of relative
    error in the FP64 answer of the 15-primary network (measured against
    40-iteration Newton roots), so two correct solvers that stop at
    different iterates would differ by that much; and (b)
    MAXIMUM_REACTION_ITERATIONS 50 (default 20): a few cells of the random
    network need 26 Newton iterations and no sub-stepping helps them."""
    o = SolverOptions(
        use_log_formulation=True,
        act_coef_update_frequency=abi.ACT_COEF_FREQUENCY_TIMESTEP,
        act_coef_update_algorithm=abi.ACT_COEF_ALGORITHM_LAG,
        maximum_reaction_iterations=50,
        maximum_reaction_cuts=10,
        max_dlnC_rreact=5.0,
        max_residual_tolerance=1.0e-30,
        max_rel_residual_tolerance=1.0e-14,
        max_relative_change_tolerance=1.0e-13,
    )
    for k, v in kw.items():
        setattr(o, k, v)
    return o


# ---------------------------------------------------------------------------
# vectorized mass action (generator only; the oracle and kernels have their own)
# ---------------------------------------------------------------------------
def _lag_gamma(net: ReactionNetwork, m: np.ndarray, sec: np.ndarray):
    """Debye-Hueckel LAG pass (same formula as reaction.F90:4555-4600)."""
    Zp = net.primary_spec_Z[:, None]
    Zc = net.eqcplx_Z[:, None]
    I = 0.5 * ((m * Zp * Zp).sum(0) + (sec * Zc * Zc).sum(0))
    sI = np.sqrt(I)

    def g(Z, a0):
        # one evaluation per distinct (Z^2, a0): the same values, computed once
        Z = Z[:, 0]
        keys = np.stack([Z * Z, np.broadcast_to(a0[:, 0], Z.shape)], axis=1)
        uniq, inv = np.unique(keys, axis=0, return_inverse=True)
        uz = np.sqrt(uniq[:, 0])[:, None]
        ua = uniq[:, 1][:, None]
        lg = -uz * uz * sI * net.debyeA / (1.0 + ua * net.debyeB * sI) + net.debyeBdot * I
        out = np.exp(lg * LOG_TO_LN)[inv.reshape(-1)]
        return np.where(np.abs(Z)[:, None] > 1e-10, out, 1.0)

    return g(Zp, net.primary_spec_a0[:, None]), g(Zc, net.eqcplx_a0[:, None]), I


def _speciate(net: ReactionNetwork, m: np.ndarray, gp: np.ndarray, gc: np.ndarray,
              logK: Optional[np.ndarray] = None):
    """sec molality and totals (mol/kg) at free molality m with fixed gammas."""
    S = net.dense_stoichiometry()
    lnact = np.log(m) + np.log(gp)
    lk = net.eqcplx_logK if logK is None else logK
    lnQK = -lk[:, None] * LOG_TO_LN + S @ lnact
    if net.eqcplxh2ostoich is not None:
        pass  # ln a_w = 0 (use_activity_h2o off)
    sec = np.exp(lnQK) / gc
    tot = m + S.T @ sec
    return sec, tot


def _equilibrate_gamma(net, m, passes=3):
    """column-parallel over cell chunks (element-wise numpy work releases
    the GIL; every cell's result is independent of the chunking)"""
    n = m.shape[1]
    chunk = 32768
    if n <= chunk:
        return _equilibrate_gamma_cols(net, m, passes)
    from concurrent.futures import ThreadPoolExecutor
    parts = [(a, min(n, a + chunk)) for a in range(0, n, chunk)]
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        res = list(ex.map(lambda ab: _equilibrate_gamma_cols(net, m[:, ab[0]:ab[1]], passes),
                          parts))
    return tuple(np.concatenate([r[q] for r in res], axis=1) for q in range(4))


def _equilibrate_gamma_cols(net, m, passes=3):
    sec = np.zeros((net.neqcplx, m.shape[1]))
    gp = np.ones_like(m)
    gc = np.ones_like(sec)
    for _ in range(passes):
        sec, _ = _speciate(net, m, gp, gc)
        gp, gc, I = _lag_gamma(net, m, sec)
    sec, tot = _speciate(net, m, gp, gc)
    return sec, tot, gp, gc


def _draw_positive_totals(net, draw, ncell, rounds=200):
    """Draw free molalities (draw(k) -> (naq, k)) until every total is
    positive.  RStep replaces totals <= 1e-40 by 1e-40 (reaction.F90:3617-3631),
    which would turn a negative total into a different problem."""
    m0 = draw(ncell)
    check = np.arange(ncell)  # cells drawn since their totals were checked
    for _ in range(rounds):
        _, T0, _, _ = _equilibrate_gamma(net, m0[:, check])
        bad = check[np.where((T0 <= 0.0).any(axis=0))[0]]
        if bad.size == 0:
            return m0
        m0[:, bad] = draw(bad.size)
        check = bad
    raise RuntimeError("could not draw cells with positive totals")


def _fill_common(st: CellState, net: ReactionNetwork, m0, sec0, gp, gc, T_molal,
                 den, porosity, sat=1.0, volume=1.0, temp=25.0):
    n = st.ncell
    st["pri_molal"][:, :n] = m0
    st["sec_molal"][:, :n] = sec0
    st["pri_act_coef"][:, :n] = gp
    st["sec_act_coef"][:, :n] = gc
    st["total"][:, :n] = T_molal * den * 1.0e-3  # mol/kg -> mol/L
    st["porosity"][0, :n] = porosity
    st["saturation"][0, :n] = sat
    st["den_kg"][0, :n] = den
    st["volume"][0, :n] = volume
    st["temperature"][0, :n] = temp


# ---------------------------------------------------------------------------
# calcite system on a 10x10x10 grid
# ---------------------------------------------------------------------------
def calcite_network(seed: int = 20261203) -> ReactionNetwork:
    rng = _rng(seed)
    # H+, HCO3-, Ca++  (calcite.dat:4-6: a0, Z)
    Zp = [1.0, -1.0, 2.0]
    a0p = [9.0, 4.0, 6.0]
    names = ["OH-", "CO3--", "CO2(aq)", "CaHCO3+", "CaCO3(aq)"]
    species = [[0], [0, 1], [0, 1], [2, 1], [0, 2, 1]]
    stoich = [[-1.0], [-1.0, 1.0], [1.0, 1.0], [1.0, 1.0], [-1.0, 1.0, 1.0]]
    h2o = [1.0, 0.0, -1.0, 0.0, 0.0]
    Zc = [-1.0, -2.0, 0.0, 1.0, 0.0]
    a0c = [3.5, 4.5, 3.0, 4.0, 3.0]
    # reference composition: pH 7, HCO3- = Ca++ = 1e-3 molal, unit activity
    xref = np.array([-7.0, -3.0, -3.0])
    logK = np.zeros(5)
    for k in range(5):
        nu = np.zeros(3)
        for j, v in zip(species[k], stoich[k]):
            nu[j] += v
        while True:
            lk = rng.uniform(-14.0, 10.0)
            lm = -lk + nu @ xref
            if -14.0 <= lm <= -2.0:
                logK[k] = lk
                break
    nu_cal = np.array([-1.0, 1.0, 1.0])
    while True:
        lk_cal = rng.uniform(-14.0, 10.0)
        if -2.0 <= -lk_cal + nu_cal @ xref <= 2.0:
            break
    net = ReactionNetwork(
        primary_species_names=["H+", "HCO3-", "Ca++"],
        primary_spec_Z=Zp, primary_spec_a0=a0p,
        eqcplx_names=names, eqcplx_species=species, eqcplx_stoich=stoich,
        eqcplxh2ostoich=h2o, eqcplx_logK=logK, eqcplx_Z=Zc, eqcplx_a0=a0c,
        kinmnrl_names=["Calcite"], kinmnrl_species=[[0, 2, 1]],
        kinmnrl_stoich=[[-1.0, 1.0, 1.0]], kinmnrl_logK=[lk_cal],
        kinmnrl_rate_constant=[1.0e-9], kinmnrl_molar_vol=[36.934e-6],
    )
    return net


def calcite_workload(ncell: int = 1000, seed: int = 20261203, dt: float = 3600.0,
                     wmax: float = 0.5, ld: Optional[int] = None) -> Workload:
    net = calcite_network(seed)
    rng = _rng(seed + 1)

    def draw(k):
        pH = rng.uniform(6.0, 8.0, k)
        lhco3 = rng.uniform(np.log10(1e-6), np.log10(1e-2), k)
        lca = rng.uniform(np.log10(1e-6), np.log10(1e-2), k)
        return np.vstack([10.0 ** (-pH), 10.0 ** lhco3, 10.0 ** lca])

    # no positivity filter here: at pH 6-8 with the drawn log K the H+ total
    # is negative in almost every cell, and RStep then solves with
    # total H+ = 1e-40 (reaction.F90:3617-3631), as the reference does
    m0 = draw(ncell)
    sec0, T0, gp, gc = _equilibrate_gamma(net, m0)
    # neighbour along x of the 10x10x10 grid (periodic), mixing weight w
    nb = (np.arange(ncell) + 1) % ncell
    w = rng.uniform(0.0, wmax, ncell)
    T = (1.0 - w) * T0 + w * T0[:, nb]
    st = CellState(net, ncell, ld)
    _fill_common(st, net, m0, sec0, gp, gc, T, WATER_DENSITY_25C, 0.3)
    st["mnrl_volfrac"][0, :ncell] = 0.1
    st["mnrl_area"][0, :ncell] = 1.0
    notes = ("calcite system, log K drawn U[-14,10] under a sanity filter; "
             "calcite log K drawn so log QK at pH 7 / 1e-3 molal is in [-2,2]; "
             "initial calcite volume fraction 0.1")
    return Workload("calcite_1000", net, bench_options(), st, dt, seed, notes)


# ---------------------------------------------------------------------------
# random 15-primary / 60-complex / 5-mineral network
# ---------------------------------------------------------------------------
def random_network(naq: int = 15, ncplx: int = 60, nmnrl: int = 5, seed: int = 20261204,
                   logk_range=(-20.0, 15.0), logrk_range=(-12.0, -6.0),
                   nbackground: int = 2) -> ReactionNetwork:
    """Primaries 0..nbackground-1 are a background electrolyte (Z = +1, -1)
    that sets the ionic strength and takes part in no reaction."""
    rng = _rng(seed)
    Z = rng.integers(-2, 3, naq).astype(float)
    if nbackground >= 2:
        Z[0], Z[1] = 1.0, -1.0
    a0 = np.full(naq, 4.0)
    active = np.arange(nbackground, naq)
    xref = rng.uniform(-6.0, -2.0, naq)
    species, stoich, logK, Zc, a0c = [], [], [], [], []
    while len(species) < ncplx:
        nsp = int(rng.integers(2, 5))
        sp = rng.choice(active, size=nsp, replace=False)
        nu = rng.integers(1, 4, nsp) * rng.choice([-1, 1], nsp)
        if np.all(nu < 0):
            continue
        if abs(float(nu @ Z[sp])) > 3.0:
            continue  # keep complex charges physical (|Z| <= 3)
        lk = rng.uniform(*logk_range)
        lm = -lk + float(nu @ xref[sp])
        if not (-16.0 <= lm <= -3.0):
            continue
        # a negative coefficient may remove at most 2% of that primary at the
        # reference composition, so totals stay positive (see
        # _draw_positive_totals)
        neg = nu < 0
        if np.any(np.log10(-nu[neg]) + lm > xref[sp][neg] + np.log10(0.02)):
            continue
        species.append([int(s) for s in sp])
        stoich.append([float(v) for v in nu])
        logK.append(lk)
        zk = float(nu @ Z[sp])
        Zc.append(zk)
        a0c.append(4.0 if zk != 0 else 3.0)
    msp, mst, mlogK = [], [], []
    while len(msp) < nmnrl:
        nsp = int(rng.integers(2, 5))
        sp = rng.choice(active, size=nsp, replace=False)
        nu = rng.integers(1, 4, nsp) * rng.choice([-1, 1], nsp, p=[0.25, 0.75])
        if np.all(nu < 0):
            continue
        lq = float(nu @ xref[sp])
        msp.append([int(s) for s in sp])
        mst.append([float(v) for v in nu])
        mlogK.append(lq - rng.uniform(-1.0, 1.0))  # log QK at reference in [-1, 1]
    rate = 10.0 ** rng.uniform(*logrk_range, nmnrl)
    net = ReactionNetwork(
        primary_species_names=["P%d" % i for i in range(naq)],
        primary_spec_Z=Z, primary_spec_a0=a0,
        eqcplx_names=["C%d" % k for k in range(ncplx)], eqcplx_species=species,
        eqcplx_stoich=stoich, eqcplx_logK=logK, eqcplx_Z=Zc, eqcplx_a0=a0c,
        kinmnrl_names=["M%d" % m for m in range(nmnrl)], kinmnrl_species=msp,
        kinmnrl_stoich=mst, kinmnrl_logK=mlogK, kinmnrl_rate_constant=rate,
        kinmnrl_molar_vol=np.full(nmnrl, 3.0e-5),
    )
    net._xref = xref
    net._nbackground = nbackground
    return net


def random_network_workload(ncell: int = 1 << 20, naq: int = 15, ncplx: int = 60,
                            nmnrl: int = 5, seed: int = 20261204, dt: float = 3600.0,
                            wmax: float = 0.3, ld: Optional[int] = None,
                            spread: float = 0.25, variable_ncomp: bool = False) -> Workload:
    net = random_network(naq, ncplx, nmnrl, seed)
    rng = _rng(seed + 1)
    nb_ = net._nbackground
    m0 = _draw_positive_totals(
        net, lambda k: 10.0 ** (net._xref[:, None] + rng.uniform(-spread, spread, (naq, k))),
        ncell)
    I_target = 10.0 ** rng.uniform(np.log10(1e-3), np.log10(0.5), ncell)
    # background electrolyte sets the ionic strength
    m0[:nb_] = 1e-6
    for _ in range(2):
        sec, _, gp, gc = _equilibrate_gamma(net, m0, passes=2)
        Zp = net.primary_spec_Z[:, None]
        Zc = net.eqcplx_Z[:, None]
        I_other = 0.5 * ((m0[nb_:] * Zp[nb_:] ** 2).sum(0) + (sec * Zc ** 2).sum(0))
        mb = np.maximum(I_target - I_other, 1e-3 * I_target)
        m0[0] = mb
        m0[1] = mb
    sec0, T0, gp, gc = _equilibrate_gamma(net, m0)
    nb = (np.arange(ncell) + 1) % ncell
    w = rng.uniform(0.0, wmax, ncell)
    T = (1.0 - w) * T0 + w * T0[:, nb]
    st = CellState(net, ncell, ld, variable_ncomp=variable_ncomp)
    _fill_common(st, net, m0, sec0, gp, gc, T, WATER_DENSITY_25C, 0.3)
    # active-mineral subset per cell: count ~ U{0..nmnrl}
    cnt = rng.integers(0, nmnrl + 1, ncell)
    keys = rng.random((nmnrl, ncell))
    rank = np.argsort(np.argsort(keys, axis=0), axis=0)
    act = rank < cnt[None, :]
    st["mnrl_area"][:, :ncell] = np.where(act, 1.0, 0.0)
    st["mnrl_volfrac"][:, :ncell] = np.where(act, rng.uniform(0.0, 0.05, (nmnrl, ncell)), 0.0)
    notes = ("random network: Z~U{-2..2} (P0/P1 background electrolyte Z=+1/-1 setting "
             "I~logU[1e-3,0.5]), a0=4; complexes on 2-4 of P2..P14 with nu in "
             "+-{1,2,3} and |Z|<=3, log K~U[-20,15] kept if log m_k at the reference "
             "composition is in [-16,-3]; minerals log k~U[-12,-6], log QK at "
             "reference in [-1,1], area 1 m^2/m^3 when active, active count "
             "U{0..5} per cell")
    return Workload("random_n%d_c%d_m%d" % (naq, ncplx, nmnrl), net, bench_options(), st,
                    dt, seed, notes)


def variable_network_workload(ncell: int = 262_144, seed: int = 20261206, dt: float = 3600.0,
                              nmin: int = 8, ld: Optional[int] = None) -> Workload:
    """Large network of 64 primaries, 300 complexes and 10 kinetic minerals
    (the random-network generator) in which each cell has its own number of
    active primaries, U{nmin..64}: the cell solves the network restricted to
    primaries 0..nact-1 (complexes and minerals all of whose species are
    among them), so its system is nact x nact."""
    wl = random_network_workload(ncell, naq=64, ncplx=300, nmnrl=10, seed=seed, dt=dt, ld=ld,
                                 variable_ncomp=True)
    rng = _rng(seed + 2)
    wl.state["ncomp_active"][0, :ncell] = rng.integers(nmin, 65, ncell)
    # relative-residual tolerance 1e-13: sums of up to 64 terms leave some
    # cells' residuals above 1e-14 of their initial value at the rounding
    # level, so at 1e-14 their Newton iteration would stop only by cutting
    # the step
    opt = dataclasses.replace(wl.opt, variable_ncomp=True, max_rel_residual_tolerance=1.0e-13)
    notes = wl.notes + ("; 64 primaries, 300 complexes, 10 minerals; per-cell active "
                        "primaries U{%d..64} (prefix 0..nact-1)" % nmin)
    return Workload("variable_n64_c300_m10", wl.net, opt, wl.state, dt, seed, notes)


# ---------------------------------------------------------------------------
# SomDec decomposition networks (ELM-CN litter / SOM cascade)
# ---------------------------------------------------------------------------
SOMDEC_PRIMARIES = ["CO2(aq)", "NH4+", "NO3-", "N2O(aq)", "H+"]
SOMDEC_IMMOBILE = ["SOM1", "SOM2", "SOM3", "SOM4", "Lit1C", "Lit2C", "Lit3C",
                   "Lit1N", "Lit2N", "Lit3N"]


def somdec_network(rates=None, som_cn=(12.0, 12.0, 10.0, 10.0), half_saturation=1.0e-4,
                   temperature_function=0, Q10=2.0,
                   oxygen: Optional[float] = None, complexes: bool = False) -> ReactionNetwork:
    """The CLM-CN cascade of the clm_cn3 regression deck
    (regression_tests/ngee/CLMCNplus/clm_cn3.in): litter pools Lit1-3 with
    variable C:N, SOM1-4 with fixed C:N, Monod NH4+/NO3- limitation on the
    litter reactions.  rates: 7 first-order rate constants (1/s) in the
    deck's reaction order, default the deck's turnover times.  oxygen: a
    Monod oxygen response of that half saturation (mol/L) on every reaction
    through an O2(aq) primary, whose consumption is tracked
    (O2_SPECIES_NAME); None = no oxygen.  complexes: two synthetic aqueous
    ion pairs, NH4NO3(aq) = NH4+ + NO3- (log K -4) and NH4CO2(aq) = NH4+ +
    CO2(aq) (log K -3.5), so the sandbox's totals and dtotal carry
    complexation."""
    from . import somdec as SD
    h, d, y = 3600.0, 86400.0, 365.0 * 86400.0
    turnover = [20 * h, 14 * d, 71 * d, 14 * d, 71 * d, 2 * y, 27.4 * y]
    if rates is None:
        rates = [1.0 / t for t in turnover]
    pools = [SD.Pool("SOM%d" % (i + 1), som_cn[i]) for i in range(4)] + \
        [SD.Pool("Lit1"), SD.Pool("Lit2"), SD.Pool("Lit3")]
    mon = [SD.Monod("NH4+", half_saturation), SD.Monod("NO3-", half_saturation)]
    cascade = [("Lit1", [("SOM1", 0.61)], mon), ("Lit2", [("SOM2", 0.45)], mon),
               ("Lit3", [("SOM3", 0.71)], mon), ("SOM1", [("SOM2", 0.72)], []),
               ("SOM2", [("SOM3", 0.54)], []), ("SOM3", [("SOM4", 0.45)], []),
               ("SOM4", [], [])]
    ox = {} if oxygen is None else dict(Ox_response_function=SD.OX_MONOD,
                                        Ox_half_saturation=oxygen, Ox_species="O2(aq)")
    rx = [SD.Reaction(up, down, rate_constant=k, monod=list(m),
                      temperature_response_function=temperature_function, Q10=Q10, **ox)
          for (up, down, m), k in zip(cascade, rates)]
    prim = list(SOMDEC_PRIMARIES) + ([] if oxygen is None else ["O2(aq)"])
    sd = SD.build_somdec(pools, rx, prim, SOMDEC_IMMOBILE,
                         o2_species=None if oxygen is None else "O2(aq)")
    cx = {}
    if complexes:
        ip = {name: i for i, name in enumerate(prim)}
        cx = dict(eqcplx_names=["NH4NO3(aq)", "NH4CO2(aq)"],
                  eqcplx_species=[[ip["NH4+"], ip["NO3-"]], [ip["NH4+"], ip["CO2(aq)"]]],
                  eqcplx_stoich=[[1.0, 1.0], [1.0, 1.0]], eqcplx_logK=np.array([-4.0, -3.5]))
    return ReactionNetwork(primary_species_names=prim, primary_spec_Z=np.zeros(len(prim)),
                           primary_spec_a0=np.zeros(len(prim)),
                           immobile_names=list(SOMDEC_IMMOBILE), somdec=sd, **cx)


def somdec_options(**kw) -> SolverOptions:
    return bench_options(act_coef_update_frequency=abi.ACT_COEF_FREQUENCY_OFF, **kw)


def somdec_workload(ncell: int = 500_000 * 15, seed: int = 20261205, dt: float = 1800.0,
                    ld: Optional[int] = None, oxygen: Optional[float] = None,
                    complexes: bool = False) -> Workload:
    """ELM-CN decomposition columns (500,000 columns x 15 layers by default):
    per-reaction first-order rates drawn U[1e-8, 1e-5] /s, SOM C:N (mass)
    U[8, 30], Monod half-saturation 1e-6 mol/m^3; per cell a temperature
    U[0, 35] C acting through Q10 = 2 and a moisture scalar f_w U[0, 1]
    (the ELM w_scalar input); litter C (mol/m^3) logU[1e-2, 1] with C:N
    (mass) U[20, 80], SOM C logU[1e-3, 1], NH4+ / NO3- logU[1e-7, 1e-4]
    mol/L, CO2 logU[1e-6, 1e-3] mol/L, N2O 1e-20 mol/L, pH 6; with oxygen
    (a Monod half saturation, mol/L) an O2(aq) primary logU[2e-4, 1e-3]
    mol/L limits every reaction and is consumed as CO2 is released; with
    complexes the two ion pairs of somdec_network (free-molality guess =
    totals, unit activity coefficients)."""
    rng = _rng(seed)
    rates = 10.0 ** rng.uniform(-8.0, -5.0, 7)
    som_cn = rng.uniform(8.0, 30.0, 4)
    net = somdec_network(rates=list(rates), som_cn=tuple(som_cn), half_saturation=1.0e-6,
                         temperature_function=2, Q10=2.0, oxygen=oxygen, complexes=complexes)
    st = CellState(net, ncell, ld, somdec_scalars=True)
    n = ncell
    lu = lambda lo, hi, k: 10.0 ** rng.uniform(np.log10(lo), np.log10(hi), k)
    T = np.vstack([lu(1e-6, 1e-3, n), lu(1e-7, 1e-4, n), lu(1e-7, 1e-4, n),
                   np.full(n, 1e-20), np.full(n, 1e-6)] +
                  ([] if oxygen is None else [lu(2e-4, 1e-3, n)]))
    st["total"][:, :n] = T
    st["pri_molal"][:, :n] = T          # no complexes, unit density below
    st["pri_act_coef"][:, :n] = 1.0
    if complexes:
        st["sec_act_coef"][:, :n] = 1.0
    litc = np.vstack([lu(1e-2, 1.0, n) for _ in range(3)])
    litcn = rng.uniform(20.0, 80.0, (3, n))
    from .somdec import CN_RATIO_MASS_TO_MOL
    im = np.vstack([lu(1e-3, 1.0, n) for _ in range(4)] + [litc] +
                   [litc / litcn / CN_RATIO_MASS_TO_MOL])
    st["immobile"][:, :n] = im
    st["porosity"][0, :n] = rng.uniform(0.3, 0.6, n)
    st["saturation"][0, :n] = 1.0
    st["den_kg"][0, :n] = 1000.0
    st["volume"][0, :n] = 1.0
    st["temperature"][0, :n] = rng.uniform(0.0, 35.0, n)
    st["somdec_f_t"][0, :n] = 1.0
    st["somdec_f_w"][0, :n] = rng.uniform(0.0, 1.0, n)
    notes = ("SomDec ELM-CN cascade (clm_cn3 structure); rates U[1e-8,1e-5]/s per reaction, "
             "SOM C:N U[8,30], Q10 = 2 with per-cell temperature, per-cell f_w U[0,1]" +
             ("" if oxygen is None else ", Monod O2 response K = %g mol/L with O2 tracking" % oxygen) +
             (", NH4NO3(aq) / NH4CO2(aq) ion pairs" if complexes else ""))
    return Workload("somdec_%d" % ncell, net, somdec_options(), st, dt, seed, notes)
