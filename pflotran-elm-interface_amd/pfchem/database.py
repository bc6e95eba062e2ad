"""Thermodynamic database reader and basis builder for the reaction network.

Reads PFLOTRAN's text database format (reaction_database.F90 DatabaseRead:
26-...; e.g. calcite.dat, carbonate.dat):

    'temperature points' n T_1 .. T_n
    'name' a0 Z mw                                   primary species ...
    'null' 0 0 0
    'name' nsp (nu 'sp')*nsp logK*n a0 Z mw          aqueous complexes ...
    'null' 1 0. '0' 0. ...
    'name' V_m nsp (nu 'sp')*nsp logK*n mw           gases ...
    'null' 0. 1 1. '0' 0. ...
    'name' V_m nsp (nu 'sp')*nsp logK*n mw           minerals ...
    'null' 0. 1 0. '0' 0. ...
    (surface complexes, not used on this path)

and builds a `ReactionNetwork` for chosen primary / secondary species and
kinetic minerals, the way BasisInit does for a basis that is the chosen
primary set (reaction_database.F90:812, compressed arrays :1704-1772):
water is never a component (its coefficient goes to eqcplxh2ostoich /
kinmnrlh2ostoich); log K at the reference temperature comes from the
five-term least-squares fit over the database temperatures
(ReactionFitLogKCoef + ReactionInitializeLogK, reaction_aux.F90:1159-1281),
or, for isothermal runs, from the tabulated value at that temperature
(reaction_database.F90:1775-1781); mineral molar volumes are converted
from cm^3/mol to m^3/mol (reaction_mineral.F90:459).

The chosen primaries need not be the database's primaries (basis
swapping, BasisInit :1168-1448): every chosen species that has a database
reaction (a chosen primary that is a database secondary, and every chosen
secondary that is one) gives one row [-1 at itself, nu at each species of
its reaction], split into the columns of H2O + the chosen primaries and the
columns of the chosen secondaries; with S^-1 the inverse of the square
secondary block, the secondaries' stoichiometry in the chosen basis is
-S^-1 P and their log K vectors (all database temperatures) -S^-1 logK.  A
kinetic mineral's reaction then has each chosen secondary it names replaced
by that secondary's reaction, log K += nu * log K (BasisSubSpeciesIn-
MineralRxn, reaction_database_aux.F90:494-586; stoichiometries that cancel
below 1e-10 are dropped).
"""
from __future__ import annotations

import shlex
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from .network import ReactionNetwork

# missing-data marker of the database (reaction_aux.F90:1197)
NO_DATA = 500.0
WATER = "H2O"


@dataclass
class DbReaction:
    name: str
    species: List[str]
    stoich: List[float]
    logK: np.ndarray            # at the database temperatures
    a0: float = 0.0
    Z: float = 0.0
    molar_weight: float = 0.0
    molar_volume: float = 0.0   # minerals / gases, as read (cm^3/mol)


@dataclass
class Database:
    temperatures: np.ndarray
    primary: Dict[str, DbReaction] = field(default_factory=dict)
    secondary: Dict[str, DbReaction] = field(default_factory=dict)
    gases: Dict[str, DbReaction] = field(default_factory=dict)
    minerals: Dict[str, DbReaction] = field(default_factory=dict)


def _tokens(path: str) -> List[List[str]]:
    """one token list per non-empty record; quoted names keep their spaces"""
    out = []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line or line.startswith("#") or line.startswith("!"):
                continue
            out.append(shlex.split(line, posix=True))
    return out


def read_database(path: str) -> Database:
    recs = _tokens(path)
    if not recs or recs[0][0] != "temperature points":
        raise ValueError("%s: first record must be 'temperature points'" % path)
    nt = int(float(recs[0][1]))
    temps = np.array([float(v) for v in recs[0][2:2 + nt]])
    db = Database(temperatures=temps)
    section = 0  # 0 primary, 1 secondary, 2 gas, 3 mineral, 4 surface
    for r in recs[1:]:
        name = r[0]
        if name == "null":
            section += 1
            continue
        if section == 0:
            db.primary[name] = DbReaction(name, [], [], np.zeros(0), a0=float(r[1]),
                                          Z=float(r[2]), molar_weight=float(r[3]))
        elif section in (1, 2, 3):
            p = 1
            vm = 0.0
            if section in (2, 3):
                vm = float(r[p])
                p += 1
            nsp = int(float(r[p]))
            p += 1
            sp, nu = [], []
            for _ in range(nsp):
                nu.append(float(r[p]))
                sp.append(r[p + 1])
                p += 2
            logK = np.array([float(v) for v in r[p:p + nt]])
            p += nt
            if section == 1:
                rx = DbReaction(name, sp, nu, logK, a0=float(r[p]), Z=float(r[p + 1]),
                                molar_weight=float(r[p + 2]))
                db.secondary[name] = rx
            else:
                rx = DbReaction(name, sp, nu, logK, molar_volume=vm,
                                molar_weight=float(r[p]))
                (db.gases if section == 2 else db.minerals)[name] = rx
        # surface complexation records are not used on this path
    return db


# ---------------------------------------------------------------------------
# log K fit (ReactionFitLogKCoef / ReactionInitializeLogK /
# ReactionInterpolateLogK, reaction_aux.F90:1159-1312)
# ---------------------------------------------------------------------------
def _crout_lu_solve(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Crout LU with implicit partial pivoting, zero pivots replaced by
    1e-20, then forward / back substitution (LUDecomposition1 /
    LUBackSubstitution, utility.F90:597-735), in plain FP64 scalar order."""
    a = [list(map(float, row)) for row in a]
    b = list(map(float, b))
    n = len(b)
    vv = [0.0] * n
    indx = [0] * n
    for i in range(n):
        aamax = 0.0
        for j in range(n):
            aamax = max(aamax, abs(a[i][j]))
        if aamax == 0.0:
            raise ValueError("singular matrix (zero row) in the log K fit")
        vv[i] = 1.0 / aamax
    imax = 0
    for j in range(n):
        for i in range(j):
            s = a[i][j]
            for k in range(i):
                s = s - a[i][k] * a[k][j]
            a[i][j] = s
        aamax = 0.0
        for i in range(j, n):
            s = a[i][j]
            for k in range(j):
                s = s - a[i][k] * a[k][j]
            a[i][j] = s
            dum = vv[i] * abs(s)
            if dum >= aamax:
                imax = i
                aamax = dum
        if j != imax:
            a[imax], a[j] = a[j], a[imax]
            vv[imax] = vv[j]
        indx[j] = imax
        if a[j][j] == 0.0:
            a[j][j] = 1.0e-20
        if j != n - 1:
            dum = 1.0 / a[j][j]
            for i in range(j + 1, n):
                a[i][j] = a[i][j] * dum
    ii = -1
    for i in range(n):
        ll = indx[i]
        s = b[ll]
        b[ll] = b[i]
        if ii >= 0:
            for j in range(ii, i):
                s = s - a[i][j] * b[j]
        elif s != 0.0:
            ii = i
        b[i] = s
    for i in range(n - 1, -1, -1):
        s = b[i]
        for j in range(i + 1, n):
            s = s - a[i][j] * b[j]
        b[i] = s / a[i][i]
    return np.array(b)


def fit_logk_coef(temperatures: Sequence[float], logK: Sequence[float]) -> np.ndarray:
    """five coefficients of logK(T) = c1 ln T + c2 + c3 T + c4/T + c5/T^2
    (T in K) by normal equations over the points that are not NO_DATA"""
    nt = len(temperatures)
    vec = np.zeros((5, nt))
    for i, tc in enumerate(temperatures):
        tk = float(tc) + 273.15
        vec[:, i] = [np.log(tk), 1.0, tk, 1.0 / tk, 1.0 / (tk * tk)]
    use = [abs(float(v) - NO_DATA) >= 1.0e-10 for v in logK]
    coefs = np.zeros(5)
    for j in range(5):
        s = 0.0
        for i in range(nt):
            if use[i]:
                s = s + vec[j, i] * float(logK[i])
        coefs[j] = s
    a = np.zeros((5, 5))
    for j in range(5):
        for k in range(j, 5):
            s = 0.0
            for i in range(nt):
                if use[i]:
                    s = s + vec[j, i] * vec[k, i]
            a[j, k] = s
            a[k, j] = s
    return _crout_lu_solve(a, coefs)


def interpolate_logk(coefs: Sequence[float], temp_c: float) -> float:
    tk = temp_c + 273.15
    c = [float(v) for v in coefs]
    return c[0] * np.log(tk) + c[1] + c[2] * tk + c[3] / tk + c[4] / (tk * tk)


def logk_at(temperatures, logK, temp_c: float, isothermal: bool):
    """(logK at temp_c, fit coefficients) the way BasisInit sets them"""
    coefs = fit_logk_coef(temperatures, logK)
    if isothermal:
        # tabulated value at a database temperature, else linear
        # interpolation between neighbours (reaction_database.F90:1775-1781)
        t = np.asarray(temperatures, dtype=np.float64)
        hit = np.where(np.abs(t - temp_c) < 1.0e-10)[0]
        if hit.size:
            return float(logK[hit[0]]), coefs
        hi = int(np.searchsorted(t, temp_c))
        lo = hi - 1
        w = (temp_c - t[lo]) / (t[hi] - t[lo])
        return float(logK[lo] + w * (logK[hi] - logK[lo])), coefs
    return interpolate_logk(coefs, temp_c), coefs


# ---------------------------------------------------------------------------
# basis
# ---------------------------------------------------------------------------
@dataclass
class KineticMineralSpec:
    name: str
    rate_constant: float                  # mol/m^2/s
    activation_energy: float = 0.0        # J/mol
    affinity_power: Optional[float] = None
    Temkin_const: Optional[float] = None
    min_scale_factor: Optional[float] = None
    irreversible: int = 0


def _species(db: Database, name: str) -> DbReaction:
    for table in (db.primary, db.secondary):
        if name in table:
            return table[name]
    raise KeyError("species %r not in the database" % name)


def basis_reactions(db: Database, primaries: Sequence[str], secondaries: Sequence[str]):
    """Reactions of the chosen secondaries in the chosen basis: a list, in
    the order of `secondaries`, of (stoichiometry over H2O + primaries,
    an array of len(primaries) + 1; log K over the database temperatures)."""
    cols = [WATER] + list(primaries)
    pcol = {n: i for i, n in enumerate(cols)}
    scol = {n: i for i, n in enumerate(secondaries)}
    for n in list(primaries) + list(secondaries):
        _species(db, n)
    rows = ([n for n in primaries if n in db.secondary] +
            [n for n in secondaries if n in db.secondary])
    ns = len(secondaries)
    if len(rows) != ns:
        raise ValueError("basis: %d species with database reactions for %d secondaries; "
                         "the chosen primaries and secondaries are not a basis"
                         % (len(rows), ns))
    if ns == 0:
        return []
    P = np.zeros((ns, len(cols)))
    S = np.zeros((ns, ns))
    K = np.zeros((ns, db.temperatures.size))
    for r, name in enumerate(rows):
        rx = db.secondary[name]
        K[r] = rx.logK
        if name in pcol:
            P[r, pcol[name]] -= 1.0
        else:
            S[r, scol[name]] -= 1.0
        for sp, nu in zip(rx.species, rx.stoich):
            if sp in pcol:
                P[r, pcol[sp]] += nu
            elif sp in scol:
                S[r, scol[sp]] += nu
            else:
                raise ValueError("reaction %r names %r, which is neither a chosen primary "
                                 "nor a chosen secondary species" % (name, sp))
    Sinv = np.linalg.inv(S)
    stoich = -Sinv @ P
    logk = -Sinv @ K
    return [(stoich[i], logk[i]) for i in range(ns)]


def _in_basis(rx: DbReaction, cols, sec_rxn):
    """a mineral reaction in the chosen basis: (stoichiometry over H2O +
    primaries, log K over the database temperatures)"""
    pcol = {n: i for i, n in enumerate(cols)}
    st = np.zeros(len(cols))
    logk = np.array(rx.logK, dtype=np.float64)
    for sp, nu in zip(rx.species, rx.stoich):
        if sp in pcol:
            st[pcol[sp]] += nu
        elif sp in sec_rxn:
            s_st, s_lk = sec_rxn[sp]
            st += nu * s_st
            logk = logk + nu * s_lk
        else:
            raise ValueError("reaction %r names %r, which is neither a chosen primary nor a "
                             "chosen secondary species" % (rx.name, sp))
    st[np.abs(st) <= 1e-10] = 0.0
    return st, logk


def build_network(db: Database, primaries: Sequence[str], secondaries: Sequence[str] = (),
                  kinetic_minerals: Sequence[KineticMineralSpec] = (),
                  temperature: float = 25.0, isothermal: bool = False) -> ReactionNetwork:
    """ReactionNetwork of the chosen species, 0-based ids in the order given
    (any basis of the database reactions, see the module docstring)"""
    cols = [WATER] + list(primaries)
    for name in primaries:
        if name not in db.primary and name not in db.secondary:
            raise KeyError("primary species %r not in the database" % name)
    sec = basis_reactions(db, primaries, secondaries)
    sec_rxn = dict(zip(secondaries, sec))

    def split(st):
        """(species ids, stoichiometry, H2O coefficient) of a basis row"""
        sp = [j - 1 for j in range(1, len(cols)) if abs(st[j]) > 1e-40]
        return sp, [float(st[i + 1]) for i in sp], float(st[0])

    c_sp, c_nu, c_h2o, c_logK, c_coef, c_Z, c_a0 = [], [], [], [], [], [], []
    for name, (st, lkv) in zip(secondaries, sec):
        rx = _species(db, name)
        sp, nu, h2o = split(st)
        lk, cf = logk_at(db.temperatures, lkv, temperature, isothermal)
        c_sp.append(sp), c_nu.append(nu), c_h2o.append(h2o)
        c_logK.append(lk), c_coef.append(cf), c_Z.append(rx.Z), c_a0.append(rx.a0)
    m_sp, m_nu, m_h2o, m_logK, m_coef, m_vm = [], [], [], [], [], []
    for spec in kinetic_minerals:
        rx = db.minerals[spec.name]
        st, lkv = _in_basis(rx, cols, sec_rxn)
        sp, nu, h2o = split(st)
        lk, cf = logk_at(db.temperatures, lkv, temperature, isothermal)
        m_sp.append(sp), m_nu.append(nu), m_h2o.append(h2o)
        m_logK.append(lk), m_coef.append(cf), m_vm.append(rx.molar_volume * 1.0e-6)
    nm = len(kinetic_minerals)

    def opt(attr):
        vals = [getattr(s, attr) for s in kinetic_minerals]
        if all(v is None for v in vals):
            return None
        return np.array([1.0 if v is None else v for v in vals])

    net = ReactionNetwork(
        primary_species_names=list(primaries),
        primary_spec_Z=[_species(db, n).Z for n in primaries],
        primary_spec_a0=[_species(db, n).a0 for n in primaries],
        eqcplx_names=list(secondaries),
        eqcplx_species=c_sp, eqcplx_stoich=c_nu,
        eqcplxh2ostoich=np.array(c_h2o) if secondaries and any(c_h2o) else None,
        eqcplx_logK=np.array(c_logK), eqcplx_logKcoef=np.array(c_coef).reshape(-1),
        eqcplx_Z=np.array(c_Z), eqcplx_a0=np.array(c_a0),
        kinmnrl_names=[s.name for s in kinetic_minerals],
        kinmnrl_species=m_sp, kinmnrl_stoich=m_nu,
        kinmnrlh2ostoich=np.array(m_h2o) if nm and any(m_h2o) else None,
        kinmnrl_logK=np.array(m_logK) if nm else None,
        kinmnrl_logKcoef=np.array(m_coef).reshape(-1) if nm else None,
        kinmnrl_rate_constant=np.array([s.rate_constant for s in kinetic_minerals]) if nm else None,
        kinmnrl_activation_energy=(np.array([s.activation_energy for s in kinetic_minerals])
                                   if nm and any(s.activation_energy for s in kinetic_minerals)
                                   else None),
        kinmnrl_molar_vol=np.array(m_vm) if nm else None,
        kinmnrl_affinity_power=opt("affinity_power"),
        kinmnrl_Temkin_const=opt("Temkin_const"),
        kinmnrl_min_scale_factor=opt("min_scale_factor"),
        kinmnrl_irreversible=(np.array([s.irreversible for s in kinetic_minerals])
                              if nm and any(s.irreversible for s in kinetic_minerals) else None),
    )
    net.validate()
    return net


def constraint_reactions(db: Database, primaries: Sequence[str], secondaries: Sequence[str],
                         names: Sequence[str], temperature: float = 25.0,
                         isothermal: bool = False):
    """Mineral or passive-gas reactions of the database (its 'mineral' and
    'gas' entries) written in the basis of `primaries`, for MINERAL and GAS
    constraints (mnrlspecid / paseqspecid as BasisInit builds them,
    reaction_database.F90:1800-1830, 1925-1960): one (species ids, stoichiometry, H2O
    coefficient, log K at `temperature`, log K fit coefficients) per name."""
    cols = [WATER] + list(primaries)
    sec_rxn = dict(zip(secondaries, basis_reactions(db, primaries, secondaries)))
    out = []
    for name in names:
        if name in db.minerals:
            rx = db.minerals[name]
        elif name in db.gases:
            rx = db.gases[name]
        else:
            raise KeyError("%r is neither a mineral nor a gas of the database" % name)
        st, lkv = _in_basis(rx, cols, sec_rxn)
        sp = [j - 1 for j in range(1, len(cols)) if abs(st[j]) > 1e-40]
        lk, cf = logk_at(db.temperatures, lkv, temperature, isothermal)
        out.append((sp, [float(st[i + 1]) for i in sp], float(st[0]), lk, cf))
    return out
