"""Aqueous constraints for the constraint speciation (pfchem_equilibrate).

Mirrors the CONCENTRATIONS block of a PFLOTRAN CONSTRAINT card
(TranConstraintRTRead, transport_constraint_rt.F90:261-360): every primary
species gets a type letter (T total, F free, L log activity, P pH, M mineral,
G gas, Z charge balance) and, for M and G, the mineral or gas name.  The
values themselves are per cell (an SoA array handed to
`Context.equilibrate`), so one constraint definition speciates a whole field
of cells with different concentrations.
"""
import ctypes as C
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import abi, database
from .network import ReactionNetwork, _pack_reactions

# constraint letters and keywords, transport_constraint_rt.F90:301-336
TYPE_OF = {
    "F": abi.CONSTRAINT_FREE, "FREE": abi.CONSTRAINT_FREE,
    "T": abi.CONSTRAINT_TOTAL, "TOTAL": abi.CONSTRAINT_TOTAL,
    "P": abi.CONSTRAINT_PH, "PH": abi.CONSTRAINT_PH,
    "E": abi.CONSTRAINT_PE, "PE": abi.CONSTRAINT_PE,
    "L": abi.CONSTRAINT_LOG, "LOG": abi.CONSTRAINT_LOG,
    "M": abi.CONSTRAINT_MINERAL, "MINERAL": abi.CONSTRAINT_MINERAL,
    "MNRL": abi.CONSTRAINT_MINERAL,
    "G": abi.CONSTRAINT_GAS, "GAS": abi.CONSTRAINT_GAS,
    "SC": abi.CONSTRAINT_SUPERCRIT_CO2,
    "Z": abi.CONSTRAINT_CHARGE_BAL, "CHARGE_BALANCE": abi.CONSTRAINT_CHARGE_BAL,
    "TOTAL_SORB": abi.CONSTRAINT_TOTAL_SORB,
    "TOTAL_AQ_PLUS_SORB": abi.CONSTRAINT_TOTAL_AQ_PLUS_SORB,
}
SUPPORTED = (abi.CONSTRAINT_NULL, abi.CONSTRAINT_FREE, abi.CONSTRAINT_TOTAL,
             abi.CONSTRAINT_LOG, abi.CONSTRAINT_PH, abi.CONSTRAINT_MINERAL,
             abi.CONSTRAINT_GAS, abi.CONSTRAINT_CHARGE_BAL)


@dataclass
class Constraint:
    """One aqueous constraint over the network's primaries (0-based).

    types: PFCHEM_CONSTRAINT_* per primary; rxn: per primary, the row of the
    reaction table for M / G (else -1); the table holds the mineral / gas
    reactions in the network's basis.  h_ion_id: >= 0 the primary that is
    H+, -(k+1) when H+ is complex k, None if H+ is absent."""

    types: List[int]
    rxn: List[int]
    rxn_species: List[List[int]] = field(default_factory=list)
    rxn_stoich: List[List[float]] = field(default_factory=list)
    rxn_h2o: Optional[np.ndarray] = None
    rxn_logK: Optional[np.ndarray] = None
    rxn_logKcoef: Optional[np.ndarray] = None
    h_ion_id: Optional[int] = None
    initialize_with_molality: bool = False
    use_prev_soln_as_guess: bool = False
    max_iterations: int = 10000
    max_dlnC: float = 5.0
    free_ion_guess: Optional[np.ndarray] = None
    rxn_names: List[str] = field(default_factory=list)

    def validate(self, net: ReactionNetwork):
        n = net.naqcomp
        if len(self.types) != n or len(self.rxn) != n:
            raise ValueError("constraint needs one type and one reaction row per primary")
        for i, t in enumerate(self.types):
            if t not in SUPPORTED:
                raise NotImplementedError(
                    "constraint type %d of %s is not supported (redox, sorption and "
                    "supercritical CO2 constraints are outside this library)"
                    % (t, net.primary_species_names[i] if net.primary_species_names else i))
            if t in (abi.CONSTRAINT_MINERAL, abi.CONSTRAINT_GAS):
                if not 0 <= self.rxn[i] < len(self.rxn_species):
                    raise ValueError("component %d: no reaction row for its M/G constraint" % i)
            if t == abi.CONSTRAINT_PH:
                h = self.h_ion_id
                if h is None:
                    raise ValueError("pH constraint, but H+ is not among the species")
                if h >= 0 and h != i:
                    names = net.primary_species_names or []
                    if i >= len(names) or names[i] != "OH-":
                        raise ValueError("pH constraint on a species other than H+ or OH-")

    def to_abi(self, net: ReactionNetwork) -> abi.Constraint:
        """pfchem_constraint struct; keeps its arrays alive on this object"""
        self.validate(net)
        keep = []

        def ip(a):
            a = np.ascontiguousarray(a, dtype=np.int32)
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(C.c_int32))

        def dp(a):
            if a is None:
                return C.cast(None, C.POINTER(C.c_double))
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(C.c_double))

        s = abi.Constraint()
        s.constraint_type = ip(self.types)
        s.constraint_rxn = ip(self.rxn)
        nr = len(self.rxn_species)
        if nr:
            ids, st, ms = _pack_reactions(self.rxn_species, self.rxn_stoich)
            s.nrxn, s.rxn_maxsp = nr, ms
            s.rxnspecid, s.rxnstoich = ip(ids), dp(st)
            s.rxnh2ostoich = dp(self.rxn_h2o)
            s.rxn_logK = dp(self.rxn_logK)
            s.rxn_logKcoef = dp(self.rxn_logKcoef)
        else:
            s.nrxn, s.rxn_maxsp = 0, 1
            s.rxnspecid = C.cast(None, C.POINTER(C.c_int32))
            s.rxnstoich = s.rxnh2ostoich = s.rxn_logK = s.rxn_logKcoef = dp(None)
        s.h_ion_id = -(1 << 30) if self.h_ion_id is None else int(self.h_ion_id)
        s.initialize_with_molality = int(self.initialize_with_molality)
        s.use_prev_soln_as_guess = int(self.use_prev_soln_as_guess)
        s.max_iterations = int(self.max_iterations)
        s.max_dlnC = float(self.max_dlnC)
        s.free_ion_guess = dp(self.free_ion_guess)
        self._keep = keep
        return s


def h_ion_id(net: ReactionNetwork) -> Optional[int]:
    """species_idx%h_ion_id (reaction_database.F90:3590-3656): the primary
    named H+, else -(k+1) for the complex k named H+"""
    if net.primary_species_names and "H+" in net.primary_species_names:
        return net.primary_species_names.index("H+")
    if net.eqcplx_names and "H+" in net.eqcplx_names:
        return -(net.eqcplx_names.index("H+") + 1)
    return None


def from_deck(db: database.Database, net: ReactionNetwork,
              spec: Dict[str, Tuple[str, ...]], temperature: float = 25.0,
              isothermal: bool = False, **kw) -> Constraint:
    """Constraint from deck-style entries {primary name: (letter[, mineral
    or gas name])}, e.g. {"H+": ("P",), "HCO3-": ("G", "CO2(g)"),
    "Ca++": ("M", "Calcite")}.  Primaries not named are TOTAL."""
    names = net.primary_species_names
    types, rxn, rnames = [], [], []
    for name in names:
        entry = spec.get(name, ("T",))
        t = TYPE_OF[entry[0].upper()]
        types.append(t)
        if t in (abi.CONSTRAINT_MINERAL, abi.CONSTRAINT_GAS):
            if len(entry) < 2:
                raise ValueError("%s: an M or G constraint names its mineral or gas" % name)
            if entry[1] not in rnames:
                rnames.append(entry[1])
            rxn.append(rnames.index(entry[1]))
        else:
            rxn.append(-1)
    unknown = set(spec) - set(names)
    if unknown:
        raise KeyError("constraint names species that are not primaries: %s" % sorted(unknown))
    rows = database.constraint_reactions(db, names, net.eqcplx_names or [], rnames,
                                         temperature, isothermal) if rnames else []
    con = Constraint(types=types, rxn=rxn, h_ion_id=h_ion_id(net), rxn_names=rnames, **kw)
    if rows:
        con.rxn_species = [r[0] for r in rows]
        con.rxn_stoich = [r[1] for r in rows]
        h2o = np.array([r[2] for r in rows])
        con.rxn_h2o = h2o if np.any(h2o) else None
        con.rxn_logK = np.array([r[3] for r in rows])
        con.rxn_logKcoef = np.array([r[4] for r in rows]).reshape(-1)
    return con
