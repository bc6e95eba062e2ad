"""Reaction network, solver options and per-cell state containers.

`ReactionNetwork` keeps the field names of the reference's
`reaction_rt_type` (reaction_aux.F90:123-311) and its kinetic-mineral part
(`mineral_type`, reaction_mineral_aux.F90:82-140), 0-based.  It owns numpy
arrays and produces the `pfchem_network` C struct (include/pfchem.h).

`SolverOptions` carries the Newton / activity controls whose defaults are
set in ReactionCreate (reaction_aux.F90:392-393, 426-430, 533-537).

`CellState` is the structure-of-arrays per-cell state (`pfchem_state`):
the rt_auxvar / global_auxvar / material_auxvar fields RStep reads and
writes (reactive_transport_aux.F90:21-74), field-major `[k][cell]`.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import abi

# Debye-Hueckel defaults, reaction_aux.F90:486-488
DEBYE_A = 0.5114
DEBYE_B = 0.3288
DEBYE_BDOT = 0.0410


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def _pack_reactions(species: List[List[int]], stoich: List[List[float]]):
    """Compressed (id, stoich) rows like eqcplxspecid(0:max, n) /
    eqcplxstoich(max, n) (reaction_database.F90:1704-1772): entry 0 of each
    id row is the species count."""
    n = len(species)
    maxsp = max([len(s) for s in species] + [1])
    ids = np.zeros((n, maxsp + 1), dtype=np.int32)
    st = np.zeros((n, maxsp), dtype=np.float64)
    for k, (sp, nu) in enumerate(zip(species, stoich)):
        if len(sp) != len(nu):
            raise ValueError("species/stoichiometry length mismatch in reaction %d" % k)
        ids[k, 0] = len(sp)
        ids[k, 1:1 + len(sp)] = sp
        st[k, :len(nu)] = nu
    return ids, st, maxsp


@dataclass
class ReactionNetwork:
    """Aqueous speciation + kinetic minerals + immobile species network."""

    primary_species_names: List[str]
    primary_spec_Z: np.ndarray
    primary_spec_a0: np.ndarray
    # complexes
    eqcplx_names: List[str] = field(default_factory=list)
    eqcplx_species: List[List[int]] = field(default_factory=list)
    eqcplx_stoich: List[List[float]] = field(default_factory=list)
    eqcplxh2ostoich: Optional[np.ndarray] = None
    eqcplx_logK: Optional[np.ndarray] = None
    eqcplx_logKcoef: Optional[np.ndarray] = None
    eqcplx_Z: Optional[np.ndarray] = None
    eqcplx_a0: Optional[np.ndarray] = None
    # kinetic minerals
    kinmnrl_names: List[str] = field(default_factory=list)
    kinmnrl_species: List[List[int]] = field(default_factory=list)
    kinmnrl_stoich: List[List[float]] = field(default_factory=list)
    kinmnrlh2ostoich: Optional[np.ndarray] = None
    kinmnrl_logK: Optional[np.ndarray] = None
    kinmnrl_logKcoef: Optional[np.ndarray] = None
    kinmnrl_rate_constant: Optional[np.ndarray] = None
    kinmnrl_activation_energy: Optional[np.ndarray] = None
    kinmnrl_molar_vol: Optional[np.ndarray] = None
    kinmnrl_Temkin_const: Optional[np.ndarray] = None
    kinmnrl_min_scale_factor: Optional[np.ndarray] = None
    kinmnrl_affinity_power: Optional[np.ndarray] = None
    kinmnrl_affinity_threshold: Optional[np.ndarray] = None
    kinmnrl_rate_limiter: Optional[np.ndarray] = None
    kinmnrl_irreversible: Optional[np.ndarray] = None
    # immobile species and their first-order decay
    immobile_names: List[str] = field(default_factory=list)
    decayspecid: Optional[np.ndarray] = None
    decay_rate_constant: Optional[np.ndarray] = None
    h2o_aq_id: int = -1
    debyeA: float = DEBYE_A
    debyeB: float = DEBYE_B
    debyeBdot: float = DEBYE_BDOT
    somdec: Optional[object] = None  # pfchem.somdec.SomDecNetwork
    # rate prefactors per kinetic mineral (None or [] = the rate constant)
    kinmnrl_prefactors: Optional[List[List["Prefactor"]]] = None

    def __post_init__(self):
        self.primary_spec_Z = _f64(self.primary_spec_Z)
        self.primary_spec_a0 = _f64(self.primary_spec_a0)
        nc = len(self.eqcplx_species)
        nm = len(self.kinmnrl_species)
        if self.eqcplx_logK is None:
            self.eqcplx_logK = np.zeros(nc)
        if self.eqcplx_Z is None:
            self.eqcplx_Z = np.zeros(nc)
        if self.eqcplx_a0 is None:
            self.eqcplx_a0 = np.zeros(nc)
        for name in ("eqcplx_logK", "eqcplx_Z", "eqcplx_a0"):
            setattr(self, name, _f64(getattr(self, name)))
        if nm:
            if self.kinmnrl_logK is None or self.kinmnrl_rate_constant is None:
                raise ValueError("kinetic minerals need logK and rate constants")
            if self.kinmnrl_molar_vol is None:
                self.kinmnrl_molar_vol = np.zeros(nm)
        for name in ("kinmnrl_logK", "kinmnrl_rate_constant", "kinmnrl_molar_vol",
                     "kinmnrl_activation_energy", "kinmnrl_Temkin_const",
                     "kinmnrl_min_scale_factor", "kinmnrl_affinity_power",
                     "kinmnrl_affinity_threshold", "kinmnrl_rate_limiter",
                     "eqcplxh2ostoich", "kinmnrlh2ostoich", "eqcplx_logKcoef",
                     "kinmnrl_logKcoef", "decay_rate_constant"):
            v = getattr(self, name)
            if v is not None:
                setattr(self, name, _f64(v))
        for name in ("kinmnrl_irreversible", "decayspecid"):
            v = getattr(self, name)
            if v is not None:
                setattr(self, name, _i32(v))
        self._abi_keep = None

    # sizes ---------------------------------------------------------------
    @property
    def naqcomp(self) -> int:
        return len(self.primary_spec_Z)

    @property
    def nimmobile(self) -> int:
        return len(self.immobile_names)

    @property
    def ncomp(self) -> int:
        return self.naqcomp + self.nimmobile

    @property
    def neqcplx(self) -> int:
        return len(self.eqcplx_species)

    @property
    def nkinmnrl(self) -> int:
        return len(self.kinmnrl_species)

    @property
    def ndecay_rxn(self) -> int:
        return 0 if self.decayspecid is None else len(self.decayspecid)

    def validate(self):
        n = self.naqcomp
        for k, sp in enumerate(self.eqcplx_species):
            if any(s < 0 or s >= n for s in sp):
                raise ValueError("complex %d names a species outside 0..%d" % (k, n - 1))
        for m, sp in enumerate(self.kinmnrl_species):
            if any(s < 0 or s >= n for s in sp):
                raise ValueError("mineral %d names a species outside 0..%d" % (m, n - 1))
        if self.primary_spec_a0.shape != (n,):
            raise ValueError("primary_spec_a0 has the wrong length")

    def dense_stoichiometry(self) -> np.ndarray:
        """S[k, j] = nu of primary j in complex k (dense)."""
        S = np.zeros((self.neqcplx, self.naqcomp))
        for k, (sp, nu) in enumerate(zip(self.eqcplx_species, self.eqcplx_stoich)):
            for j, v in zip(sp, nu):
                S[k, j] += v
        return S

    def dense_mineral_stoichiometry(self) -> np.ndarray:
        S = np.zeros((self.nkinmnrl, self.naqcomp))
        for m, (sp, nu) in enumerate(zip(self.kinmnrl_species, self.kinmnrl_stoich)):
            for j, v in zip(sp, nu):
                S[m, j] += v
        return S

    # ABI -------------------------------------------------------------------
    def to_abi(self) -> abi.Network:
        """Build (and cache) the pfchem_network struct; the returned struct
        references arrays owned by this object."""
        self.validate()
        keep = []

        def dp(a):
            if a is None:
                return C.cast(None, C.POINTER(C.c_double))
            a = _f64(a)
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(C.c_double))

        def ip(a):
            if a is None:
                return C.cast(None, C.POINTER(C.c_int32))
            a = _i32(a)
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(C.c_int32))

        s = abi.Network()
        s.naqcomp = self.naqcomp
        s.nimmobile = self.nimmobile
        s.h2o_aq_id = self.h2o_aq_id
        s.primary_spec_Z = dp(self.primary_spec_Z)
        s.primary_spec_a0 = dp(self.primary_spec_a0)
        if self.neqcplx:
            ids, st, ms = _pack_reactions(self.eqcplx_species, self.eqcplx_stoich)
            s.neqcplx = self.neqcplx
            s.eqcplx_maxsp = ms
            s.eqcplxspecid = ip(ids)
            s.eqcplxstoich = dp(st)
        s.eqcplxh2ostoich = dp(self.eqcplxh2ostoich)
        s.eqcplx_logK = dp(self.eqcplx_logK)
        s.eqcplx_logKcoef = dp(self.eqcplx_logKcoef)
        s.eqcplx_Z = dp(self.eqcplx_Z)
        s.eqcplx_a0 = dp(self.eqcplx_a0)
        if self.nkinmnrl:
            ids, st, ms = _pack_reactions(self.kinmnrl_species, self.kinmnrl_stoich)
            s.nkinmnrl = self.nkinmnrl
            s.kinmnrl_maxsp = ms
            s.kinmnrlspecid = ip(ids)
            s.kinmnrlstoich = dp(st)
        s.kinmnrlh2ostoich = dp(self.kinmnrlh2ostoich)
        s.kinmnrl_logK = dp(self.kinmnrl_logK)
        s.kinmnrl_logKcoef = dp(self.kinmnrl_logKcoef)
        s.kinmnrl_rate_constant = dp(self.kinmnrl_rate_constant)
        s.kinmnrl_activation_energy = dp(self.kinmnrl_activation_energy)
        s.kinmnrl_molar_vol = dp(self.kinmnrl_molar_vol)
        s.kinmnrl_Temkin_const = dp(self.kinmnrl_Temkin_const)
        s.kinmnrl_min_scale_factor = dp(self.kinmnrl_min_scale_factor)
        s.kinmnrl_affinity_power = dp(self.kinmnrl_affinity_power)
        s.kinmnrl_affinity_threshold = dp(self.kinmnrl_affinity_threshold)
        s.kinmnrl_rate_limiter = dp(self.kinmnrl_rate_limiter)
        s.kinmnrl_irreversible = ip(self.kinmnrl_irreversible)
        s.ndecay_rxn = self.ndecay_rxn
        s.decayspecid = ip(self.decayspecid)
        s.decay_rate_constant = dp(self.decay_rate_constant)
        s.debyeA = self.debyeA
        s.debyeB = self.debyeB
        s.debyeBdot = self.debyeBdot
        pref = self.kinmnrl_prefactors
        if pref and any(pref):
            nm = self.nkinmnrl
            mp = max(len(p) for p in pref)
            msp = max(max((len(q.species) for q in p), default=0) for p in pref)
            msp = max(msp, 1)
            npf = np.zeros(nm, np.int32)
            pid = np.zeros((nm, mp, msp + 1), np.int32)
            alpha = np.zeros((nm, mp, msp))
            beta = np.zeros((nm, mp, msp))
            atten = np.zeros((nm, mp, msp))
            prate = np.zeros((nm, mp))
            pea = np.zeros((nm, mp))
            for m, plist in enumerate(pref):
                npf[m] = len(plist)
                for q, pf in enumerate(plist):
                    prate[m, q] = pf.rate
                    pea[m, q] = pf.activation_energy
                    pid[m, q, 0] = len(pf.species)
                    for a_, (sid, al, be, at) in enumerate(pf.species):
                        pid[m, q, 1 + a_] = sid
                        alpha[m, q, a_] = al
                        beta[m, q, a_] = be
                        atten[m, q, a_] = at
            s.kinmnrl_max_pref = mp
            s.kinmnrl_max_pref_spec = msp
            s.kinmnrl_num_prefactors = ip(npf)
            s.kinmnrl_prefactor_id = ip(pid)
            s.kinmnrl_pref_alpha = dp(alpha)
            s.kinmnrl_pref_beta = dp(beta)
            s.kinmnrl_pref_atten_coef = dp(atten)
            s.kinmnrl_pref_rate = dp(prate)
            s.kinmnrl_pref_activation_energy = dp(pea)
        if self.somdec is not None:
            sd = self.somdec.to_abi(keep)
            keep.append(sd)
            s.somdec = C.pointer(sd)
        else:
            s.somdec = C.cast(None, C.POINTER(abi.Somdec))
        keep.append(s)
        self._abi_keep = keep
        return s


@dataclass
class Prefactor:
    """one parallel rate term of a kinetic mineral (a PREFACTOR block of
    MINERAL_KINETICS, reaction_mineral.F90:820-870): rate constant (mol/m^2/s),
    activation energy (J/mol, 0 = none) and species terms (primary id, ALPHA,
    BETA, ATTENUATION_COEF): a^alpha / (1 + atten a^beta)"""
    rate: float
    species: List[tuple] = field(default_factory=list)
    activation_energy: float = 0.0


@dataclass
class SolverOptions:
    """reaction_rt_type solver controls; defaults of ReactionCreate
    (reaction_aux.F90:392-393, 426-430, 533-537)."""

    use_full_geochemistry: bool = True
    use_log_formulation: bool = True
    act_coef_update_frequency: int = abi.ACT_COEF_FREQUENCY_OFF
    act_coef_update_algorithm: int = abi.ACT_COEF_ALGORITHM_LAG
    use_activity_h2o: bool = False
    use_total_as_guess: bool = False
    isothermal: bool = True
    maximum_reaction_iterations: int = 20
    maximum_reaction_cuts: int = 10
    max_dlnC_rreact: float = 5.0
    max_residual_tolerance: float = 1.0e-12
    max_rel_residual_tolerance: float = 1.0e-8
    max_relative_change_tolerance: float = 1.0e-6
    # per-cell active primary counts in the state (pfchem_state.ncomp_active)
    variable_ncomp: bool = False
    # RReact calls per RStep before the cell fails (0: 10000); the
    # reference's sub-step loop has no bound (reaction.F90:3721-3725)
    maximum_rreact_calls: int = 0

    def to_abi(self) -> abi.Options:
        o = abi.Options()
        for name, _ in abi.Options._fields_:
            if name.startswith("pad"):
                continue
            v = getattr(self, name)
            setattr(o, name, int(v) if isinstance(v, (bool, int)) else float(v))
        return o


def state_fields(net: ReactionNetwork, somdec_scalars: bool = False,
                 variable_ncomp: bool = False) -> Dict[str, tuple]:
    """name -> (rows, dtype) of every SoA array of pfchem_state."""
    f = {
        "pri_molal": (net.naqcomp, "f8"),
        "total": (net.naqcomp, "f8"),
        "immobile": (net.nimmobile, "f8"),
        "sec_molal": (net.neqcplx, "f8"),
        "pri_act_coef": (net.naqcomp, "f8"),
        "sec_act_coef": (net.neqcplx, "f8"),
        "ln_act_h2o": (1, "f8"),
        "mnrl_volfrac": (net.nkinmnrl, "f8"),
        "mnrl_area": (net.nkinmnrl, "f8"),
        "mnrl_rate": (net.nkinmnrl, "f8"),
        "porosity": (1, "f8"),
        "saturation": (1, "f8"),
        "den_kg": (1, "f8"),
        "volume": (1, "f8"),
        "temperature": (1, "f8"),
        "num_iterations": (1, "i4"),
        "num_sub_steps": (1, "i4"),
        "status": (1, "i4"),
    }
    if somdec_scalars:
        f["somdec_f_t"] = (1, "f8")
        f["somdec_f_w"] = (1, "f8")
    if variable_ncomp:
        f["ncomp_active"] = (1, "i4")
    return f


class CellState:
    """Host (numpy) SoA state: arrays[name] has shape (rows, ld)."""

    def __init__(self, net: ReactionNetwork, ncell: int, ld: Optional[int] = None,
                 somdec_scalars: bool = False, variable_ncomp: bool = False):
        self.net = net
        self.ncell = int(ncell)
        self.ld = int(ld if ld is not None else ncell)
        if self.ld < self.ncell:
            raise ValueError("ld < ncell")
        self.fields = state_fields(net, somdec_scalars, variable_ncomp)
        self.arrays: Dict[str, np.ndarray] = {}
        for name, (rows, dt) in self.fields.items():
            self.arrays[name] = np.zeros((rows, self.ld), dtype=dt)

    def __getitem__(self, name):
        return self.arrays[name]

    def copy(self) -> "CellState":
        out = CellState.__new__(CellState)
        out.net, out.ncell, out.ld, out.fields = self.net, self.ncell, self.ld, self.fields
        out.arrays = {k: v.copy() for k, v in self.arrays.items()}
        return out

    def subset(self, idx) -> "CellState":
        idx = np.asarray(idx)
        out = CellState.__new__(CellState)
        out.net, out.fields = self.net, self.fields
        out.ncell = out.ld = int(idx.size)
        out.arrays = {k: np.ascontiguousarray(v[:, idx]) for k, v in self.arrays.items()}
        return out

    def to_abi(self) -> abi.State:
        s = abi.State()
        s.ld = self.ld
        for name, _ in abi.State._fields_:
            if name == "ld":
                continue
            a = self.arrays.get(name)
            if a is None or a.shape[0] == 0:
                setattr(s, name, None)
            else:
                assert a.flags.c_contiguous
                setattr(s, name, a.ctypes.data)
        return s
