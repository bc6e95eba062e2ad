"""SomDec decomposition sandbox descriptor (the `pfchem_somdec` C struct) and a
builder that does what SomDecSetup does (reaction_sandbox_somdec.F90:
987-1501) from a deck-like description.

Pools: a pool with a C:N ratio (mass C / mass N) has one species (its C;
N = C x nc), a pool without one has two, name+'C' and name+'N' (variable
C:N).  Reactions name an upstream pool, downstream pools with their
stoichiometry, and one of TURNOVER_TIME (rate = 1/turnover, :796),
RATE_CONSTANT or RATE_DECOMPOSITION.  Setup-time stoichiometry of fixed-C:N
reactions (:1358-1398): respiration fraction 1 - sum(downstream), N
mineralization upstream_nc - sum(stoich * downstream_nc).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import abi

# elm_rspfuncs.F90 constants
CN_RATIO_MASS_TO_MOL = 1.166156023644992
TEMPERATURE_OFF, TEMPERATURE_CLMCN, TEMPERATURE_Q10, TEMPERATURE_DLEM, TEMPERATURE_ARRHENIUS = \
    0, 1, 2, 3, 4
MOISTURE_OFF, MOISTURE_CLMCN, MOISTURE_DLEM, MOISTURE_LOGTHETA = 0, 1, 2, 3
OX_OFF, OX_MONOD, OX_WFPS = 0, 1, 2
INHIBITION_THRESHOLD, INHIBITION_MONOD, INHIBITION_INVERSE_MONOD = 1, 3, 4
UNINITIALIZED_DOUBLE = -999.0  # pflotran_constants.F90:327

MD, MM, MI = abi.SOMDEC_MAX_DOWN, abi.SOMDEC_MAX_MONOD, abi.SOMDEC_MAX_INHIB


@dataclass
class Pool:
    name: str
    cn_ratio: Optional[float] = None  # mass C / mass N; None = variable C:N


@dataclass
class Monod:
    species: str                       # aqueous primary or immobile species name
    half_saturation: float             # HALF_SATURATION_CONSTANT
    threshold: float = 0.0             # THRESHOLD_CONCENTRATION
    pool_normalized: bool = False


@dataclass
class Inhibition:
    species: str
    itype: int = INHIBITION_MONOD
    constant: float = 0.0
    constant2: Optional[float] = None


@dataclass
class Reaction:
    upstream: str
    downstream: List[Tuple[str, float]] = field(default_factory=list)
    turnover_time: Optional[float] = None      # s
    rate_constant: Optional[float] = None      # 1/s
    rate_decomposition: Optional[float] = None  # 1/s
    rate_ad_factor: float = 1.0
    monod: List[Monod] = field(default_factory=list)
    inhibition: List[Inhibition] = field(default_factory=list)
    temperature_response_function: int = TEMPERATURE_OFF
    Q10: float = 1.5
    Ea: float = 0.0
    moisture_response_function: int = MOISTURE_OFF
    Ox_response_function: int = OX_OFF
    Ox_half_saturation: float = 0.0
    Ox_species: Optional[str] = None  # oxidizer of the Monod oxygen response


@dataclass
class SomDecNetwork:
    """the compiled descriptor (per-reaction arrays, 0-based species ids)"""
    nrxn: int
    species_id_co2: int
    species_id_nh4: int
    species_id_no3: int = -1
    species_id_n2o: int = -1
    species_id_proton: int = -1
    species_id_hr: int = -1
    species_id_nmin: int = -1
    species_id_nimm: int = -1
    species_id_nimp: int = -1
    species_id_ngasmin: int = -1
    x0eps: float = 1.0e-20
    n2o_frac_mineralization: float = 0.02
    inhibition_nh4_no3: float = 1.0
    species_o2: int = 0  # tracked O2: aqueous id + 1, -(immobile id + 1), 0 = none
    arrays: dict = field(default_factory=dict)

    def to_abi(self, keep: list) -> abi.Somdec:
        s = abi.Somdec()
        for name in ("nrxn", "species_id_co2", "species_id_nh4", "species_id_no3",
                     "species_id_n2o", "species_id_proton", "species_id_hr", "species_id_nmin",
                     "species_id_nimm", "species_id_nimp", "species_id_ngasmin", "x0eps",
                     "n2o_frac_mineralization", "inhibition_nh4_no3", "species_o2"):
            setattr(s, name, getattr(self, name))
        for fname, ftype in abi.Somdec._fields_:
            if fname not in self.arrays:
                continue
            a = self.arrays[fname]
            if ftype == C.POINTER(C.c_double):
                a = np.ascontiguousarray(a, dtype=np.float64)
                setattr(s, fname, a.ctypes.data_as(C.POINTER(C.c_double)))
            else:
                a = np.ascontiguousarray(a, dtype=np.int32)
                setattr(s, fname, a.ctypes.data_as(C.POINTER(C.c_int32)))
            keep.append(a)
        return s


def build_somdec(pools: Sequence[Pool], reactions: Sequence[Reaction],
                 primary_names: Sequence[str], immobile_names: Sequence[str],
                 x0eps: float = 1.0e-20, n2o_frac_mineralization: float = 0.02,
                 inhibition_nh4_no3: float = 1.0,
                 o2_species: Optional[str] = None) -> SomDecNetwork:
    """SomDecSetup: species ids, setup-time N:C ratios and stoichiometry;
    o2_species: O2_SPECIES_NAME (its consumption is tracked)"""
    pid = {n: i for i, n in enumerate(primary_names)}
    iid = {n: i for i, n in enumerate(immobile_names)}

    def imm_or(name):
        return iid.get(name, -1)

    pool_c, pool_n, pool_nc = {}, {}, {}
    for p in pools:
        if p.cn_ratio is None or p.cn_ratio <= 0.0:
            if p.name + "C" not in iid or p.name + "N" not in iid:
                raise ValueError("variable C:N pool %s needs immobile species %sC and %sN"
                                 % (p.name, p.name, p.name))
            pool_c[p.name], pool_n[p.name] = iid[p.name + "C"], iid[p.name + "N"]
            pool_nc[p.name] = UNINITIALIZED_DOUBLE
        else:
            if p.name not in iid:
                raise ValueError("pool %s is not an immobile species" % p.name)
            pool_c[p.name], pool_n[p.name] = iid[p.name], -1
            pool_nc[p.name] = 1.0 / p.cn_ratio / CN_RATIO_MASS_TO_MOL
    nr = len(reactions)
    A = {}

    def arr(name, shape, dtype, fill):
        A[name] = np.full(shape, fill, dtype=dtype)
        return A[name]

    rc = arr("rate_constant", nr, np.float64, -1.0)
    rd = arr("rate_decomposition", nr, np.float64, -1.0)
    ad = arr("rate_ad_factor", nr, np.float64, 1.0)
    uc = arr("upstream_c_id", nr, np.int32, -1)
    un = arr("upstream_n_id", nr, np.int32, -1)
    uhr = arr("upstream_hr_id", nr, np.int32, -1)
    unmin = arr("upstream_nmin_id", nr, np.int32, -1)
    unimp = arr("upstream_nimp_id", nr, np.int32, -1)
    unimm = arr("upstream_nimm_id", nr, np.int32, -1)
    unc = arr("upstream_nc", nr, np.float64, UNINITIALIZED_DOUBLE)
    uvary = arr("upstream_is_varycn", nr, np.int32, 0)
    nd = arr("n_downstream_pools", nr, np.int32, 0)
    dc = arr("downstream_c_id", (nr, MD), np.int32, -1)
    dn = arr("downstream_n_id", (nr, MD), np.int32, -1)
    ds = arr("downstream_stoich", (nr, MD), np.float64, 0.0)
    dnc = arr("downstream_nc", (nr, MD), np.float64, UNINITIALIZED_DOUBLE)
    dvary = arr("downstream_is_varycn", (nr, MD), np.int32, 0)
    mcs = arr("mineral_c_stoich", nr, np.float64, 0.0)
    mns = arr("mineral_n_stoich", nr, np.float64, 0.0)
    tf = arr("temperature_response_function", nr, np.int32, 0)
    q10 = arr("Q10", nr, np.float64, 0.0)
    ea = arr("Ea", nr, np.float64, 0.0)
    mf = arr("moisture_response_function", nr, np.int32, 0)
    of = arr("Ox_response_function", nr, np.int32, 0)
    oh = arr("Ox_half_saturation", nr, np.float64, 0.0)
    nmo = arr("n_monod", nr, np.int32, 0)
    msp = arr("monod_specid", (nr, MM), np.int32, 0)
    mhs = arr("monod_half_saturation", (nr, MM), np.float64, 0.0)
    mth = arr("monod_threshold", (nr, MM), np.float64, 0.0)
    mpn = arr("monod_pool_normalized", (nr, MM), np.int32, 0)
    nin = arr("n_inhibition", nr, np.int32, 0)
    isp = arr("inhibition_specid", (nr, MI), np.int32, 0)
    ity = arr("inhibition_type", (nr, MI), np.int32, 0)
    ic1 = arr("inhibition_constant", (nr, MI), np.float64, 0.0)
    ic2 = arr("inhibition_constant2", (nr, MI), np.float64, 0.0)
    oxs = arr("Ox_specid", nr, np.int32, 0)

    def spec(name):
        if name in pid:
            return pid[name]
        if name in iid:
            return -iid[name] - 1
        raise ValueError("species %s is neither primary nor immobile" % name)

    def code(name):  # 1-based species code: aqueous id + 1, -(immobile id + 1)
        s = spec(name)
        return s + 1 if s >= 0 else s

    for r, rx in enumerate(reactions):
        if rx.turnover_time is not None:
            rc[r] = 1.0 / rx.turnover_time
        elif rx.rate_constant is not None:
            rc[r] = rx.rate_constant
        elif rx.rate_decomposition is not None:
            rd[r] = rx.rate_decomposition
        else:
            raise ValueError("reaction %d needs TURNOVER_TIME, RATE_CONSTANT or "
                             "RATE_DECOMPOSITION" % r)
        ad[r] = rx.rate_ad_factor
        uc[r], un[r], unc[r] = pool_c[rx.upstream], pool_n[rx.upstream], pool_nc[rx.upstream]
        uvary[r] = 1 if un[r] >= 0 else 0
        uhr[r] = imm_or(rx.upstream + "CHR")
        unmin[r] = imm_or(rx.upstream + "NMIN")
        unimp[r] = imm_or(rx.upstream + "NIMP")
        unimm[r] = imm_or(rx.upstream + "NIMM")
        if len(rx.downstream) > MD:
            raise ValueError("at most %d downstream pools" % MD)
        nd[r] = len(rx.downstream)
        for j, (pname, st) in enumerate(rx.downstream):
            dc[r, j], dn[r, j], dnc[r, j] = pool_c[pname], pool_n[pname], pool_nc[pname]
            ds[r, j] = st
            dvary[r, j] = 1 if dn[r, j] >= 0 else 0
        tf[r], q10[r], ea[r] = rx.temperature_response_function, rx.Q10, rx.Ea
        mf[r], of[r], oh[r] = rx.moisture_response_function, rx.Ox_response_function, \
            rx.Ox_half_saturation
        if rx.Ox_species is not None:
            oxs[r] = code(rx.Ox_species)
        if len(rx.monod) > MM or len(rx.inhibition) > MI:
            raise ValueError("too many Monod / inhibition terms")
        nmo[r] = len(rx.monod)
        for q, m in enumerate(rx.monod):
            msp[r, q], mhs[r, q], mth[r, q], mpn[r, q] = spec(m.species), m.half_saturation, \
                m.threshold, int(m.pool_normalized)
        nin[r] = len(rx.inhibition)
        for q, h in enumerate(rx.inhibition):
            isp[r, q], ity[r, q], ic1[r, q] = spec(h.species), h.itype, h.constant
            if h.constant2 is not None:
                ic2[r, q] = h.constant2
                ity[r, q] = INHIBITION_THRESHOLD if h.itype == INHIBITION_THRESHOLD else h.itype
            elif h.itype == INHIBITION_THRESHOLD:
                raise ValueError("threshold inhibition needs constant2")
        # setup-time stoichiometry of fixed C:N reactions (:1358-1398)
        if not uvary[r]:
            st_c = 1.0 - ds[r, :nd[r]].sum()
            st_n = unc[r]
            for j in range(nd[r]):
                st_n = st_n - ds[r, j] * dnc[r, j]
            if abs(st_c) < 1e-15:
                st_c = 0.0
            if abs(st_n) < 1e-15:
                st_n = 0.0
            if st_c < 0.0 or st_n < 0.0:
                raise ValueError("fixed C:N reaction %d has a negative respiration or N "
                                 "mineralization fraction" % r)
            mcs[r], mns[r] = st_c, st_n

    def pid_or(name):
        return pid.get(name, -1)

    return SomDecNetwork(
        nrxn=nr, species_id_co2=pid["CO2(aq)"], species_id_nh4=pid["NH4+"],
        species_id_no3=pid_or("NO3-"), species_id_n2o=pid_or("N2O(aq)"),
        species_id_proton=pid_or("H+"), species_id_hr=imm_or("HRimm"),
        species_id_nmin=imm_or("Nmin"), species_id_nimm=imm_or("Nimm"),
        species_id_nimp=imm_or("Nimp"), species_id_ngasmin=imm_or("NGASmin"),
        x0eps=x0eps, n2o_frac_mineralization=n2o_frac_mineralization,
        inhibition_nh4_no3=inhibition_nh4_no3,
        species_o2=code(o2_species) if o2_species is not None else 0, arrays=A)
