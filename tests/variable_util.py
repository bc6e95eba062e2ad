"""Oracle side of per-cell active primaries (options.variable_ncomp): a
cell with nact active primaries is the reference's RStep on the network
restricted to primaries 0..nact-1 and the complexes / kinetic minerals all
of whose species are among them.  The oracle is run on those sub-networks,
one group of cells per nact, and the results are put back in the full
state (entries outside the cell's sub-network keep their input values)."""
import numpy as np

import oracle_lib as O
from pfchem.network import CellState, ReactionNetwork


def sub_network(net, nact):
    """(sub-network, complex ids kept, mineral ids kept)"""
    kc = [k for k, sp in enumerate(net.eqcplx_species) if max(sp) < nact]
    km = [m for m, sp in enumerate(net.kinmnrl_species) if max(sp) < nact]

    def pick(a, idx):
        return None if a is None else np.asarray(a)[idx]

    sub = ReactionNetwork(
        primary_species_names=list(net.primary_species_names[:nact]),
        primary_spec_Z=net.primary_spec_Z[:nact], primary_spec_a0=net.primary_spec_a0[:nact],
        eqcplx_names=[net.eqcplx_names[k] for k in kc],
        eqcplx_species=[list(net.eqcplx_species[k]) for k in kc],
        eqcplx_stoich=[list(net.eqcplx_stoich[k]) for k in kc],
        eqcplxh2ostoich=pick(net.eqcplxh2ostoich, kc), eqcplx_logK=pick(net.eqcplx_logK, kc),
        eqcplx_logKcoef=pick(net.eqcplx_logKcoef, kc), eqcplx_Z=pick(net.eqcplx_Z, kc),
        eqcplx_a0=pick(net.eqcplx_a0, kc),
        kinmnrl_names=[net.kinmnrl_names[m] for m in km],
        kinmnrl_species=[list(net.kinmnrl_species[m]) for m in km],
        kinmnrl_stoich=[list(net.kinmnrl_stoich[m]) for m in km],
        kinmnrlh2ostoich=pick(net.kinmnrlh2ostoich, km), kinmnrl_logK=pick(net.kinmnrl_logK, km),
        kinmnrl_logKcoef=pick(net.kinmnrl_logKcoef, km),
        kinmnrl_rate_constant=pick(net.kinmnrl_rate_constant, km),
        kinmnrl_activation_energy=pick(net.kinmnrl_activation_energy, km),
        kinmnrl_molar_vol=pick(net.kinmnrl_molar_vol, km),
        h2o_aq_id=net.h2o_aq_id if net.h2o_aq_id < nact else -1,
        debyeA=net.debyeA, debyeB=net.debyeB, debyeBdot=net.debyeBdot)
    return sub, np.array(kc, dtype=np.int64), np.array(km, dtype=np.int64)


ROWS = {"pri_molal": "aq", "total": "aq", "pri_act_coef": "aq", "sec_molal": "c",
        "sec_act_coef": "c", "mnrl_volfrac": "m", "mnrl_area": "m", "mnrl_rate": "m"}


def run_oracle_variable(wl, state=None, nthreads=0):
    """(result state as a dict of arrays, dict nact -> (cells, records))"""
    st = state or wl.state
    n = st.ncell
    nact_all = np.clip(st.arrays["ncomp_active"][0, :n], 1, wl.net.naqcomp)
    out = {k: v[:, :n].copy() for k, v in st.arrays.items()}
    recs = {}
    opt = wl.opt.__class__(**{**wl.opt.__dict__, "variable_ncomp": False})
    for nact in np.unique(nact_all):
        cells = np.flatnonzero(nact_all == nact)
        sub, kc, km = sub_network(wl.net, int(nact))
        ss = CellState(sub, cells.size)
        for name, a in ss.arrays.items():
            if name not in st.arrays:
                continue
            kind = ROWS.get(name)
            src = st.arrays[name]
            if kind == "aq":
                a[:, :] = src[:nact][:, cells]
            elif kind == "c":
                a[:, :] = src[kc][:, cells]
            elif kind == "m":
                a[:, :] = src[km][:, cells]
            elif a.shape[0]:
                a[:, :] = src[: a.shape[0]][:, cells]
        rec = O.Records(cells.size)
        O.step_batch(sub, opt, ss, wl.dt, nthreads, rec)
        recs[int(nact)] = (cells, rec, sub, ss)
        for name, a in ss.arrays.items():
            if name not in out:
                continue
            kind = ROWS.get(name)
            if kind == "aq":
                out[name][:nact, cells] = a
            elif kind == "c":
                out[name][np.ix_(kc, cells)] = a
            elif kind == "m":
                out[name][np.ix_(km, cells)] = a
            elif a.shape[0]:
                out[name][: a.shape[0], cells] = a
    return out, recs


def rel_residuals_variable(wl, got, recs):
    """oracle FP64 relative residual of a solver's answer `got` (dict of
    full-state arrays) on each group's sub-network, for cells whose oracle
    run took one sub-step; returns (rr, the oracle's own final residual)"""
    rr, own = [], []
    opt = wl.opt.__class__(**{**wl.opt.__dict__, "variable_ncomp": False})
    for nact, (cells, rec, sub, ss) in recs.items():
        one = np.flatnonzero(ss.arrays["num_sub_steps"][0] == 1)
        if one.size == 0:
            continue
        g = ss.copy()
        _, kc, km = sub_network(wl.net, nact)
        g.arrays["pri_molal"][:, :] = got["pri_molal"][:nact][:, cells]
        g.arrays["total"][:, :] = got["total"][:nact][:, cells]
        if kc.size:
            g.arrays["sec_molal"][:, :] = got["sec_molal"][kc][:, cells]
        if km.size:
            g.arrays["mnrl_volfrac"][:, :] = got["mnrl_volfrac"][km][:, cells]
        rr.append(O.rel_residuals(sub, opt, g, rec, one))
        own.append(O.record_fields(rec, "final_rel_residual")[one])
    return np.concatenate(rr), np.concatenate(own)
