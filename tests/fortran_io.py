"""Text input / output of the Fortran host-module driver
(pflotran-elm-interface_amd/fortran/pfchem_rstep_driver.F90): a network in
reaction_rt_type form (1-based species ids, eqcplxspecid(0:maxsp, k) with
the count in entry 0) and per-cell auxvars, and the driver's per-cell
results."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(os.path.dirname(HERE), "pflotran-elm-interface_amd", "fortran",
                      "pfchem_rstep_driver")


def _fmt(vals):
    return " ".join(repr(float(v)) if not isinstance(v, (int, np.integer)) else str(int(v))
                    for v in vals)


def _rows(species, stoich, width):
    ids, st = [], []
    for sp, nu in zip(species, stoich):
        ids.append([len(sp)] + [s + 1 for s in sp] + [0] * (width - len(sp)))
        st.append(list(nu) + [0.0] * (width - len(nu)))
    return ids, st


def write_input(path, net, opt, state, dt, variant):
    """network + options + the cells of `state` (a CellState) in the
    driver's list-directed format"""
    for name in ("kinmnrl_activation_energy", "kinmnrl_Temkin_const", "kinmnrl_min_scale_factor",
                 "kinmnrl_affinity_power", "kinmnrl_affinity_threshold", "kinmnrl_rate_limiter",
                 "kinmnrl_irreversible", "eqcplx_logKcoef", "kinmnrl_logKcoef"):
        if getattr(net, name, None) is not None:
            raise ValueError("the driver's input format has no %s" % name)
    if net.somdec is not None:
        raise ValueError("the driver's input format has no SomDec descriptor")
    naq, nimm = net.naqcomp, net.nimmobile
    nc, nm = len(net.eqcplx_species), len(net.kinmnrl_species)
    cms = max([len(s) for s in net.eqcplx_species] + [1])
    mms = max([len(s) for s in net.kinmnrl_species] + [1])
    ndecay = 0 if net.decayspecid is None else len(net.decayspecid)
    L = []
    L.append(_fmt([naq, nimm, nc, cms, nm, mms, ndecay, net.h2o_aq_id + 1]))
    L.append(_fmt([int(opt.use_full_geochemistry), int(opt.use_log_formulation),
                   int(opt.use_total_as_guess), int(opt.use_activity_h2o),
                   opt.act_coef_update_frequency, opt.act_coef_update_algorithm,
                   opt.maximum_reaction_iterations, opt.maximum_reaction_cuts,
                   int(opt.isothermal)]))
    L.append(_fmt([opt.max_dlnC_rreact, opt.max_residual_tolerance, opt.max_rel_residual_tolerance,
                   opt.max_relative_change_tolerance, net.debyeA, net.debyeB, net.debyeBdot]))
    L.append(_fmt(net.primary_spec_Z))
    L.append(_fmt(net.primary_spec_a0))
    h2o = net.eqcplxh2ostoich if net.eqcplxh2ostoich is not None else np.zeros(nc)
    ids, st = _rows(net.eqcplx_species, net.eqcplx_stoich, cms)
    for k in range(nc):
        L.append(_fmt(ids[k]))
        L.append(_fmt(st[k]))
        L.append(_fmt([1 if h2o[k] != 0 else 0, h2o[k], net.eqcplx_logK[k], net.eqcplx_Z[k],
                       net.eqcplx_a0[k]]))
    mh2o = net.kinmnrlh2ostoich if net.kinmnrlh2ostoich is not None else np.zeros(nm)
    ids, st = _rows(net.kinmnrl_species, net.kinmnrl_stoich, mms)
    for m in range(nm):
        L.append(_fmt(ids[m]))
        L.append(_fmt(st[m]))
        L.append(_fmt([1 if mh2o[m] != 0 else 0, mh2o[m], net.kinmnrl_logK[m],
                       net.kinmnrl_rate_constant[m], net.kinmnrl_molar_vol[m]]))
    if ndecay:
        L.append(_fmt(np.asarray(net.decayspecid) + 1))
        L.append(_fmt(net.decay_rate_constant))
    n = state.ncell
    L.append(_fmt([n, float(dt), variant]))
    a = state.arrays
    law = a["ln_act_h2o"][0, :n] if "ln_act_h2o" in a and a["ln_act_h2o"].shape[0] else np.zeros(n)
    temp = a["temperature"][0, :n] if a.get("temperature") is not None and a["temperature"].shape[0] \
        else np.full(n, 25.0)
    blocks = [a["pri_molal"][:, :n], a["total"][:, :n], a["pri_act_coef"][:, :n],
              a["immobile"][:, :n], a["sec_molal"][:, :n], a["sec_act_coef"][:, :n],
              a["mnrl_volfrac"][:, :n], a["mnrl_area"][:, :n], law[None, :],
              a["porosity"][:, :n], a["saturation"][:, :n], a["den_kg"][:, :n],
              a["volume"][:, :n], temp[None, :]]
    cols = np.vstack([b for b in blocks if b.shape[0]])
    for c in range(n):
        L.append(_fmt(cols[:, c]))
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def read_output(path, net, ncell):
    """(first-cell RStep result, dict of per-cell arrays [rows, ncell])"""
    naq, nimm = net.naqcomp, net.nimmobile
    nc, nm = len(net.eqcplx_species), len(net.kinmnrl_species)
    with open(path) as f:
        lines = f.read().split("\n")
    head = lines[0].split()
    assert head[0] == "rstep"
    rstep = dict(num_sub_steps=int(head[1]), num_iterations=int(head[2]),
                 pri_molal=np.array([float(v) for v in head[3:3 + naq]]))
    rows = np.array([[float(v) for v in ln.split()] for ln in lines[1:1 + ncell]])
    out = {"status": rows[:, 0].astype(int)[None, :], "num_iterations": rows[:, 1].astype(int)[None, :],
           "num_sub_steps": rows[:, 2].astype(int)[None, :]}
    o = 3
    for name, k in (("pri_molal", naq), ("total", naq), ("immobile", nimm), ("sec_molal", nc),
                    ("mnrl_volfrac", nm), ("mnrl_rate", nm)):
        out[name] = rows[:, o:o + k].T.copy()
        o += k
    return rstep, out
