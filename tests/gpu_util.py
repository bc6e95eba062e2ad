"""Helpers shared by the GPU parity tests: run a workload through the oracle
(CPU, the checker) and through the HIP kernels, and compare."""
import numpy as np

import oracle_lib as O
from pfchem import abi
from pfchem.lib import Context, DeviceState

FIELDS_F64 = ("pri_molal", "total", "sec_molal", "immobile", "mnrl_volfrac", "mnrl_rate")


def run_oracle(wl, nthreads=0, state=None):
    st = (state or wl.state).copy()
    rec = O.Records(st.ncell)
    nfail = O.step_batch(wl.net, wl.opt, st, wl.dt, nthreads, rec)
    return st, rec, nfail


def run_gpu(wl, variant, lanes=None, block=None, bpc=None, state=None, ctx=None, repeat=1,
            rows=None):
    import torch
    ctx = ctx or Context(wl.net, wl.opt)
    if lanes is not None:
        cur = ctx.get_launch(variant)
        ctx.set_launch(variant, lanes, block or cur["block_size"], bpc or cur["blocks_per_cu"],
                       rows or 1)
    host = state or wl.state
    out = None
    for _ in range(repeat):
        ds = DeviceState(host)
        ctx.step(ds, wl.dt, variant)
        torch.cuda.synchronize()
        out = ds.to_host()
    return out, ctx


def rel_err(a, b, floor=0.0):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.abs(a - b)
    s = np.maximum(np.abs(b), floor)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(d == 0, 0.0, d / s)
    return r


def rate_scale(net, state, ncell):
    """per kinetic mineral: k A, the TST rate at unit affinity (largest area
    over the cells), mol/m^3/s"""
    k = np.asarray(net.kinmnrl_rate_constant, dtype=np.float64)
    a = state["mnrl_area"][:, :ncell].max(axis=1) if ncell else np.zeros(k.size)
    return (k * a)[:, None]


def total_term_scale(net, state, ncell):
    """per component and cell: the magnitude of the terms its total is summed
    from, (|m_i| + sum_k |nu_ik| m_k) rho / 1000 (RTotal, reaction.F90:4958-
    5066).  A total is a signed sum; where it cancels, its relative error is
    the terms' error times |terms| / |total|, so its error is measured
    against this scale."""
    naq = net.naqcomp
    m = np.abs(state["pri_molal"][:naq, :ncell])
    scale = m.copy()
    if len(net.eqcplx_species):
        nu = np.zeros((naq, len(net.eqcplx_species)))
        for k, (sp, st) in enumerate(zip(net.eqcplx_species, net.eqcplx_stoich)):
            for s_, v in zip(sp, st):
                nu[s_, k] += abs(v)
        scale += nu @ np.abs(state["sec_molal"][:len(net.eqcplx_species), :ncell])
    return scale * state["den_kg"][0, :ncell][None, :] * 1.0e-3


def compare(ref, got, ncell, fields=FIELDS_F64, floor_frac=1e-12, net=None):
    """max relative error per field over the first ncell cells.  Totals can be
    near zero, so their relative error uses a floor of floor_frac times the
    largest magnitude of that species over the cells; with the network given,
    a total's error is measured against the magnitude of its terms
    (total_term_scale).  Mineral rates
    Im = -A k (1 - QK) vanish at equilibrium, where their relative error is
    the concentration error amplified by QK / |1 - QK|: with the network
    given, they are compared relative to max(|Im|, k A) (k A: the rate at
    unit affinity, |1 - QK| = 1), else with the totals' floor."""
    out = {}
    for f in fields:
        a = ref[f][:, :ncell]
        b = got[f][:, :ncell]
        if a.size == 0:
            continue
        floor = 0.0
        if f == "total" or (f == "mnrl_rate" and net is None):
            floor = floor_frac * np.abs(a).max(axis=1, keepdims=True)
            if f == "total" and net is not None and not net.nimmobile and net.somdec is None:
                floor = np.maximum(floor, total_term_scale(net, ref, ncell))
        elif f == "mnrl_rate":
            floor = rate_scale(net, ref, ncell)
        out[f] = float(rel_err(b, a, floor).max())
    return out


def gate_rel_residual(wl, gpu_state, rec, cells, ref_state):
    """||F(x_gpu)||/||F(x0)|| of the oracle's last sub-step system, for cells
    whose oracle run used exactly one sub-step (same system for both), and
    the bar for each cell: 1e-10, or twice the oracle's own final relative
    residual where FP64 rounding keeps the oracle above 1e-10 (cells whose
    initial residual is already near rounding level, so that the residual of
    any answer is rounding noise of that size)."""
    ns = ref_state["num_sub_steps"][0]
    cells = np.array([c for c in cells if ns[c] == 1], dtype=np.int64)
    if cells.size == 0:
        return np.zeros(0), np.zeros(0)
    rr = O.rel_residuals(wl.net, wl.opt, gpu_state, rec, cells)
    own = O.record_fields(rec, "final_rel_residual")[cells]
    return rr, np.where(own > 1e-10, 2.0 * own, 1e-10)
