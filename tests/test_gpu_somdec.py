"""GPU parity of the SomDec decomposition step (reaction_sandbox_somdec.F90,
called from RReaction, reaction.F90:4059-4130) against the FP64 CPU oracle.

The sandbox's Jacobian is the reference's, which is not the exact
derivative (aqueous Monod columns without the theta*1000 factor, the
NH4+/NO3- split's derivative switched off), so the Newton iteration
converges linearly in the aqueous species.  At the bench tolerances the
oracle's own answer then moves when its tolerance moves: tightening it ten
times changes N2O by up to 8e-3 and CO2 by 2e-6 relative on these columns.
Round-off (GPU vs host libm) decides at which iteration the relative-
residual test fires, so one iteration more or less than the oracle is a
legitimate outcome.  The bar is therefore, per cell and species, for cells
that took the oracle's sub-step path:

  |gpu - oracle| <= 1e-8 |oracle| + |oracle(tol*10) - oracle| + |oracle(tol/10) - oracle|

i.e. within 1e-8 of the oracle or within the spread of the oracle's own
answers under a ten-fold change of its convergence tolerances; plus the
oracle's FP64 relative residual of the GPU answer <= 1e-10 on one-sub-step
cells, or no more than the oracle itself accepted for that cell in those
runs (its relative-change exit stops a linearly converging iteration at
residuals up to ~6e-10 here) except for at most 0.1% of cells, which stay
below the largest residual the oracle accepted in any cell, C and N
conserved to the oracle's own bar in every cell, and at most
0.1% of cells on another sub-step path (a failed attempt whose outcome at the
iteration limit is decided by round-off).
"""
import dataclasses

import numpy as np
import pytest

from pfchem import abi, workloads

import gpu_util as G
import oracle_lib as O
import somdec_util as SU

pytestmark = pytest.mark.gpu

CONC_TOL = 1e-8
RESID_TOL = 1e-10
VARIANTS = [abi.KERNEL_BATCHED_FP64, abi.KERNEL_BATCHED_MIXED]
# the one-cell-per-thread kernel's SomDec instantiation faulted the GPU with no
# cause found in its code; the library refuses it until one is
PLAIN_SKIP = ("plain kernel + SomDec faulted on MI355X, cause not found; pfchem_step refuses "
              "the combination")
FIELDS = ("pri_molal", "total", "immobile")


def _with_tolerance(wl, scale):
    o = wl.opt
    return dataclasses.replace(wl, opt=dataclasses.replace(
        o, max_rel_residual_tolerance=o.max_rel_residual_tolerance * scale,
        max_relative_change_tolerance=o.max_relative_change_tolerance * scale))


def _oracle_band(wl, ref, state=None):
    """per-entry spread of the oracle's answers under a ten-fold looser and
    a ten-fold tighter convergence tolerance; "resid" holds the largest
    relative residual the oracle accepted per cell over the three runs (the
    relative-change exit fires on small steps of a linearly converging
    iteration, above the relative-residual tolerance)"""
    n = wl.state.ncell
    band = {}
    resid = None
    for scale in (10.0, 0.1):
        other, orec, _ = G.run_oracle(_with_tolerance(wl, scale), state=state)
        for f in FIELDS:
            d = np.abs(other[f][:, :n] - ref[f][:, :n])
            band[f] = band.get(f, 0.0) + d
        same = other["num_sub_steps"][0, :n] == ref["num_sub_steps"][0, :n]
        r = np.where(same, O.record_fields(orec, "final_rel_residual")[:n], 0.0)
        resid = r if resid is None else np.maximum(resid, r)
    band["resid"] = resid
    return band


@pytest.fixture(scope="module")
def columns():
    wl = workloads.somdec_workload(8192)
    ref, rec, nfail = G.run_oracle(wl)
    assert nfail == 0
    return wl, ref, rec, _oracle_band(wl, ref)


def _check(wl, ref, rec, band, got, label, state=None):
    n = wl.state.ncell
    print("case", label, flush=True)
    assert np.all(got["status"][0, :n] == 0), "%s: failed cells" % label
    same = ref["num_sub_steps"][0, :n] == got["num_sub_steps"][0, :n]
    print(label, "cells on another sub-step path:", int((~same).sum()), flush=True)
    assert (~same).sum() <= max(2, n // 1000)
    worst = {}
    for f in FIELDS:
        a = ref[f][:, :n]
        b = got[f][:, :n]
        err = np.abs(b - a)
        over = (err > CONC_TOL * np.abs(a) + band[f]) & same[None, :]
        rel = np.where(a != 0, err / np.abs(a), 0.0)[:, same]
        worst[f] = float(rel.max())
        assert not over.any(), "%s: %s outside the oracle band at %s" % (
            label, f, np.argwhere(over)[:5].tolist())
    print(label, "max rel err (same path)", worst, flush=True)
    # cells whose initial residual is zero (dry, or no reaction with f_w = 0)
    # have no relative residual
    moving = O.record_fields(rec, "r0_norm")[:n] > 0.0
    one = np.flatnonzero(same & moving & (ref["num_sub_steps"][0, :n] == 1))
    assert one.size > 0.9 * (same & moving).sum()
    rr = O.rel_residuals(wl.net, wl.opt, got, rec, one)
    own = np.maximum(O.record_fields(rec, "final_rel_residual")[one], band["resid"][one])
    bar = np.maximum(RESID_TOL, own)
    bad = np.flatnonzero(rr > bar)
    print(label, "max rel residual %.3e over %d cells (%d above %.0e, oracle accepted up to %.3e)"
          % (rr.max(), rr.size, (rr > RESID_TOL).sum(), RESID_TOL, own.max()), flush=True)
    for k in bad[:5]:
        print("   cell %d rr %.3e oracle %.3e its oracle/gpu %d/%d" % (
            one[k], rr[k], own[k], ref["num_iterations"][0, one[k]], got["num_iterations"][0, one[k]]))
    # per cell: within what the oracle accepted for it; a handful of cells
    # may exit on the relative-change test at another iterate, but none
    # above the largest residual the oracle accepted anywhere
    assert bad.size <= max(2, n // 1000)
    assert rr.max() <= max(RESID_TOL, own.max())
    init = (state or wl.state).arrays
    C0, N0 = SU.carbon_nitrogen(wl.net, init, n)
    C1, N1 = SU.carbon_nitrogen(wl.net, got, n)
    np.testing.assert_allclose(C1, C0, rtol=1e-11)
    np.testing.assert_allclose(N1, N0, rtol=1e-9)


@pytest.mark.parametrize("variant", VARIANTS)
def test_somdec_parity(columns, variant):
    wl, ref, rec, band = columns
    got, _ = G.run_gpu(wl, variant)
    _check(wl, ref, rec, band, got, "somdec/" + abi.KERNEL_NAMES[variant])


@pytest.mark.skip(reason=PLAIN_SKIP)
def test_somdec_parity_plain(columns):
    wl, ref, rec, band = columns
    got, _ = G.run_gpu(wl, abi.KERNEL_PLAIN_FP64)
    _check(wl, ref, rec, band, got, "somdec/plain")
    assert np.mean(ref["num_iterations"][0] == got["num_iterations"][0]) > 0.99


def test_somdec_plain_is_refused(columns):
    from pfchem.lib import Context, DeviceState, PfchemError
    wl = columns[0]
    ctx = Context(wl.net, wl.opt)
    with pytest.raises(PfchemError) as ei:
        ctx.step(DeviceState(wl.state), wl.dt, abi.KERNEL_PLAIN_FP64)
    assert ei.value.code == abi.PFCHEM_ERR_UNSUPPORTED


@pytest.mark.parametrize("lanes,rows,block", [(16, 1, 128), (8, 2, 128), (32, 1, 256)])
def test_somdec_team_shapes(columns, lanes, rows, block):
    wl, ref, rec, band = columns
    for variant in (abi.KERNEL_BATCHED_FP64, abi.KERNEL_BATCHED_MIXED):
        got, _ = G.run_gpu(wl, variant, lanes=lanes, block=block, rows=rows)
        _check(wl, ref, rec, band, got, "somdec/%dx%d/%s" % (lanes, rows, abi.KERNEL_NAMES[variant]))


def test_somdec_shapes_without_kernel_are_refused(columns):
    from pfchem.lib import Context
    wl = columns[0]
    ctx = Context(wl.net, wl.opt)
    assert not ctx.try_set_launch(abi.KERNEL_BATCHED_FP64, 4, 128, 4, 4)
    assert ctx.get_launch(abi.KERNEL_BATCHED_MIXED)["lanes_per_cell"] == 16


def test_somdec_dry_and_frozen_cells(columns):
    """f_w = 0 stops decomposition (pools unchanged); a cell below the
    minimum saturation skips the reactions entirely"""
    wl = columns[0]
    st = wl.state.copy()
    n = st.ncell
    st.arrays["somdec_f_w"][0, : n // 2] = 0.0
    st.arrays["saturation"][0, n // 2: n // 2 + 64] = 0.0
    ref, rec, nfail = G.run_oracle(wl, state=st)
    assert nfail == 0
    band = _oracle_band(wl, ref, state=st)
    for variant in VARIANTS:
        got, _ = G.run_gpu(wl, variant, state=st)
        _check(wl, ref, rec, band, got, "somdec-dry/" + abi.KERNEL_NAMES[variant], state=st)
        np.testing.assert_allclose(got["immobile"][:, : n // 2], st.arrays["immobile"][:, : n // 2],
                                   rtol=1e-12)


@pytest.fixture(scope="module")
def oxygen_columns():
    wl = workloads.somdec_workload(4096, oxygen=1.0e-5)
    ref, rec, nfail = G.run_oracle(wl)
    assert nfail == 0
    return wl, ref, rec, _oracle_band(wl, ref)


@pytest.mark.parametrize("variant", VARIANTS)
def test_somdec_oxygen_parity(oxygen_columns, variant):
    """the Monod oxygen response on every reaction (oxidizer O2(aq), K =
    1e-5 mol/L) with O2 consumption tracked: the oracle band as above, and
    total O2 + CO2 unchanged in every cell"""
    wl, ref, rec, band = oxygen_columns
    got, _ = G.run_gpu(wl, variant)
    _check(wl, ref, rec, band, got, "somdec-oxygen/" + abi.KERNEL_NAMES[variant])
    n = wl.state.ncell
    names = wl.net.primary_species_names
    ico2, io2 = names.index("CO2(aq)"), names.index("O2(aq)")
    s0 = wl.state["total"][ico2, :n] + wl.state["total"][io2, :n]
    s1 = got["total"][ico2, :n] + got["total"][io2, :n]
    np.testing.assert_allclose(s1, s0, rtol=1e-12)
    assert np.all(got["total"][io2, :n] < wl.state["total"][io2, :n])


@pytest.fixture(scope="module")
def complex_columns():
    wl = workloads.somdec_workload(4096, complexes=True)
    ref, rec, nfail = G.run_oracle(wl)
    assert nfail == 0
    return wl, ref, rec, _oracle_band(wl, ref)


@pytest.mark.parametrize("variant", VARIANTS)
def test_somdec_with_complexes_parity(complex_columns, variant):
    """the sandbox on a network with aqueous ion pairs (NH4NO3(aq),
    NH4CO2(aq)): totals and the dtotal entries it reads include the
    complexes; the oracle band as above, secondary molalities too"""
    wl, ref, rec, band = complex_columns
    got, _ = G.run_gpu(wl, variant)
    _check(wl, ref, rec, band, got, "somdec-complexes/" + abi.KERNEL_NAMES[variant])
    n = wl.state.ncell
    a, b = ref["sec_molal"][:, :n], got["sec_molal"][:, :n]
    same = ref["num_sub_steps"][0, :n] == got["num_sub_steps"][0, :n]
    rel = np.abs(b - a)[:, same] / np.abs(a)[:, same]
    print("somdec-complexes sec_molal max rel err", float(rel.max()), flush=True)
    assert rel.max() < 1e-6
