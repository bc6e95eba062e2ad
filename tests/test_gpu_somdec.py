"""GPU parity of the SomDec decomposition step (reaction_sandbox_somdec.F90,
called from RReaction, reaction.F90:4059-4130) against the FP64 CPU oracle.

Bar: relative concentration error <= 1e-8 against the oracle on the same
seeded ELM-CN columns, the oracle's FP64 relative residual of the kernel's
answer <= 1e-10 (or the oracle's own where rounding keeps it above), and C
and N conserved by every kernel to the oracle's own bar (1e-11 / 1e-9).
"""
import numpy as np
import pytest

from pfchem import abi, workloads

import gpu_util as G
import somdec_util as SU

pytestmark = pytest.mark.gpu

CONC_TOL = 1e-8
VARIANTS = [abi.KERNEL_BATCHED_FP64, abi.KERNEL_BATCHED_MIXED]
# the one-cell-per-thread kernel's SomDec instantiation faulted the GPU with no
# cause found in its code; the library refuses it until one is
PLAIN_SKIP = ("plain kernel + SomDec faulted on MI355X, cause not found; pfchem_step refuses "
              "the combination")
FIELDS = ("pri_molal", "total", "immobile")


@pytest.fixture(scope="module")
def columns():
    wl = workloads.somdec_workload(8192)
    ref, rec, nfail = G.run_oracle(wl)
    assert nfail == 0
    return wl, ref, rec


def _check(wl, ref, rec, got, label):
    n = wl.state.ncell
    print("case", label, flush=True)
    assert np.all(got["status"][0, :n] == 0), "%s: failed cells" % label
    errs = G.compare(ref, got, n, fields=FIELDS)
    print(label, errs, flush=True)
    for f, e in errs.items():
        assert e <= CONC_TOL, "%s: %s rel err %.3e > %.1e" % (label, f, e, CONC_TOL)
    rr, bar = G.gate_rel_residual(wl, got, rec, range(n), ref)
    assert rr.size > 0.9 * n
    print(label, "max rel residual %.3e over %d cells" % (rr.max(), rr.size), flush=True)
    assert np.all(rr <= bar)
    C0, N0 = SU.carbon_nitrogen(wl.net, wl.state.arrays, n)
    C1, N1 = SU.carbon_nitrogen(wl.net, got, n)
    np.testing.assert_allclose(C1, C0, rtol=1e-11)
    np.testing.assert_allclose(N1, N0, rtol=1e-9)


@pytest.mark.parametrize("variant", VARIANTS)
def test_somdec_parity(columns, variant):
    wl, ref, rec = columns
    got, _ = G.run_gpu(wl, variant)
    _check(wl, ref, rec, got, "somdec/" + abi.KERNEL_NAMES[variant])


@pytest.mark.skip(reason=PLAIN_SKIP)
def test_somdec_parity_plain(columns):
    wl, ref, rec = columns
    got, _ = G.run_gpu(wl, abi.KERNEL_PLAIN_FP64)
    _check(wl, ref, rec, got, "somdec/plain")
    assert np.mean(ref["num_iterations"][0] == got["num_iterations"][0]) > 0.99


def test_somdec_plain_is_refused(columns):
    from pfchem.lib import Context, DeviceState, PfchemError
    wl, _, _ = columns
    ctx = Context(wl.net, wl.opt)
    with pytest.raises(PfchemError) as ei:
        ctx.step(DeviceState(wl.state), wl.dt, abi.KERNEL_PLAIN_FP64)
    assert ei.value.code == abi.PFCHEM_ERR_UNSUPPORTED


@pytest.mark.parametrize("lanes,rows,block", [(16, 1, 128), (8, 2, 128), (32, 1, 256)])
def test_somdec_team_shapes(columns, lanes, rows, block):
    wl, ref, rec = columns
    for variant in (abi.KERNEL_BATCHED_FP64, abi.KERNEL_BATCHED_MIXED):
        got, _ = G.run_gpu(wl, variant, lanes=lanes, block=block, rows=rows)
        _check(wl, ref, rec, got, "somdec/%dx%d/%s" % (lanes, rows, abi.KERNEL_NAMES[variant]))


def test_somdec_shapes_without_kernel_are_refused(columns):
    from pfchem.lib import Context
    wl, _, _ = columns
    ctx = Context(wl.net, wl.opt)
    assert not ctx.try_set_launch(abi.KERNEL_BATCHED_FP64, 4, 128, 4, 4)
    assert ctx.get_launch(abi.KERNEL_BATCHED_MIXED)["lanes_per_cell"] == 16


def test_somdec_dry_and_frozen_cells(columns):
    """f_w = 0 stops decomposition (pools unchanged); a cell below the
    minimum saturation skips the reactions entirely"""
    wl, _, _ = columns
    st = wl.state.copy()
    n = st.ncell
    st.arrays["somdec_f_w"][0, : n // 2] = 0.0
    st.arrays["saturation"][0, n // 2: n // 2 + 64] = 0.0
    ref, rec, nfail = G.run_oracle(wl, state=st)
    assert nfail == 0
    for variant in VARIANTS:
        got, _ = G.run_gpu(wl, variant, state=st)
        assert np.all(got["status"][0, :n] == 0)
        errs = G.compare(ref, got, n, fields=FIELDS)
        assert max(errs.values()) <= CONC_TOL, errs
        np.testing.assert_allclose(got["immobile"][:, : n // 2], st.arrays["immobile"][:, : n // 2],
                                   rtol=1e-12)
