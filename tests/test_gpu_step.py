"""GPU parity of the reaction step (RStep, reaction.F90:3564) against the
FP64 CPU oracle on the same seeded inputs.

Bar (stated per test): relative concentration error <= 1e-8 against the
oracle (free and secondary molalities relative to themselves; totals
relative to the magnitude of the terms they are summed from, since a total
that cancels amplifies its terms' rounding: on the uneven draw a total of
1.1e-9 mol/L summed from terms near 1e-4 differs from the oracle's by
1.2e-9 of itself in the plain FP64 kernel and 4e-7 in the mixed one, whose
free molalities agree to 5e-12), and the oracle's FP64 relative residual of
the kernel's answer <= 1e-10.  Calls go through the C ABI (libpfchem.so via
ctypes).
"""
import numpy as np
import pytest

from pfchem import abi, workloads

import gpu_util as G

pytestmark = pytest.mark.gpu

CONC_TOL = 1e-8      # relative concentration error vs the FP64 oracle
RESID_TOL = 1e-10    # FP64 relative residual of the kernel's answer (or the
                     # oracle's own, where rounding keeps it above 1e-10)

VARIANTS = [abi.KERNEL_PLAIN_FP64, abi.KERNEL_BATCHED_FP64, abi.KERNEL_BATCHED_MIXED]


@pytest.fixture(scope="module")
def calcite():
    wl = workloads.calcite_workload(1000)
    ref, rec, nfail = G.run_oracle(wl)
    assert nfail == 0
    return wl, ref, rec


@pytest.fixture(scope="module")
def randnet():
    wl = workloads.random_network_workload(ncell=8192)
    ref, rec, nfail = G.run_oracle(wl)
    assert nfail == 0
    return wl, ref, rec


def _check(wl, ref, rec, got, label):
    n = wl.state.ncell
    print("case", label, flush=True)
    assert np.all(got["status"][0, :n] == 0), "%s: failed cells" % label
    errs = G.compare(ref, got, n, net=wl.net)
    plain = G.compare(ref, got, n, fields=("total",))
    print(label, errs, "totals relative to themselves:", plain, flush=True)
    for f, e in errs.items():
        assert e <= CONC_TOL, "%s: %s rel err %.3e > %.1e" % (label, f, e, CONC_TOL)
    rr, bar = G.gate_rel_residual(wl, got, rec, range(n), ref)
    assert rr.size > 0.9 * n
    print(label, "max rel residual %.3e over %d cells (bar max %.3e)" % (rr.max(), rr.size, bar.max()),
          flush=True)
    assert np.all(rr <= bar)


@pytest.mark.parametrize("variant", VARIANTS)
def test_calcite_parity(calcite, variant):
    wl, ref, rec = calcite
    got, _ = G.run_gpu(wl, variant)
    _check(wl, ref, rec, got, "calcite/" + abi.KERNEL_NAMES[variant])
    if variant == abi.KERNEL_PLAIN_FP64:
        # same algorithm and tolerances: same Newton iteration counts
        it_ref = ref["num_iterations"][0]
        it_gpu = got["num_iterations"][0]
        assert np.mean(it_ref == it_gpu) > 0.99


@pytest.mark.parametrize("variant", VARIANTS)
def test_random_network_parity(randnet, variant):
    wl, ref, rec = randnet
    got, _ = G.run_gpu(wl, variant)
    _check(wl, ref, rec, got, "random15/" + abi.KERNEL_NAMES[variant])


@pytest.fixture(scope="module")
def uneven():
    # the same network, initial states drawn +-0.5 decades around the
    # reference composition and mixed with the next cell's at w ~ U[0, 1]:
    # 1 to ~40 Newton iterations per cell
    wl = workloads.random_network_workload(ncell=16384, spread=0.5, wmax=1.0)
    ref, rec, nfail = G.run_oracle(wl)
    assert nfail == 0
    it = ref["num_iterations"][0, :wl.state.ncell]
    assert it.max() >= 3 * np.median(it)
    return wl, ref, rec


@pytest.mark.parametrize("variant", VARIANTS)
def test_uneven_iteration_counts_parity(uneven, variant):
    wl, ref, rec = uneven
    got, _ = G.run_gpu(wl, variant)
    _check(wl, ref, rec, got, "random15-uneven/" + abi.KERNEL_NAMES[variant])


def test_long_hand_over_list():
    """the mixed step on 131,072 cells of the uneven draw: the FP32 pass hands
    over about a tenth of the cells, more than the 4,096 above which the
    FP64 one-lane kernel steps the listed cells before the team pass"""
    wl = workloads.random_network_workload(ncell=131072, spread=0.5, wmax=1.0)
    ref, rec, nfail = G.run_oracle(wl)
    assert nfail == 0
    got, _ = G.run_gpu(wl, abi.KERNEL_BATCHED_MIXED)
    _check(wl, ref, rec, got, "random15-uneven-131072/batched_mixed")


def _run_shape(wl, variant, lanes, rows, block):
    """run with the requested team shape, halving the workgroup size until
    the team tiles fit the device's LDS; None when no size fits"""
    from pfchem.lib import Context
    ctx = Context(wl.net, wl.opt)
    while block >= 64 and not ctx.try_set_launch(variant, lanes, block, 4, rows):
        block //= 2
    if block < 64:
        return None, None
    got, _ = G.run_gpu(wl, variant, ctx=ctx)
    return got, block


@pytest.mark.parametrize("lanes,rows,block", [(16, 1, 64), (16, 1, 256), (32, 1, 128), (8, 2, 128),
                                             (4, 4, 128), (2, 8, 64), (8, 4, 64), (32, 1, 64)])
def test_launch_shapes_agree(randnet, lanes, rows, block):
    """The autotuner's choices (lanes per cell, rows per lane, workgroup size)
    must not change the answer beyond the parity bar."""
    wl, ref, rec = randnet
    ran = 0
    for variant in (abi.KERNEL_BATCHED_MIXED, abi.KERNEL_BATCHED_FP64):
        got, b = _run_shape(wl, variant, lanes, rows, block)
        if got is None:
            continue
        ran += 1
        _check(wl, ref, rec, got, "random15/%s/L%dR%d/B%d" % (abi.KERNEL_NAMES[variant], lanes,
                                                              rows, b))
    if not ran:
        pytest.skip("team tiles of L%dR%d exceed the device's LDS" % (lanes, rows))


ONE_LANE_SKIP = ("one-lane teams (1x4, 1x8) returned wrong answers and faulted on MI355X, cause "
                 "not found; the library no longer offers them")
EIGHT_SKIP = ("8x1 teams did not finish on the calcite cells on MI355X, cause not found; the "
              "library no longer offers the shape")


@pytest.mark.parametrize("lanes,rows", [
    pytest.param(1, 4, marks=pytest.mark.skip(reason=ONE_LANE_SKIP)), (4, 1), (2, 4),
    pytest.param(8, 1, marks=pytest.mark.skip(reason=EIGHT_SKIP)),
    pytest.param(1, 8, marks=pytest.mark.skip(reason=ONE_LANE_SKIP))])
def test_small_team_shapes(calcite, lanes, rows):
    """Teams smaller than a quad on the three-component calcite system."""
    wl, ref, rec = calcite
    for variant in (abi.KERNEL_BATCHED_MIXED, abi.KERNEL_BATCHED_FP64):
        got, b = _run_shape(wl, variant, lanes, rows, 128)
        assert got is not None
        _check(wl, ref, rec, got, "calcite/%s/L%dR%d/B%d" % (abi.KERNEL_NAMES[variant], lanes,
                                                             rows, b))


def test_deterministic_and_order_independent(randnet):
    """Per-cell results do not depend on run or on cell order (each cell's
    state is private to its team; no shared sandbox state)."""
    wl, ref, rec = randnet
    a, ctx = G.run_gpu(wl, abi.KERNEL_BATCHED_MIXED)
    b, _ = G.run_gpu(wl, abi.KERNEL_BATCHED_MIXED, ctx=ctx)
    for f in G.FIELDS_F64:
        assert np.array_equal(a[f], b[f]), f
    perm = np.random.default_rng(7).permutation(wl.state.ncell)
    sub = wl.state.subset(perm)
    c, _ = G.run_gpu(wl, abi.KERNEL_BATCHED_MIXED, state=sub, ctx=ctx)
    for f in ("pri_molal", "total", "mnrl_volfrac"):
        assert np.array_equal(c[f], a[f][:, perm]), f


def test_ld_padding_and_partial_range(randnet):
    """ld > ncell and stepping a prefix of the cells leaves the rest intact."""
    wl, ref, rec = randnet
    n = wl.state.ncell
    padded = workloads.random_network_workload(ncell=n, ld=n + 64)
    ctx = None
    import torch
    from pfchem.lib import Context, DeviceState
    ctx = Context(padded.net, padded.opt)
    ds = DeviceState(padded.state)
    half = n // 2
    ctx.step(ds, padded.dt, abi.KERNEL_BATCHED_MIXED, ncell=half)
    torch.cuda.synchronize()
    got = ds.to_host()
    errs = G.compare(ref, got, half, net=wl.net)
    assert max(errs.values()) <= CONC_TOL
    assert np.array_equal(got["pri_molal"][:, half:], padded.state["pri_molal"][:, half:])


def test_pass_timing_hook(randnet):
    """pfchem_set_pass_timing / pfchem_pass_time: no time before a timed
    step; after one, the one-lane pass's duration, and the step's answer is
    the untimed one's"""
    from pfchem.lib import Context, DeviceState, PfchemError
    wl, ref, rec = randnet
    ctx = Context(wl.net, wl.opt)
    ctx.set_pass_timing(True)
    with pytest.raises(PfchemError):
        ctx.pass_time()
    got, _ = G.run_gpu(wl, abi.KERNEL_BATCHED_MIXED, ctx=ctx)
    ms = ctx.pass_time()
    print("one-lane pass %.3f ms" % ms, flush=True)
    assert 0.0 < ms < 1.0e4
    _check(wl, ref, rec, got, "random15/timed-pass")
    ctx.set_pass_timing(False)
