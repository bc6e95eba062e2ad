"""GPU parity of the variable-size step: a network of 64 primaries, 300
complexes and 10 kinetic minerals in which every cell has its own number of
active primaries (options.variable_ncomp, pfchem_state.ncomp_active), so
cells solve systems from 8x8 to 64x64.  Each cell's answer is the
reference's RStep (reaction.F90:3564) on the network restricted to its
active primaries, computed by the FP64 oracle per group of cells.

Bar: relative concentration error <= 1e-8 against the oracle, oracle FP64
relative residual of the kernel's answer <= 1e-10 (or the oracle's own),
and entries outside a cell's sub-network untouched.
"""
import numpy as np
import pytest

from pfchem import abi, workloads

import gpu_util as G
import variable_util as V

pytestmark = pytest.mark.gpu

CONC_TOL = 1e-8
RESID_TOL = 1e-10
FIELDS = ("pri_molal", "total", "sec_molal", "mnrl_volfrac", "mnrl_rate")


@pytest.fixture(scope="module")
def cells():
    wl = workloads.variable_network_workload(4096)
    ref, recs = V.run_oracle_variable(wl)
    assert np.all(ref["status"][0] == 0)
    return wl, ref, recs


@pytest.mark.parametrize("variant", [abi.KERNEL_PLAIN_FP64, abi.KERNEL_BATCHED_FP64,
                                     abi.KERNEL_BATCHED_MIXED])
def test_variable_parity(cells, variant):
    wl, ref, recs = cells
    n = wl.state.ncell
    label = "variable/" + abi.KERNEL_NAMES[variant]
    print("case", label, flush=True)
    got, ctx = G.run_gpu(wl, variant)
    print(label, ctx.get_launch(variant), flush=True)
    assert np.all(got["status"][0, :n] == 0)
    errs = G.compare(ref, got.arrays, n, fields=FIELDS)
    print(label, errs, flush=True)
    for f, e in errs.items():
        assert e <= CONC_TOL, (f, e)
    rr, own = V.rel_residuals_variable(wl, got.arrays, recs)
    print(label, "max rel residual %.3e over %d cells" % (rr.max(), rr.size), flush=True)
    assert np.all(rr <= np.maximum(RESID_TOL, own))
    # primaries beyond a cell's active count are untouched
    nact = wl.state["ncomp_active"][0, :n]
    rows = np.arange(wl.net.naqcomp)[:, None]
    off = rows >= nact[None, :]
    assert np.array_equal(got["pri_molal"][:, :n][off], wl.state["pri_molal"][:, :n][off])
