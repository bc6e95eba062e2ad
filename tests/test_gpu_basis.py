"""GPU reaction step on a network read from a PFLOTRAN database in two bases
(calcite.dat: primaries H+, HCO3-, Ca++ and the swapped H+, CO3--, Ca++):
the same physical cells, written in either basis, end the step in the same
state (species molalities matched by name, mineral rate and volume
fraction), at 1e-8 relative for the FP64 and mixed batched kernels and the
plain kernel."""
import numpy as np
import pytest

import test_database_basis as TB
from pfchem import abi
from pfchem.network import CellState

pytestmark = pytest.mark.gpu


def _cells(net, pri_of, ncell, names):
    st = CellState(net, ncell)
    n = net.naqcomp
    for c in range(ncell):
        one, _ = TB._one_cell(net, [pri_of[c][k] for k in names])
        for f, a in one.arrays.items():
            if a.shape[0]:
                st[f][:, c] = a[:, 0]
    return st


@pytest.mark.parametrize("variant", [abi.KERNEL_PLAIN_FP64, abi.KERNEL_BATCHED_FP64,
                                     abi.KERNEL_BATCHED_MIXED])
def test_step_is_basis_invariant_on_gpu(variant):
    import torch
    from pfchem.lib import Context, DeviceState
    _, a, b = TB.nets(TB.C_PRI, TB.C_SEC)
    rng = np.random.default_rng(5)
    ncell = 128
    pri_of = []
    # pH 6-7.5 keeps every component total positive in both bases (RStep
    # replaces totals <= 1e-40 by 1e-40, which would change the problem)
    for c in range(ncell):
        m = TB._molalities(a, [10 ** rng.uniform(-7.5, -6.0), 10 ** rng.uniform(-3.5, -2.0),
                               10 ** rng.uniform(-3.5, -2.5)])
        pri_of.append(m)
    sa = _cells(a, pri_of, ncell, TB.A_PRI)
    sb = _cells(b, pri_of, ncell, TB.C_PRI)
    assert np.all(sa["total"][:3, :ncell] > 0) and np.all(sb["total"][:3, :ncell] > 0)
    _, opt = TB._one_cell(a, [pri_of[0][k] for k in TB.A_PRI])
    dt = 3.0e4
    out = []
    print("case basis", abi.KERNEL_NAMES[variant], flush=True)
    for net, st in ((a, sa), (b, sb)):
        ctx = Context(net, opt)
        ds = DeviceState(st)
        ctx.step(ds, dt, variant)
        torch.cuda.synchronize()
        out.append(ds.to_host())
    ga, gb = out
    assert np.all(ga["status"][0, :ncell] == 0) and np.all(gb["status"][0, :ncell] == 0)
    fa = dict(zip(TB.A_PRI, ga["pri_molal"][:3, :ncell]))
    fa.update(zip(TB.A_SEC, ga["sec_molal"][:, :ncell]))
    fb = dict(zip(TB.C_PRI, gb["pri_molal"][:3, :ncell]))
    fb.update(zip(TB.C_SEC, gb["sec_molal"][:, :ncell]))
    worst = 0.0
    for name in fa:
        worst = max(worst, float(np.max(np.abs(fb[name] - fa[name]) / np.abs(fa[name]))))
    rr = np.abs(gb["mnrl_rate"][0, :ncell] - ga["mnrl_rate"][0, :ncell]) / \
        np.abs(ga["mnrl_rate"][0, :ncell])
    print("basis invariance max rel diff: species %.2e rate %.2e" % (worst, rr.max()), flush=True)
    assert worst <= 1e-8
    assert rr.max() <= 1e-8
