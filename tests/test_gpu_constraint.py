"""GPU constraint speciation (pfchem_equilibrate, one cell per thread, FP64)
against the oracle's restatement of ReactionEquilibrateConstraint on the
same cells: free-ion molalities, totals, secondaries and activity
coefficients to 1e-9 relative, status codes equal, Newton iteration counts
equal in at least 99% of cells (device and host exp/log/pow may differ in
the last bit, which can move a cell across a convergence threshold)."""
import numpy as np
import pytest

import golden_util as GU
import oracle_lib as O
import test_oracle_constraint as T
from pfchem import abi, constraint

pytestmark = pytest.mark.gpu
NCELL = 2048


def _case(which, rng):
    n = NCELL
    if which == "ca-carbonate-totals":
        db, net = T._ca_carbonate()
        con = constraint.from_deck(db, net, {}, 25.0, True)
        v = np.stack([10 ** rng.uniform(-3.5, -2.0, n), 10 ** rng.uniform(-4.0, -2.5, n),
                      10 ** rng.uniform(-4.0, -2.5, n)])
    elif which == "ph-total-charge":
        db = T._db_with_co2_gas()
        net = T._calcite_net(db)
        con = constraint.from_deck(db, net, {"H+": ("P",), "HCO3-": ("T",), "Ca++": ("Z",)},
                                   25.0, True)
        # pH >= 5 with these totals: the free HCO3- outweighs H+, so charge
        # balance is reachable with a positive Ca++ molality
        v = np.stack([rng.uniform(5.0, 7.5, n), 10 ** rng.uniform(-3.0, -2.0, n),
                      np.full(n, 1e-6)])
    elif which == "free-gas-mineral":
        db = T._db_with_co2_gas()
        net = T._calcite_net(db)
        con = constraint.from_deck(db, net, {"H+": ("F",), "HCO3-": ("G", "CO2(g)"),
                                             "Ca++": ("M", "Calcite")}, 25.0, True)
        v = np.stack([10 ** rng.uniform(-8.5, -7.0, n), rng.uniform(-4.0, -2.0, n),
                      np.full(n, 5e-4)])
    else:
        raise ValueError(which)
    opt = GU.lag_options(act_coef_update_frequency=abi.ACT_COEF_FREQUENCY_TIMESTEP)
    return net, opt, con, np.ascontiguousarray(v)


def _compare(net, ref, got, n):
    worst = 0.0
    for f in ("pri_molal", "total", "sec_molal", "pri_act_coef", "sec_act_coef"):
        a, b = ref[f][:, :n], got[f][:, :n]
        scale = np.maximum(np.abs(a), 1e-300)
        if f == "total":  # totals may cancel to ~0 (charge-balanced H+ rows)
            scale = np.maximum(scale, 1e-12 * np.max(np.abs(a), axis=1, keepdims=True))
        worst = max(worst, float(np.max(np.abs(b - a) / scale)))
    same_it = float(np.mean(ref["num_iterations"][0, :n] == got["num_iterations"][0, :n]))
    return worst, same_it


@pytest.mark.parametrize("which", ["ca-carbonate-totals", "ph-total-charge", "free-gas-mineral"])
def test_equilibrate_matches_oracle(which):
    import torch
    from pfchem.lib import Context, DeviceState
    print("case equilibrate", which, NCELL, flush=True)
    rng = np.random.default_rng(11)
    net, opt, con, v = _case(which, rng)
    st = T._state(net, NCELL)
    ref = st.copy()
    assert O.equilibrate_batch(net, opt, con, ref, v, 0) == 0
    ctx = Context(net, opt)
    ds = DeviceState(st)
    conc = torch.from_numpy(v).cuda()
    ctx.equilibrate(ds, con, conc)
    got = ds.to_host()
    assert np.all(got["status"][0, :NCELL] == 0)
    worst, same_it = _compare(net, ref, got, NCELL)
    print("equilibrate %s: max rel diff %.2e, same iteration count %.4f" % (which, worst, same_it),
          flush=True)
    assert worst <= 1e-9
    assert same_it >= 0.99
    # restart from the answer: fewer iterations, same state
    con.use_prev_soln_as_guess = True
    ref2 = ref.copy()
    assert O.equilibrate_batch(net, opt, con, ref2, v, 0) == 0
    ctx.equilibrate(ds, con, conc)
    got2 = ds.to_host()
    worst2, _ = _compare(net, ref2, got2, NCELL)
    assert worst2 <= 1e-9
    assert np.mean(got2["num_iterations"][0, :NCELL]) < np.mean(got["num_iterations"][0, :NCELL])


def test_equilibrate_failures_and_argument_checks():
    import torch
    from pfchem.lib import Context, DeviceState, PfchemError
    print("case equilibrate failures", flush=True)
    rng = np.random.default_rng(3)
    net, opt, con, v = _case("ca-carbonate-totals", rng)
    ctx = Context(net, opt)
    st = T._state(net, NCELL)
    ds = DeviceState(st)
    conc = torch.from_numpy(v).cuda()
    con.max_iterations = 2
    with pytest.raises(PfchemError) as e:
        ctx.equilibrate(ds, con, conc)
    assert e.value.code == abi.PFCHEM_ERR_CELLS
    got = ds.to_host()
    assert np.all(got["status"][0, :NCELL] == abi.PFCHEM_CELL_CONSTRAINT_MAX_ITERATIONS)
    ref = T._state(net, NCELL)
    O.equilibrate_batch(net, opt, con, ref, v, 0)
    assert np.all(ref["status"][0, :NCELL] == abi.PFCHEM_CELL_CONSTRAINT_MAX_ITERATIONS)
    # a redox constraint is refused by the host mirror and by the C ABI,
    # before any launch
    import ctypes as C
    con.max_iterations = 10000
    k = con.to_abi(net)
    types = (C.c_int32 * 3)(abi.CONSTRAINT_PE, abi.CONSTRAINT_TOTAL, abi.CONSTRAINT_TOTAL)
    k.constraint_type = C.cast(types, C.POINTER(C.c_int32))
    s = ds.to_abi()
    rc = ctx.L.pfchem_equilibrate(ctx.h, NCELL, C.byref(s), C.byref(k), conc.data_ptr(), None)
    assert rc == abi.PFCHEM_ERR_UNSUPPORTED
    con.types = [abi.CONSTRAINT_PE] + list(con.types[1:])
    with pytest.raises(NotImplementedError):
        ctx.equilibrate(ds, con, conc)
    # a pH constraint in a network without H+
    con2 = constraint.from_deck(*T._ca_carbonate(), {}, 25.0, True)
    k = con2.to_abi(net)
    k.h_ion_id = -(1 << 30)
    types = (C.c_int32 * 3)(abi.CONSTRAINT_TOTAL, abi.CONSTRAINT_TOTAL, abi.CONSTRAINT_PH)
    k.constraint_type = C.cast(types, C.POINTER(C.c_int32))
    rc = ctx.L.pfchem_equilibrate(ctx.h, NCELL, C.byref(s), C.byref(k), conc.data_ptr(), None)
    assert rc == abi.PFCHEM_ERR_INVALID
