"""Host side of per-cell active primaries: the sub-network a cell with nact
active primaries solves, and the grouped oracle run."""
import numpy as np

import gpu_util as G
import variable_util as V
from pfchem import workloads


def test_sub_network_keeps_reactions_within_the_active_primaries():
    wl = workloads.variable_network_workload(64)
    net = wl.net
    for nact in (8, 16, 33, 64):
        sub, kc, km = V.sub_network(net, nact)
        assert sub.naqcomp == nact
        mx = np.array([max(s) for s in net.eqcplx_species])
        assert np.array_equal(kc, np.flatnonzero(mx < nact))
        mm = np.array([max(s) for s in net.kinmnrl_species])
        assert np.array_equal(km, np.flatnonzero(mm < nact))
        for j, k in enumerate(kc):
            assert sub.eqcplx_species[j] == list(net.eqcplx_species[k])
            assert sub.eqcplx_logK[j] == net.eqcplx_logK[k]
    sub, kc, km = V.sub_network(net, 64)
    assert kc.size == len(net.eqcplx_species) and km.size == len(net.kinmnrl_species)


def test_all_active_cells_equal_the_full_network_oracle():
    wl = workloads.variable_network_workload(256)
    st = wl.state.copy()
    st.arrays["ncomp_active"][0, :] = 64
    out, _ = V.run_oracle_variable(wl, state=st)
    import dataclasses
    full = dataclasses.replace(wl, opt=dataclasses.replace(wl.opt, variable_ncomp=False))
    ref, _, nfail = G.run_oracle(full, state=st)
    assert nfail == 0
    for f in ("pri_molal", "total", "sec_molal", "mnrl_volfrac"):
        np.testing.assert_array_equal(out[f], ref[f][:, :256])


def test_inactive_entries_keep_their_inputs():
    wl = workloads.variable_network_workload(256)
    out, _ = V.run_oracle_variable(wl)
    nact = wl.state["ncomp_active"][0]
    for c in range(0, 256, 17):
        np.testing.assert_array_equal(out["pri_molal"][nact[c]:, c],
                                      wl.state["pri_molal"][nact[c]:, c])
        sub, kc, km = V.sub_network(wl.net, int(nact[c]))
        off = np.setdiff1d(np.arange(len(wl.net.eqcplx_species)), kc)
        np.testing.assert_array_equal(out["sec_molal"][off, c], wl.state["sec_molal"][off, c])
