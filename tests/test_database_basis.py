"""Basis swapping in the network builder (BasisInit, reaction_database.F90:
1168-1448; BasisSubSpeciesInMineralRxn, reaction_database_aux.F90:494-586)
against what the chemistry fixes: the swapped reactions are the textbook
rearrangements, mass action gives the same species molalities in either
basis, and a kinetic reaction step run by the oracle in either basis ends at
the same physical state."""
import numpy as np
import pytest

import golden_util as GU
import oracle_lib as O
from pfchem import abi, database
from pfchem.network import CellState, SolverOptions

A_PRI = ["H+", "HCO3-", "Ca++"]
B_PRI = ["H+", "CO2(aq)", "Ca++"]
A_SEC = ["OH-", "CO3--", "CO2(aq)", "CaOH+", "CaHCO3+", "CaCO3(aq)"]
B_SEC = ["OH-", "CO3--", "HCO3-", "CaOH+", "CaHCO3+", "CaCO3(aq)"]
# a basis in which every component total stays positive (RStep replaces
# totals <= 1e-40 by 1e-40, so a negative proton total in the CO2(aq) basis
# would change the problem)
C_PRI = ["H+", "CO3--", "Ca++"]
C_SEC = ["OH-", "HCO3-", "CO2(aq)", "CaOH+", "CaHCO3+", "CaCO3(aq)"]
CALCITE = database.KineticMineralSpec("Calcite", rate_constant=1.0e-9)


def nets(pri=B_PRI, sec=B_SEC):
    db = GU.db("calcite.dat")
    a = database.build_network(db, A_PRI, A_SEC, [CALCITE], temperature=25.0, isothermal=True)
    b = database.build_network(db, pri, sec, [CALCITE], temperature=25.0, isothermal=True)
    return db, a, b


def test_swapped_reactions_are_the_rearranged_database_reactions():
    db, a, b = nets()
    lk = {n: k for n, k in zip(A_SEC, a.eqcplx_logK)}
    # CO2(aq) = -H2O + H+ + HCO3-  ==>  HCO3- = CO2(aq) + H2O - H+
    k = B_SEC.index("HCO3-")
    got = dict(zip(b.eqcplx_species[k], b.eqcplx_stoich[k]))
    assert got == {0: -1.0, 1: 1.0}
    assert b.eqcplxh2ostoich[k] == 1.0
    assert b.eqcplx_logK[k] == -lk["CO2(aq)"]
    # CO3-- = -H+ + HCO3-  ==>  CO3-- = CO2(aq) + H2O - 2 H+
    k = B_SEC.index("CO3--")
    assert dict(zip(b.eqcplx_species[k], b.eqcplx_stoich[k])) == {0: -2.0, 1: 1.0}
    assert b.eqcplxh2ostoich[k] == 1.0
    assert np.isclose(b.eqcplx_logK[k], lk["CO3--"] - lk["CO2(aq)"], rtol=0, atol=1e-14)
    # Calcite = -H+ + Ca++ + HCO3-  ==>  -2 H+ + Ca++ + CO2(aq) + H2O
    got = dict(zip(b.kinmnrl_species[0], b.kinmnrl_stoich[0]))
    assert got == {0: -2.0, 1: 1.0, 2: 1.0}
    assert b.kinmnrlh2ostoich[0] == 1.0
    assert np.isclose(b.kinmnrl_logK[0], a.kinmnrl_logK[0] - lk["CO2(aq)"], rtol=0, atol=1e-14)
    # the unswapped basis reproduces the database reactions exactly
    assert dict(zip(a.kinmnrl_species[0], a.kinmnrl_stoich[0])) == {0: -1.0, 1: 1.0, 2: 1.0}


def test_not_a_basis_is_rejected():
    db = GU.db("calcite.dat")
    with pytest.raises(ValueError):
        # CO3-- named, HCO3- neither primary nor secondary
        database.build_network(db, ["H+", "Ca++"], ["CO3--"], isothermal=True)


def _molalities(net, pri):
    """all species molalities by name at unit activity coefficients"""
    opt = SolverOptions(act_coef_update_frequency=abi.ACT_COEF_FREQUENCY_OFF)
    cc = O.CellCtx(net, opt)
    n = net.naqcomp
    cc.set("pri_molal", pri)
    cc.set("pri_act_coef", np.ones(n))
    cc.set("sec_act_coef", np.ones(net.neqcplx))
    cc.c.den_kg = 997.16
    cc.total()
    out = dict(zip(net.primary_species_names, pri))
    out.update(zip(net.eqcplx_names, cc.get("sec_molal", net.neqcplx)))
    return out


@pytest.mark.parametrize("pri,sec", [(B_PRI, B_SEC), (C_PRI, C_SEC)])
def test_mass_action_is_basis_invariant(pri, sec):
    _, a, b = nets(pri, sec)
    ma = _molalities(a, [10 ** -7.5, 2.0e-3, 1.0e-3])
    mb = _molalities(b, [ma[n] for n in pri])
    for name in set(ma) | set(mb):
        assert np.isclose(mb[name], ma[name], rtol=1e-12, atol=0), name


def _one_cell(net, pri):
    opt = SolverOptions(act_coef_update_frequency=abi.ACT_COEF_FREQUENCY_OFF,
                        max_rel_residual_tolerance=1e-13, max_relative_change_tolerance=1e-14,
                        max_residual_tolerance=1e-30, maximum_reaction_iterations=50)
    st = CellState(net, 1)
    n = net.naqcomp
    st["pri_molal"][:n, 0] = pri
    st["pri_act_coef"][:n, 0] = 1.0
    st["sec_act_coef"][:, 0] = 1.0
    st["porosity"][0, 0] = 0.3
    st["saturation"][0, 0] = 1.0
    st["den_kg"][0, 0] = 997.16
    st["volume"][0, 0] = 1.0
    st["mnrl_volfrac"][0, 0] = 0.1
    st["mnrl_area"][0, 0] = 1.0
    cc = O.CellCtx(net, opt).gather(st, 0)
    cc.total()
    st["total"][:n, 0] = cc.get("total", n)
    st["sec_molal"][:, 0] = cc.get("sec_molal", net.neqcplx)
    return st, opt


def test_kinetic_step_is_basis_invariant():
    """the oracle's RStep on the same physical cell, written in either basis,
    reaches the same free and complexed molalities, rate and volume
    fraction (Newton tolerances 1e-13; bar 1e-8 relative)"""
    _, a, b = nets(C_PRI, C_SEC)
    ma = _molalities(a, [10 ** -7.5, 2.0e-3, 1.0e-3])
    sa, opt = _one_cell(a, [ma[n] for n in A_PRI])
    sb, _ = _one_cell(b, [ma[n] for n in C_PRI])
    assert np.all(sb["total"][:3, 0] > 0)
    dt = 3.0e4
    assert O.step_batch(a, opt, sa, dt, 1) == 0
    assert O.step_batch(b, opt, sb, dt, 1) == 0
    fa = dict(zip(A_PRI, sa["pri_molal"][:3, 0]))
    fa.update(zip(A_SEC, sa["sec_molal"][:, 0]))
    fb = dict(zip(C_PRI, sb["pri_molal"][:3, 0]))
    fb.update(zip(C_SEC, sb["sec_molal"][:, 0]))
    for name in fa:
        assert np.isclose(fb[name], fa[name], rtol=1e-8, atol=0), name
    # the step moved the state: calcite dissolved / precipitated measurably
    assert abs(sa["mnrl_volfrac"][0, 0] - 0.1) > 1e-12
    assert np.isclose(sb["mnrl_rate"][0, 0], sa["mnrl_rate"][0, 0], rtol=1e-8, atol=0)
    assert np.isclose(sb["mnrl_volfrac"][0, 0], sa["mnrl_volfrac"][0, 0], rtol=1e-12, atol=0)
