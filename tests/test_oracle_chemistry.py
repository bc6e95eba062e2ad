"""The oracle's chemistry against what the mathematics fixes: Debye-Hueckel
special cases and limits, mass action and its derivatives against finite
differences, TST rate signs and zeros, the analytic Jacobian against the
reference's own numerical-derivative scheme (RReactionDerivative,
reaction.F90:4134-4239), and conservation / control flow of RStep."""
import numpy as np
import pytest

import oracle_lib as O
from pfchem import abi, workloads
from pfchem.network import ReactionNetwork, SolverOptions

LAG = SolverOptions(act_coef_update_algorithm=abi.ACT_COEF_ALGORITHM_LAG)


def ion_net(Z, a0):
    return ReactionNetwork(primary_species_names=["s%d" % i for i in range(len(Z))],
                           primary_spec_Z=Z, primary_spec_a0=a0)


def gammas_at(Z, a0, I):
    """activity coefficients of ions (Z, a0) in a solution of ionic strength I
    made of a 1:1 background electrolyte (species 0, 1)"""
    Zb = [1.0, -1.0] + list(Z)
    ab = [4.0, 4.0] + list(a0)
    cc = O.CellCtx(ion_net(Zb, ab), LAG)
    cc.set("pri_molal", [I, I] + [0.0] * len(Z))
    cc.activity()
    return cc.get("pri_act_coef", len(Zb))[2:]


def test_debye_huckel_neutral_and_dilute():
    g = gammas_at([0.0, 1.0, -2.0], [3.0, 9.0, 4.0], 0.1)
    assert g[0] == 1.0                       # Z = 0: gamma = 1 (reaction.F90:4583)
    g0 = gammas_at([1.0, 2.0, -3.0], [4.0, 6.0, 5.0], 1e-24)
    np.testing.assert_allclose(g0, 1.0, atol=1e-10)


def test_debye_huckel_limiting_law():
    """ln gamma -> -ln10 A z^2 sqrt(I) as I -> 0 (Debye-Hueckel limiting law)"""
    for z in (1.0, 2.0, 3.0):
        I = 1e-10
        g = gammas_at([z], [5.0], I)[0]
        lim = -np.log(10.0) * 0.5114 * z * z * np.sqrt(I)
        assert np.log(g) == pytest.approx(lim, rel=1e-4)


def test_debye_huckel_charge_symmetry_and_monotone():
    I = np.array([1e-4, 1e-3, 1e-2, 5e-2, 1e-1])
    gp = np.array([gammas_at([2.0], [6.0], i)[0] for i in I])
    gm = np.array([gammas_at([-2.0], [6.0], i)[0] for i in I])
    np.testing.assert_array_equal(gp, gm)
    assert np.all(np.diff(gp) < 0) and np.all(gp < 1.0)
    # larger ions are screened less
    assert gammas_at([2.0], [9.0], 0.05)[0] > gammas_at([2.0], [3.0], 0.05)[0]


def test_activity_newton_reaches_the_lag_fixed_point():
    """the NEWTON algorithm (reaction.F90:4403-4551) converges I so that one
    more LAG pass at its output changes the coefficients by O(1e-5)"""
    wl = workloads.calcite_workload(8)
    opt_n = SolverOptions(act_coef_update_algorithm=abi.ACT_COEF_ALGORITHM_NEWTON)
    for cell in range(4):
        cc = O.CellCtx(wl.net, opt_n).gather(wl.state, cell)
        cc.activity()
        assert cc.c.activity_error == 0
        n = wl.net.naqcomp
        g_newton = cc.get("pri_act_coef", n)
        lag = O.CellCtx(wl.net, LAG)
        lag.c = cc.c
        cc.total()                           # sec molal at the new coefficients
        lag.activity()
        np.testing.assert_allclose(lag.get("pri_act_coef", n), g_newton, rtol=1e-5)


@pytest.fixture(scope="module")
def rand15():
    return workloads.random_network_workload(ncell=64)


def test_mass_action_in_log10_form(rand15):
    """every complex satisfies log10(g_k m_k) = sum nu log10(a_j) - log K_k"""
    net = rand15.net
    S = net.dense_stoichiometry()
    for cell in range(8):
        cc = O.CellCtx(net, LAG).gather(rand15.state, cell)
        cc.total()
        m = cc.get("pri_molal", net.naqcomp)
        g = cc.get("pri_act_coef", net.naqcomp)
        mk = cc.get("sec_molal", net.neqcplx)
        gk = cc.get("sec_act_coef", net.neqcplx)
        lhs = np.log10(gk * mk)
        rhs = S @ np.log10(g * m) - net.eqcplx_logK
        np.testing.assert_allclose(lhs, rhs, atol=1e-9)
        # totals: free + complexed, mol/kg -> mol/L
        T = (m + S.T @ mk) * cc.c.den_kg * 1e-3
        np.testing.assert_allclose(cc.get("total", net.naqcomp), T, rtol=1e-13)


def test_dtotal_against_central_differences(rand15):
    net = rand15.net
    n = net.naqcomp
    for cell in range(4):
        cc = O.CellCtx(net, LAG).gather(rand15.state, cell)
        cc.total()
        D = cc.dtotal()
        m0 = cc.get("pri_molal", n)
        fd = np.zeros((n, n))
        for j in range(n):
            h = 1e-6 * m0[j]
            for sgn in (1, -1):
                m = m0.copy()
                m[j] += sgn * h
                cc.set("pri_molal", m)
                cc.total()
                fd[:, j] += sgn * cc.get("total", n) / (2 * h)
        cc.set("pri_molal", m0)
        scale = np.abs(D).max(axis=1, keepdims=True)
        np.testing.assert_allclose(D / scale, fd / scale, atol=1e-7)


def calcite_cell(cell=0):
    wl = workloads.calcite_workload(4)
    return wl, O.CellCtx(wl.net, LAG).gather(wl.state, cell)


def test_tst_rate_zero_at_equilibrium_and_signs():
    wl, cc = calcite_cell()
    net = wl.net
    cc.total()
    # log10 Q of calcite = -H+ + Ca++ + HCO3- in activities
    a = cc.get("pri_molal", 3) * cc.get("pri_act_coef", 3)
    logQ = float(np.dot(net.dense_mineral_stoichiometry()[0], np.log10(a)))
    for shift, sign in ((0.0, 0), (0.5, -1), (-0.5, 1)):
        # K above Q: undersaturated, dissolution (rate < 0); K below Q: growth
        cc.c.kinmnrl_logK[0] = logQ + shift
        res, jac = cc.reaction(3600.0)
        rate = cc.get("mnrl_rate", 1)[0]
        if sign == 0:
            assert abs(rate) < 1e-9 * net.kinmnrl_rate_constant[0] * cc.get("mnrl_area", 1)[0]
        else:
            assert np.sign(rate) == sign
            # Res_i += nu_i Im V (reaction_mineral.F90:1000-1010)
            nu = net.dense_mineral_stoichiometry()[0]
            np.testing.assert_allclose(res, nu * rate * cc.c.volume, rtol=1e-12)
            vf0 = cc.get("mnrl_volfrac", 1)[0]
            cc.update_kinetic_state(100.0)
            assert np.sign(cc.get("mnrl_volfrac", 1)[0] - vf0) == sign
            cc.set("mnrl_volfrac", [vf0])


def test_reaction_jacobian_vs_reference_numerical_derivative(rand15):
    """analytic kinetic-mineral Jacobian vs RReactionDerivative's forward
    differences (relative perturbation 1e-5, reaction.F90:41).  The
    reference multiplies dQK/dm by rho/1000 (reaction_mineral.F90:945), so
    its analytic mineral block is the true derivative times rho/1000; the
    oracle keeps that factor, and so must this check."""
    net = rand15.net
    checked = 0
    for cell in range(32):
        cc = O.CellCtx(net, LAG).gather(rand15.state, cell)
        if not np.any(cc.get("mnrl_area", net.nkinmnrl) > 0):
            continue
        cc.total()
        _, Ja = cc.reaction(3600.0)
        _, Jn = cc.reaction_numerical(3600.0)
        Jn = Jn * cc.c.den_kg * 1e-3
        scale = max(np.abs(Ja).max(), 1e-300)
        np.testing.assert_allclose(Ja / scale, Jn / scale, atol=2e-4)
        checked += 1
    assert checked >= 8


def test_rreact_system_jacobian_central_differences(rand15):
    """accumulation + totals part of the RReact system (RTAccumulation-
    Derivative with dtotal) against central differences; the kinetic block,
    which carries the reference's rho/1000 factor, is checked above"""
    net = rand15.net
    n = net.ncomp
    dt = 3600.0
    for cell in range(4):
        cc = O.CellCtx(net, LAG).gather(rand15.state, cell)
        cc.set("mnrl_area", np.zeros(net.nkinmnrl))
        cc.total()
        fixed = np.zeros(n)
        res0, J = cc.rreact_system(fixed, dt)
        m0 = cc.get("pri_molal", n)
        fd = np.zeros((n, n))
        for j in range(n):
            h = 1e-6 * m0[j]
            for sgn in (1, -1):
                m = m0.copy()
                m[j] += sgn * h
                cc.set("pri_molal", m)
                cc.total()
                r, _ = cc.rreact_system(fixed, dt, derivative=False)
                fd[:, j] += sgn * r / (2 * h)
        cc.set("pri_molal", m0)
        scale = np.abs(J).max(axis=1, keepdims=True)
        np.testing.assert_allclose(J / scale, fd / scale, atol=1e-6)


def test_rstep_conserves_calcium_and_carbon():
    """calcite dissolved or precipitated = change of the aqueous Ca and HCO3
    totals: psv (T1 - T0) = -vol (VF1 - VF0) / V_m"""
    wl = workloads.calcite_workload(200)
    st = wl.state.copy()
    rec = O.Records(st.ncell)
    assert O.step_batch(wl.net, wl.opt, st, wl.dt, 0, rec) == 0
    s0, s1 = wl.state.arrays, st.arrays
    psv = s0["porosity"][0] * s0["saturation"][0] * 1000.0 * s0["volume"][0]
    dmin = -(s1["mnrl_volfrac"][0] - s0["mnrl_volfrac"][0]) * s0["volume"][0] / \
        wl.net.kinmnrl_molar_vol[0]
    for sp in (1, 2):                       # HCO3-, Ca++
        dtot = psv * (s1["total"][sp] - s0["total"][sp])
        np.testing.assert_allclose(dtot, dmin, rtol=1e-7, atol=1e-13 * psv.max())
    # the oracle's own convergence: final relative residual per cell
    assert np.all(O.record_fields(rec, "final_rel_residual") < 1e-9)


def test_rstep_cuts_then_fails_and_restores():
    """with one Newton iteration allowed no sub-step converges: RStep halves
    dt maximum_reaction_cuts times, then reports the failure with the
    totals of the failed attempts restored (reaction.F90:3660-3671, 4041)"""
    wl = workloads.calcite_workload(4)
    opt = SolverOptions(**{**wl.opt.__dict__, "maximum_reaction_iterations": 1})
    cc = O.CellCtx(wl.net, opt).gather(wl.state, 0)
    t0 = cc.get("total", 3)
    err, nsub, nit = cc.rstep(wl.dt)
    assert err == 1
    # the failure return skips the restore of clamped totals (:3660-3671
    # return before :3727-3736), so those stay at 1e-40
    expect = np.where(t0 <= 1e-40, 1e-40, t0)
    np.testing.assert_array_equal(cc.get("total", 3), expect)


def test_rreact_call_bound_truncates_the_sub_step_loop():
    """options.maximum_rreact_calls stops RStep's sub-step loop after that
    many RReact calls (reaction.F90:3721-3725 has no bound of its own): a
    bound only truncates the loop, so under every bound a cell either fails
    or ends bit-equal to the unbounded run, the failing cells shrink as the
    bound grows, and a cell that took more sub-steps than the bound (one
    RReact call each, at least) fails"""
    wl = workloads.random_network_workload(ncell=512)
    n = wl.state.ncell
    opt = SolverOptions(**{**wl.opt.__dict__, "maximum_reaction_iterations": 5})
    free = wl.state.copy()
    O.step_batch(wl.net, opt, free, wl.dt, 0)
    ok = free["status"][0, :n] == 0
    nsub = free["num_sub_steps"][0, :n]
    # both kinds present: cells done in one call and cells that cut dt
    assert (ok & (nsub == 1)).sum() > 50 and (ok & (nsub > 2)).sum() > 10
    fields = ("pri_molal", "total", "sec_molal", "mnrl_volfrac", "mnrl_rate", "num_iterations",
              "num_sub_steps")
    prev_fail = None
    for bound in (1, 2, 3, 5, 8, 10000):
        got = wl.state.copy()
        O.step_batch(wl.net, SolverOptions(**{**opt.__dict__, "maximum_rreact_calls": bound}),
                     got, wl.dt, 0)
        st = got["status"][0, :n]
        same = st == 0
        assert np.all(st[~ok] != 0), bound
        for f in fields:
            np.testing.assert_array_equal(got[f][:, :n][:, same], free[f][:, :n][:, same],
                                          err_msg="%s bound %d" % (f, bound))
        assert np.all(st[ok & (nsub > bound)] != 0), bound
        if bound == 1:
            assert np.all(st[ok & (nsub == 1)] == 0)
        fail = st != 0
        if prev_fail is not None:
            assert np.all(fail <= prev_fail), bound
        prev_fail = fail
    # a bound above every cell's call count is the unbounded run
    assert np.array_equal(prev_fail, ~ok)


def test_rstep_small_totals_clamped_and_restored():
    """totals <= 1e-40 (negative H+ totals here) are replaced by 1e-40 for
    the solve and restored after it (reaction.F90:3617-3631, 3727-3736)"""
    wl = workloads.calcite_workload(64)
    neg = np.where(wl.state.arrays["total"][0] <= 1e-40)[0]
    assert neg.size > 0
    st = wl.state.copy()
    assert O.step_batch(wl.net, wl.opt, st, wl.dt, 0) == 0
    np.testing.assert_array_equal(st.arrays["total"][0, neg], wl.state.arrays["total"][0, neg])
