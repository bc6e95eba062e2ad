"""Pins of oracle functions the reference's golds do not reach, each against
what the mathematics fixes independently of the oracle's code:

* RImmobileDecay (reaction_immobile.F90:244-294): one backward-Euler step of
  first-order decay has the closed form x = x0 / (1 + k dt); decay reactions
  on one species add their rate constants; undecayed immobile species and the
  aqueous solution are untouched;
* log K(T) (ReactionInterpolateLogK, reaction_aux.F90:1285-1312, with the
  least-squares fit of reaction_database.F90): the interpolation through the
  fitted coefficients reproduces calcite.dat's tabulated values at 0-300 C to
  the fit's residual, and a non-isothermal RStep at T equals the isothermal
  RStep whose log K are the interpolated values at T;
* RKineticMineral options (reaction_mineral.F90:729-908): every option at its
  neutral value (Temkin constant 1, scale factor 1, affinity power 1, no
  limiter / threshold / irreversibility, Ea = 0) reduces to plain TST; the
  rate is zero at QK = 1 under every option; each option's rate against its
  closed form written out here (Q^(1/sigma), |1 - Q|^p, the limiter quotient,
  the precipitation cut-offs, the Arrhenius factor); the analytic Jacobian of
  each option against the reference's numerical derivative scheme
  (RReactionDerivative, reaction.F90:4134-4239).
"""
import dataclasses
import math

import numpy as np
import pytest

import oracle_lib as O
from pfchem import abi, database, workloads
from pin_util import decay_workload

GOLDEN = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden")


# ---------------------------------------------------------------------------
# immobile decay
# ---------------------------------------------------------------------------
def test_decay_closed_form():
    wl, base = decay_workload()
    ref = wl.state.copy()
    assert O.step_batch(wl.net, wl.opt, ref, wl.dt, 0) == 0
    aq = base.state.copy()
    assert O.step_batch(base.net, base.opt, aq, base.dt, 0) == 0
    one = ref["num_sub_steps"][0] == 1
    assert one.mean() > 0.9
    k = wl.net.decay_rate_constant.sum()
    x0 = wl.state["immobile"][0]
    np.testing.assert_allclose(ref["immobile"][0][one], (x0 / (1.0 + k * wl.dt))[one], rtol=1e-12)
    np.testing.assert_allclose(ref["immobile"][1], wl.state["immobile"][1], rtol=0, atol=0)
    for f in ("pri_molal", "total", "sec_molal", "mnrl_volfrac"):
        np.testing.assert_allclose(ref[f][:, one], aq[f][:, one], rtol=1e-11)


def test_decay_rate_and_jacobian():
    """RImmobileDecay adds k V x to the residual and k V to the diagonal"""
    wl, _ = decay_workload(4)
    cc = O.CellCtx(wl.net, wl.opt).gather(wl.state, 1)
    cc.c.volume = 2.5
    cc.total()
    res, jac = cc.reaction(wl.dt)
    k = wl.net.decay_rate_constant.sum()
    x = cc.get("immobile", 2)
    assert res[3] == pytest.approx(k * 2.5 * x[0], rel=1e-15)
    assert res[4] == 0.0
    assert jac[3, 3] == pytest.approx(k * 2.5, rel=1e-15)
    assert jac[4, 4] == 0.0


# ---------------------------------------------------------------------------
# log K(T)
# ---------------------------------------------------------------------------
def _calcite_db():
    return database.read_database(__import__("os").path.join(GOLDEN, "calcite.dat"))


def test_logk_interpolation_reproduces_tables():
    db = _calcite_db()
    temps = np.asarray(db.temperatures, dtype=np.float64)
    worst = 0.0
    for name in ("OH-", "CO3--", "CO2(aq)", "CaHCO3+", "CaCO3(aq)", "Calcite"):
        rx = db.minerals[name] if name in db.minerals else database._species(db, name)
        lk = np.asarray(rx.logK, dtype=np.float64)
        assert lk.size == temps.size
        coef = database.fit_logk_coef(temps, lk)
        # the oracle's interpolation, through a one-complex network
        net = workloads.calcite_network()
        net = dataclasses.replace(net, eqcplx_logKcoef=np.tile(coef, (net.neqcplx, 1)),
                                  kinmnrl_logKcoef=np.tile(coef, (1, 1)))
        opt = dataclasses.replace(workloads.bench_options(), isothermal=0)
        cc = O.CellCtx(net, opt)
        for t, want in zip(temps, lk):
            cc.c.temp = float(t)
            O.lib().orc_update_temp_dependent_coefs(O.C.byref(cc.n), O.C.byref(cc.o),
                                                    O.C.byref(cc.c))
            got = cc.get("eqcplx_logK", 1)[0]
            assert cc.get("kinmnrl_logK", 1)[0] == got
            worst = max(worst, abs(got - want))
            assert abs(got - want) < 0.02, (name, t, got, want)
    # the five-term fit leaves residuals of a few 1e-3 on these tables; a
    # dropped term or a wrong sign moves some point by far more
    assert worst > 1e-6


def test_nonisothermal_step_equals_isothermal_at_interpolated_logk():
    wl = workloads.calcite_workload(256)
    db = _calcite_db()
    temps = np.asarray(db.temperatures, dtype=np.float64)
    names = ["OH-", "CO3--", "CO2(aq)", "CaHCO3+", "CaCO3(aq)"]
    ccoef = np.array([database.fit_logk_coef(temps, database._species(db, n).logK)
                      for n in names])
    mcoef = np.array([database.fit_logk_coef(temps, db.minerals["Calcite"].logK)])
    T = 60.0
    net_t = dataclasses.replace(wl.net, eqcplx_logKcoef=ccoef, kinmnrl_logKcoef=mcoef)
    opt_t = dataclasses.replace(wl.opt, isothermal=0)
    lk_c = np.array([database.interpolate_logk(c, T) for c in ccoef])
    lk_m = np.array([database.interpolate_logk(c, T) for c in mcoef])
    net_i = dataclasses.replace(wl.net, eqcplx_logK=lk_c, kinmnrl_logK=lk_m)
    st = wl.state.copy()
    st["temperature"][0, :] = T
    a = st.copy()
    b = st.copy()
    O.step_batch(net_t, opt_t, a, wl.dt, 0)
    O.step_batch(net_i, wl.opt, b, wl.dt, 0)
    np.testing.assert_array_equal(a["status"], b["status"])
    for f in ("pri_molal", "total", "sec_molal", "mnrl_volfrac", "mnrl_rate"):
        np.testing.assert_allclose(a[f], b[f], rtol=1e-12, atol=1e-300)


# ---------------------------------------------------------------------------
# kinetic mineral options
# ---------------------------------------------------------------------------
def _mineral_cell(net, opt, Q, vf=0.1, area=1.0, temp=25.0):
    """oracle cell of the calcite network whose calcite saturation index is
    exactly Q (the mineral's log K is set from the cell's ion activity
    product)"""
    wl = workloads.calcite_workload(4)
    cc0 = O.CellCtx(wl.net, wl.opt).gather(wl.state, 2)
    lnact = np.log(cc0.get("pri_molal", 3)) + np.log(cc0.get("pri_act_coef", 3))
    lnIAP = -lnact[0] + lnact[1] + lnact[2]
    logK = (lnIAP - math.log(Q)) / 2.30258509299
    net = dataclasses.replace(net, kinmnrl_logK=np.array([logK]))
    cc = O.CellCtx(net, opt).gather(wl.state, 2)
    cc.c.temp = temp
    cc.set("mnrl_volfrac", [vf])
    cc.set("mnrl_area", [area])
    cc.total()
    return cc


def _rate(net, Q, **kw):
    opt = workloads.bench_options()
    cc = _mineral_cell(net, opt, Q, **kw)
    res, jac = cc.reaction(3600.0)
    return cc.get("mnrl_rate", 1)[0], res, jac, cc


BASE = workloads.calcite_network()
K0 = float(BASE.kinmnrl_rate_constant[0])


def with_opts(**kw):
    arrays = {k: (np.array([v], dtype=np.int32) if k == "kinmnrl_irreversible"
                  else np.array([float(v)])) for k, v in kw.items()}
    return dataclasses.replace(BASE, **arrays)


@pytest.mark.parametrize("Q", [0.2, 0.97, 1.6, 5.0])
def test_neutral_options_reduce_to_tst(Q):
    r0, res0, jac0, _ = _rate(BASE, Q)
    net = with_opts(kinmnrl_Temkin_const=1.0, kinmnrl_min_scale_factor=1.0,
                    kinmnrl_affinity_threshold=0.0, kinmnrl_rate_limiter=0.0,
                    kinmnrl_irreversible=0, kinmnrl_activation_energy=0.0)
    r1, res1, jac1, _ = _rate(net, Q)
    assert r1 == pytest.approx(r0, rel=1e-13)
    np.testing.assert_allclose(res1, res0, rtol=1e-13)
    np.testing.assert_allclose(jac1, jac0, rtol=1e-12, atol=1e-30)
    # plain TST closed form: Im = -A k (1 - Q)
    assert r0 == pytest.approx(-1.0 * K0 * (1.0 - Q), rel=1e-12)
    # affinity power 1: the same rate; its Jacobian, dIm/dQK = -Im p / |1 - Q|
    # (reaction_mineral.F90:918-919), carries the sign of the affinity, so
    # under precipitation (Q > 1) it is the negative of the TST derivative
    r2, res2, jac2, _ = _rate(with_opts(kinmnrl_affinity_power=1.0), Q)
    assert r2 == pytest.approx(r0, rel=1e-13)
    np.testing.assert_allclose(jac2, jac0 * (1.0 if Q < 1.0 else -1.0), rtol=1e-12, atol=1e-30)


OPTION_SETS = [
    dict(kinmnrl_Temkin_const=2.0),
    dict(kinmnrl_min_scale_factor=3.0),
    dict(kinmnrl_affinity_power=2.0),
    dict(kinmnrl_rate_limiter=0.5),
    dict(kinmnrl_affinity_threshold=2.0),
    dict(kinmnrl_irreversible=1),
    dict(kinmnrl_activation_energy=4.0e4),
    dict(kinmnrl_Temkin_const=2.0, kinmnrl_min_scale_factor=1.5, kinmnrl_affinity_power=1.7,
         kinmnrl_rate_limiter=3.0),
]


@pytest.mark.parametrize("opts", OPTION_SETS, ids=lambda d: "+".join(sorted(d)))
def test_option_rate_zero_at_equilibrium(opts):
    r, res, _, _ = _rate(with_opts(**opts), 1.0, temp=40.0)
    assert abs(r) <= 1e-15 * K0


def _closed_form(opts, Q, area=1.0, temp=25.0):
    """RKineticMineral's rate written from the reference's formulas"""
    tk = opts.get("kinmnrl_Temkin_const")
    sc = opts.get("kinmnrl_min_scale_factor")
    sig = (sc or 1.0) * (tk or 1.0)
    aff = 1.0 - Q ** (1.0 / sig)
    sgn = math.copysign(1.0, aff)
    if opts.get("kinmnrl_irreversible", 0) == 1 and sgn < 0:
        return 0.0
    th = opts.get("kinmnrl_affinity_threshold", 0.0)
    if th > 0 and sgn < 0 and Q < th:
        return 0.0
    lim = opts.get("kinmnrl_rate_limiter", 0.0)
    if lim > 0:
        aff = aff / (1.0 + (1.0 - aff) / lim)
    Ea = opts.get("kinmnrl_activation_energy", 0.0)
    k = K0 * (math.exp(Ea / 8.31446 * (1.0 / 298.15 - 1.0 / (temp + 273.15))) if Ea > 0 else 1.0)
    p = opts.get("kinmnrl_affinity_power", 1.0)
    return -area / (sc or 1.0) * sgn * abs(aff) ** p * k


@pytest.mark.parametrize("opts", OPTION_SETS, ids=lambda d: "+".join(sorted(d)))
@pytest.mark.parametrize("Q", [0.3, 1.5, 4.0])
def test_option_rate_closed_form(opts, Q):
    r, _, _, _ = _rate(with_opts(**opts), Q, temp=40.0)
    want = _closed_form(opts, Q, temp=40.0)
    assert r == pytest.approx(want, rel=1e-12, abs=1e-300)


def _ref_dIm_factor(opts, Q):
    """d(rate)/dQK per unit area and k0 as reaction_mineral.F90:917-966
    assembles it (Im per unit area), times the limiter factor"""
    tk = opts.get("kinmnrl_Temkin_const")
    sc = opts.get("kinmnrl_min_scale_factor")
    sig = (sc or 1.0) * (tk or 1.0)
    aff = 1.0 - Q ** (1.0 / sig)
    lim = opts.get("kinmnrl_rate_limiter", 0.0)
    if lim > 0:
        aff = aff / (1.0 + (1.0 - aff) / lim)
    im = _closed_form(opts, Q) / K0
    p = opts.get("kinmnrl_affinity_power")
    d = -im * p / abs(aff) if p is not None else 1.0 / (sc or 1.0)
    if tk is not None or sc is not None:
        d = d * (1.0 / sig) / Q * (1.0 - aff)
    if lim > 0:
        den = 1.0 + (1.0 - aff) / lim
        d = d * (1.0 + Q / lim / den) / den
    return d


@pytest.mark.parametrize("opts", OPTION_SETS, ids=lambda d: "+".join(sorted(d)))
@pytest.mark.parametrize("Q", [0.3, 4.0])
def test_option_jacobian(opts, Q):
    """analytic d(res)/d(m) of each option.  Where the reference's analytic
    derivative is the exact one (Temkin constant, scale factor, threshold,
    irreversibility, Arrhenius, affinity power under dissolution) it must
    match the reference's forward-difference scheme (relative perturbation
    1e-5); the rate limiter's factor (1 + QK/lim/den)/den and the affinity
    power's -Im p/|aff| under precipitation (reaction_mineral.F90:918-919,
    957-965) are not exact derivatives, and there the Jacobian must be the
    reference's expression: nu_i dIm/dQK(ref) nu_j QK/m_j rho/1000, against
    which the exact derivative is shown to differ."""
    net = with_opts(**opts)
    opt = workloads.bench_options()
    cc = _mineral_cell(net, opt, Q, temp=40.0)
    res, jac = cc.reaction(3600.0)
    _, jn = cc.reaction_numerical(3600.0)
    # the analytic mineral block carries rho/1000 (reaction_mineral.F90:945):
    # d res / d(molality) there, d res / d(molarity) in the numerical one
    rho = cc.c.den_kg * 1e-3
    scale = np.abs(jac).max()
    if scale == 0.0:
        assert np.abs(jn).max() == 0.0
        return
    inexact = ("kinmnrl_rate_limiter" in opts or
               ("kinmnrl_affinity_power" in opts and Q > 1.0))
    if not inexact:
        np.testing.assert_allclose(jac[:3, :3], jn[:3, :3] * rho, rtol=2e-4, atol=2e-4 * scale)
        return
    nu = np.array([-1.0, 1.0, 1.0])  # calcite: -H+ + HCO3- + Ca++ (ids 0, 2, 1)
    nu_id = np.zeros(3)
    nu_id[[0, 2, 1]] = nu
    m = cc.get("pri_molal", 3)
    Ea = opts.get("kinmnrl_activation_energy", 0.0)
    kfac = math.exp(Ea / 8.31446 * (1.0 / 298.15 - 1.0 / 313.15)) if Ea > 0 else 1.0
    dIm = _ref_dIm_factor(opts, Q) * K0 * kfac * cc.c.volume
    if "kinmnrl_affinity_power" not in opts:
        dIm = dIm * 1.0  # -Im_const k with Im_const = -area
    want = np.outer(nu_id, nu_id * Q / m) * dIm * rho
    np.testing.assert_allclose(jac[:3, :3], want, rtol=1e-11, atol=1e-11 * scale)
    assert np.abs(jac[:3, :3] - jn[:3, :3] * rho).max() > 1e-3 * scale


# ---------------------------------------------------------------------------
# rate prefactors (reaction_mineral.F90:820-870, derivative :964-1033)
import pin_util as PU
from pfchem.network import Prefactor


def test_prefactor_without_species_terms_is_the_rate_constant():
    """one prefactor with alpha = beta = 0 and no attenuation is exactly the
    rate constant: the whole step equals the plain TST one"""
    wl = workloads.calcite_workload(256)
    net = wl.net
    k = float(net.kinmnrl_rate_constant[0])
    net_p = dataclasses.replace(net, kinmnrl_prefactors=[[Prefactor(k, [(0, 0.0, 0.0, 0.0)])]],
                                kinmnrl_rate_constant=np.array([123.0]))  # unused with prefactors
    a, b = wl.state.copy(), wl.state.copy()
    assert O.step_batch(net, wl.opt, a, wl.dt, 0) == 0
    assert O.step_batch(net_p, wl.opt, b, wl.dt, 0) == 0
    for f in ("pri_molal", "sec_molal", "mnrl_volfrac", "mnrl_rate"):
        np.testing.assert_allclose(b[f], a[f], rtol=1e-14, atol=0)


def test_prefactor_rate_closed_form():
    """Im = -A (1 - QK) sum_p k_p Arr_p prod_s a_s^alpha / (1 + atten a_s^beta),
    Arr = exp(Ea/R (1/298.15 - 1/T)), evaluated in numpy at each cell's state"""
    wl = PU.prefactor_workload(64)
    net = wl.net
    pf = net.kinmnrl_prefactors[0]
    nu = net.dense_mineral_stoichiometry()[0]
    checked = 0
    for cell in range(64):
        cc = O.CellCtx(net, wl.opt).gather(wl.state, cell)
        cc.total()
        cc.reaction(3600.0)
        rate = cc.get("mnrl_rate", 1)[0]
        a = cc.get("pri_molal", 3) * cc.get("pri_act_coef", 3)
        lnQK = -net.kinmnrl_logK[0] * 2.30258509299 + float(nu @ np.log(a))
        QK = math.exp(lnQK)
        if not (cc.get("mnrl_volfrac", 1)[0] > 0 or QK > 1):
            continue
        T = cc.c.temp + 273.15
        kr = 0.0
        for p in pf:
            f = 1.0
            for sid, al, be, at in p.species:
                f *= a[sid] ** al / (1.0 + at * a[sid] ** be)
            arr = math.exp(p.activation_energy / 8.31446 * (1 / 298.15 - 1 / T))
            kr += p.rate * arr * f
        want = -cc.get("mnrl_area", 1)[0] * (1.0 - QK) * kr
        assert rate == pytest.approx(want, rel=1e-12), cell
        checked += 1
    assert checked > 32


@pytest.mark.parametrize("alpha_slip", [1.0, 1.1], ids=["as-is", "planted-alpha-slip"])
def test_prefactor_jacobian_vs_numerical_derivative(alpha_slip):
    """analytic mineral Jacobian with prefactor terms against RReactionDerivative's
    forward differences (reaction.F90:41), at unit density so the rho/1000 of
    the QK block is 1; a 10% slip in one ALPHA in the analytic evaluation only
    must be seen"""
    wl = PU.prefactor_workload(16)
    net = wl.net
    pf = net.kinmnrl_prefactors[0]
    slipped = [Prefactor(pf[0].rate, [(pf[0].species[0][0], pf[0].species[0][1] * alpha_slip,
                                       pf[0].species[0][2], pf[0].species[0][3])] +
                         pf[0].species[1:]), pf[1]]
    net_a = dataclasses.replace(net, kinmnrl_prefactors=[slipped])
    worst = 0.0
    for cell in range(16):
        cc = O.CellCtx(net, wl.opt).gather(wl.state, cell)
        cc.c.den_kg = 1000.0
        cc.total()
        _, Jn = cc.reaction_numerical(3600.0)
        ca = O.CellCtx(net_a, wl.opt).gather(wl.state, cell)
        ca.c.den_kg = 1000.0
        ca.total()
        _, Ja = ca.reaction(3600.0)
        scale = max(np.abs(Jn).max(), 1e-300)
        worst = max(worst, float(np.abs(Ja - Jn).max() / scale))
    if alpha_slip == 1.0:
        assert worst < 2e-4, worst
    else:
        assert worst > 1e-2, worst
