"""The transport oracle (oracle/oracle_transport.py, a restatement of the
OSRT transport system: RTCalculateTransportMatrix, RTCalculateRHS_t1,
TFluxCoef) against closed forms: one implicit upwind step of pure advection
through a column, mass conservation without boundaries, a uniform field
under pure diffusion."""
import numpy as np

import oracle_transport as OT
import transport_util as TU


def _column(n, q, cr_area=1.0, cb=1e-3):
    c = TU.box_case(n, 1, 1, 1, seed=1)
    c["vel"][:] = q
    c["cod"][:] = 0.0
    c["bc_cod"][:] = 0.0
    c["bc_vel"][:] = [q, -q]
    c["bc_total"][0, 0] = cb
    c["por"][:] = 0.25
    c["sat"][:] = 1.0
    c["total"][:] = 0.0
    return c


def test_advection_step_closed_form():
    """implicit upwind from c = 0 with inflow c_b: c_i = c_b (Cr/(1+Cr))^(i+1),
    Cr = q A dt / (phi s V)"""
    n, q = 50, 2e-5
    c = _column(n, q)
    got = TU.oracle_step(c)[0]
    cr = q * 1.0 * c["dt"] / (0.25 * 1.0 * 1.0)
    want = 1e-3 * (cr / (1.0 + cr)) ** np.arange(1, n + 1)
    np.testing.assert_allclose(got, want, rtol=1e-12)


def test_advection_closed_form_catches_downwinding():
    """the same column with the upwind choice reversed in the oracle's
    coefficients (flux taken from the downstream cell) misses the closed form"""
    n, q = 20, 2e-5
    c = _column(n, q)
    orig = OT.flux_coef

    def downwind(qq, cod, area):
        qq = np.asarray(qq, dtype=np.float64)
        t_up = np.where(qq > 0.0, cod, cod + qq) * area * 1000.0
        t_dn = np.where(qq > 0.0, -cod + qq, -cod) * area * 1000.0
        return t_up, t_dn
    OT.flux_coef = downwind
    try:
        got = TU.oracle_step(c)[0]
    finally:
        OT.flux_coef = orig
    cr = q * c["dt"] / 0.25
    want = 1e-3 * (cr / (1.0 + cr)) ** np.arange(1, n + 1)
    assert np.abs(got - want).max() > 1e-2 * want.max()


def test_conservation_without_boundaries():
    c = TU.box_case(6, 5, 4, 3, seed=2, dry_frac=0.05)
    c["bc_cell"] = c["bc_cell"][:0]
    for k in ("bc_area", "bc_vel", "bc_cod"):
        c[k] = c[k][:0]
    c["bc_total"] = c["bc_total"][:, :0]
    new = TU.oracle_step(c)
    psv = c["por"] * c["sat"] * 1000.0 * c["vol"]
    np.testing.assert_allclose((psv * new).sum(axis=1), (psv * c["total"]).sum(axis=1),
                               rtol=1e-12)


def test_uniform_field_under_diffusion():
    c = TU.box_case(5, 4, 3, 2, seed=4)
    c["vel"][:] = 0.0
    c["bc_cell"] = c["bc_cell"][:0]
    for k in ("bc_area", "bc_vel", "bc_cod"):
        c[k] = c[k][:0]
    c["bc_total"] = c["bc_total"][:, :0]
    c["total"][:] = 3.7e-4
    np.testing.assert_allclose(TU.oracle_step(c), 3.7e-4, rtol=1e-13)
