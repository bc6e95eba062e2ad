"""The oracle's dense solver (LUDecomposition1 / LUBackSubstitution,
utility.F90:597-735; RSolve, reaction.F90:5457-5516) against numpy's LAPACK
solve and hand cases."""
import numpy as np
import pytest

import oracle_lib as O


@pytest.mark.parametrize("n", [1, 2, 3, 7, 15, 32, 64])
def test_lu_matches_numpy(n):
    rng = np.random.default_rng(100 + n)
    a = rng.standard_normal((n, n)) + n * np.eye(n)
    b = rng.standard_normal(n)
    x = O.lu_solve(a, b)
    np.testing.assert_allclose(x, np.linalg.solve(a, b), rtol=1e-11, atol=1e-13)


@pytest.mark.parametrize("n", [4, 15, 40])
def test_lu_needs_pivoting(n):
    """zero and tiny leading pivots: only a pivoting elimination gets these"""
    rng = np.random.default_rng(7 * n)
    a = rng.standard_normal((n, n))
    a[0, 0] = 0.0
    a[1, 1] = 1e-14
    b = rng.standard_normal(n)
    x = O.lu_solve(a, b)
    np.testing.assert_allclose(a @ x, b, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(x, np.linalg.solve(a, b), rtol=1e-8, atol=1e-10)


def test_lu_permutation_and_scaling():
    p = np.eye(5)[[3, 0, 4, 1, 2]]
    d = np.diag([1e-8, 1.0, 1e8, 3.0, 1e-3])
    a = p @ d
    b = np.arange(1.0, 6.0)
    np.testing.assert_allclose(O.lu_solve(a, b), np.linalg.solve(a, b), rtol=1e-13)


def test_lu_zero_row_is_an_error():
    a = np.array([[1.0, 2.0, 3.0], [0.0, 0.0, 0.0], [4.0, 5.0, 6.0]])
    assert O.lu_solve(a, np.ones(3)) is None


def test_lu_zero_pivot_replaced():
    """a singular matrix without a zero row: the zero pivot becomes 1e-20
    (utility.F90:663) and the solve returns huge but finite values"""
    a = np.array([[1.0, 1.0], [1.0, 1.0]])
    x = O.lu_solve(a, np.array([1.0, 2.0]))
    assert x is not None and np.all(np.isfinite(x)) and np.max(np.abs(x)) > 1e15


def _rsolve(res, J, conc, log_form):
    import ctypes as C
    n = len(res)
    r = np.ascontiguousarray(res, dtype=np.float64).copy()
    j = np.ascontiguousarray(J, dtype=np.float64).copy()
    c = np.ascontiguousarray(conc, dtype=np.float64).copy()
    u = np.zeros(n)
    err = O.lib().orc_rsolve(r.ctypes.data, j.ctypes.data, c.ctypes.data, u.ctypes.data, n,
                             1 if log_form else 0)
    return err, u


@pytest.mark.parametrize("log_form", [False, True])
def test_rsolve_row_scaling_and_log_form(log_form):
    """RSolve scales rows by 1/max(1, max|J_i.|) (no effect on the solution)
    and, in the log formulation, solves for d ln c: (J diag(c)) u = r"""
    rng = np.random.default_rng(11)
    n = 6
    J = rng.standard_normal((n, n)) * 10.0 ** rng.uniform(-3, 3, size=(n, 1)) + 5 * np.eye(n)
    r = rng.standard_normal(n)
    c = 10.0 ** rng.uniform(-8, -2, size=n)
    err, u = _rsolve(r, J, c, log_form)
    assert err == 0
    A = J * c[None, :] if log_form else J
    np.testing.assert_allclose(u, np.linalg.solve(A, r), rtol=1e-9)
