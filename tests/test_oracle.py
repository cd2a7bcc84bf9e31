"""Pins the CPU oracle (oracle/solve.py) before it is trusted as the checker.

Each check is chosen so that a plausible restatement mistake (swapped pre/post
smoothing, a dropped coarse correction, W/F recursion order, the stopping
rule's comparison, an off-by-one iteration count) fails it:
  * the reference's own known answers (Jacobi, cycle complexity, coarse solvers);
  * textbook closed forms: the multigrid error propagator
      E_l = S_post (I - P (I - E_{l+1}^g) A_{l+1}^+ R A) S_pre
    built from dense matrices, independent of the oracle's recursion;
  * scipy.sparse.linalg.cg for the Krylov loop;
  * convergence to scipy.sparse.linalg.spsolve.
"""

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import rel, sa_hierarchy
from oracle.solve import (OracleMG, approximate_spectral_radius, cycle_complexity, inverse_diagonal, jacobi,
                          pcg, relres)
from amjax_amd.gallery import poisson


def _tridiag(N):
    return sp.diags([2 * np.ones(N), -np.ones(N - 1), -np.ones(N - 1)], [0, -1, 1], shape=(N, N), format="csr")


def test_jacobi_known_answers():
    """tests/relaxation/test_relaxation.py:42-80."""
    A = _tridiag(1)
    D = inverse_diagonal(A)
    np.testing.assert_almost_equal(jacobi(A, np.arange(1.0), np.zeros(1), D), [0])
    A = _tridiag(3)
    D = inverse_diagonal(A)
    np.testing.assert_almost_equal(jacobi(A, np.zeros(3), np.arange(3.0), D), [0.0, 0.5, 1.0])
    np.testing.assert_almost_equal(jacobi(A, np.arange(3.0), np.zeros(3), D), [0.5, 1.0, 0.5])
    A1 = _tridiag(1)
    np.testing.assert_almost_equal(jacobi(A1, np.arange(1.0), np.array([10.0]), inverse_diagonal(A1)), [5])
    np.testing.assert_almost_equal(jacobi(A, np.arange(3.0), np.array([10.0, 20, 30]), D), [5.5, 11.0, 15.5])
    x = jacobi(A, np.arange(3.0), np.array([10.0, 20, 30]), D, omega=1.0 / 3.0)
    np.testing.assert_almost_equal(x, 2.0 / 3.0 * np.arange(3.0) + 1.0 / 3.0 * np.array([5.5, 11.0, 15.5]))
    # dtype is preserved (test_relaxation.py:15-38)
    A32 = poisson((4,)).astype(np.float32)
    assert jacobi(A32, np.zeros(4, np.float32), np.arange(4, dtype=np.float32), inverse_diagonal(A32)).dtype == np.float32


def test_inverse_diagonal_duplicates_and_zeros():
    A = sp.coo_matrix((np.array([1.0, 1.0, 4.0, 3.0]), (np.array([0, 0, 1, 2]), np.array([0, 0, 1, 0]))),
                      shape=(3, 3))
    np.testing.assert_array_equal(inverse_diagonal(A), [0.5, 0.25, 0.0])


def test_spectral_radius():
    """tests/relaxation/test_relaxation.py:82-101."""
    A = sp.diags(2.0 * np.ones(4)).tocsr()
    np.testing.assert_almost_equal(approximate_spectral_radius(A, inverse_diagonal(A)), 1.0, decimal=5)
    A = poisson((10,))
    rho = approximate_spectral_radius(A, inverse_diagonal(A))
    ref = spla.eigsh(sp.diags(1.0 / A.diagonal()) @ A, k=1, which="LM", return_eigenvectors=False)[0]
    assert abs(rho - ref) / ref < 0.1


def _power_closed_form(M, steps, seed=0):
    """||M^t v0|| / ||M^(t-1) v0|| with v0 = g / ||g||, g = default_rng(seed).standard_normal(n):
    the t-th estimate of the reference's power iteration, whose normalisations cancel
    (amjax/relaxation/relaxation.py:122-131), computed from dense matrix powers."""
    g = np.random.default_rng(seed).standard_normal(M.shape[0])
    v = g / np.linalg.norm(g)
    hi = np.linalg.matrix_power(M, steps) @ v
    lo = np.linalg.matrix_power(M, steps - 1) @ v
    return float(np.linalg.norm(hi) / np.linalg.norm(lo))


@pytest.mark.parametrize("which", ["oracle", "product"])
def test_spectral_radius_closed_form(which):
    """15 power steps from the seed-0 normal vector, by closed form.

    A diagonal D^-1 A = diag(lam) with slowly separating eigenvalues: the 14th,
    15th and 16th estimates differ by far more than the tolerance, so a wrong
    step count, a Rayleigh-quotient estimator or a different start vector fails.
    The same on a non-diagonal operator (1-D Poisson with a non-constant
    diagonal).  In float32 the start vector and iterates are rounded: the result
    must move off the float64 value by float32 rounding, not stay on it.
    """
    from amjax_amd.relaxation import relaxation as prod
    fn = approximate_spectral_radius if which == "oracle" else prod.approximate_spectral_radius
    n = 12
    lam = np.array([1.0, 0.97, 0.93, 0.9, 0.8, 0.7, 0.6, 0.5, 0.4, 0.3, 0.2, 0.1])
    a = np.linspace(1.0, 3.0, n)
    A = sp.diags(a).tocsr()
    dinv = lam / a
    r14, r15, r16 = (_power_closed_form(np.diag(lam), t) for t in (14, 15, 16))
    assert min(abs(r15 - r14), abs(r16 - r15)) / r15 > 1e-4
    # the Rayleigh quotient of the 15th iterate: another estimator of the same rho
    g = np.random.default_rng(0).standard_normal(n)
    v15 = lam ** 15 * g / np.linalg.norm(lam ** 15 * g)
    rq = float(v15 @ (lam * v15))
    assert abs(rq - r15) / r15 > 1e-4
    assert abs(fn(A, dinv) - r15) / r15 < 1e-13
    # non-diagonal M = D^-1 A with D^-1 not the inverse diagonal of A
    B = (poisson((n,)) + sp.diags(np.linspace(0.0, 1.0, n))).tocsr()
    d = 1.0 / np.linspace(2.0, 4.0, n)
    M = np.diag(d) @ B.toarray()
    assert abs(fn(B, d) - _power_closed_form(M, 15)) / _power_closed_form(M, 15) < 1e-13
    # float32: the computation runs in the dtype of D^-1
    r32 = fn(A.astype(np.float32), dinv.astype(np.float32))
    assert 1e-9 < abs(r32 - r15) / r15 < 1e-5


def test_solve_with_x0_meets_the_shifted_tolerance():
    """amjax/multilevel.py:385-393: b - A x0 is solved from zero and x0 added back.

    The returned x must satisfy ||b - A x|| < tol ||b - A x0|| (the loop's normb
    is that of the shifted right-hand side) with the iteration count of a plain
    solve of b - A x0; a missing or doubled + x0 fails the residual check.
    """
    h = sa_hierarchy((24, 24))
    spec = ("jacobi", {"iterations": 1, "withrho": True})
    om = OracleMG(h, spec, spec)
    A = om.levels[0].A
    rng = np.random.default_rng(5)
    b, x0 = rng.standard_normal(A.shape[0]), 3.0 * rng.standard_normal(A.shape[0])
    tol = 1e-9
    x, it = om.solve(b, x0=x0, tol=tol, maxiter=200, return_info=True)
    r0 = b - A @ x0
    assert np.linalg.norm(b - A @ x) < tol * np.linalg.norm(r0)
    e, it0 = om.solve(r0, tol=tol, maxiter=200, return_info=True)
    assert it == it0 and it > 1
    assert np.linalg.norm(b - A @ (x - x0)) > 1e3 * tol * np.linalg.norm(r0)
    assert np.linalg.norm(b - A @ (x + x0)) > 1e3 * tol * np.linalg.norm(r0)


def test_pcg_ru_oracle_reduces_to_pcg_and_converges():
    """The reliable-update CG, pinned three ways:
      * in float64 the updates replace the recurrence residual by a value equal
        to it in exact arithmetic, so with few (delta -> 0) or many (delta = 0.5)
        updates the iterates and the iteration count are those of ``pcg``, itself
        pinned to scipy's cg;
      * in float32 the true float64 residual meets tol (a fold that forgets to
        reset the correction, or a residual not replaced, fails this), with about
        log(tol) / log(delta) updates;
      * maxiter folds the last correction: maxiter = 3 equals three iterations.
    """
    from oracle.solve import pcg_ru
    h = sa_hierarchy((48, 48))
    spec = ("jacobi", {"iterations": 1, "withrho": True})
    om = OracleMG(h, spec, spec)
    A = om.levels[0].A
    M = om.aspreconditioner()
    b = np.random.default_rng(11).standard_normal(A.shape[0])
    x_ref, it_ref = pcg(A, b, M, tol=1e-10, maxiter=300, return_info=True)
    for delta in (1e-300, 0.5):
        x, it, ups = pcg_ru(A, b, M, tol=1e-10, delta=delta, maxiter=300, dtype=np.float64, return_info=True)
        assert it == it_ref
        assert rel(x, x_ref) < 1e-10
        assert (ups == 1) if delta < 1e-100 else (ups >= it // 2)
    om32 = OracleMG(h, spec, spec, dtype=np.float32)

    def M32(r):
        return np.asarray(om32.aspreconditioner()(np.asarray(r, np.float32)), np.float32)

    x, it, ups = pcg_ru(A, b, M32, tol=1e-10, delta=1e-2, maxiter=300, dtype=np.float32, return_info=True)
    assert relres(A, b, x) < 1e-10
    assert abs(it - it_ref) <= 2
    assert 4 <= ups <= 7
    assert rel(x, pcg(A, b, M, tol=1e-14, maxiter=300)) < 1e-8
    x3, it3, ups3 = pcg_ru(A, b, M, tol=1e-10, maxiter=3, dtype=np.float64, return_info=True)
    assert it3 == 3 and ups3 >= 1
    assert rel(x3, pcg(A, b, M, tol=0.0, maxiter=3)) < 1e-12


def test_cycle_complexity_closed_forms():
    """tests/test_multilevel.py:110-153: nnz per level 100, 25, 9, 4."""
    nnz = [100, 25, 9, 4]
    assert cycle_complexity(nnz[:1], "V") == 1.0
    for c in "VWF":
        assert cycle_complexity(nnz[:2], c) == 2.25
    assert cycle_complexity(nnz[:3], "V") == 2.59
    assert cycle_complexity(nnz[:3], "W") == 3.18
    assert cycle_complexity(nnz[:3], "F") == 3.18
    assert cycle_complexity(nnz[:4], "V") == 2.72
    assert cycle_complexity(nnz[:4], "W") == 3.88
    assert cycle_complexity(nnz[:4], "F") == 3.66
    with pytest.raises(NotImplementedError):
        cycle_complexity(nnz, "AMLI")


class _H:
    def __init__(self, levels):
        self.levels = levels


class _L:
    pass


@pytest.mark.parametrize("coarse", ["pinv", "lu", "qr", "jacobi"])
def test_coarse_solvers(coarse):
    """A x = b on the reference's three cases (tests/test_multilevel.py:16-47)."""
    for A in [sp.diags(np.arange(1, 5, dtype=float)).tocsr(), poisson((4,)), poisson((4, 4))]:
        lvl = _L()
        lvl.A = A
        om = OracleMG(_H([lvl]), coarse_solver=coarse,
                      coarse_solver_kwargs={"iterations": 500, "omega": 2 / 3} if coarse == "jacobi" else None)
        b = np.arange(A.shape[0], dtype=float)
        x = om.coarse_solve(om.levels[0].A, np.zeros_like(b), b)
        np.testing.assert_almost_equal(A @ x, b, decimal=3)


def _dense_levels(om):
    out = []
    for l in om.levels:
        d = {"A": l.A.toarray(), "Dinv": np.diag(l.Dinv)}
        if l.P is not None:
            d["P"], d["R"] = l.P.toarray(), l.R.toarray()
        if hasattr(l, "pre"):
            d["pre"], d["post"] = l.pre, l.post
        out.append(d)
    return out


def _S(d, spec):
    kind, iters, omega = spec
    n = d["A"].shape[0]
    if kind == "none":
        return np.eye(n)
    return np.linalg.matrix_power(np.eye(n) - omega * d["Dinv"] @ d["A"], iters)


def _error_propagator(levels, l, cycle, cpl, coarse):
    """Dense multigrid error propagator of one cycle entering level l (textbook form)."""
    d = levels[l]
    L = len(levels)
    Ac = levels[l + 1]["A"]
    if l == L - 2:
        Ec = coarse(Ac)
    elif cycle == "V":
        Ec = _error_propagator(levels, l + 1, "V", cpl, coarse)
    elif cycle == "W":
        E1 = _error_propagator(levels, l + 1, "W", cpl, coarse)
        Ec = E1 @ E1
    else:
        Ec = _error_propagator(levels, l + 1, "F", cpl, coarse)
        EV = _error_propagator(levels, l + 1, "V", cpl, coarse)
        for _ in range(cpl):
            Ec = EV @ Ec
    n = d["A"].shape[0]
    CGC = np.eye(n) - d["P"] @ (np.eye(Ac.shape[0]) - Ec) @ np.linalg.pinv(Ac) @ d["R"] @ d["A"]
    return _S(d, d["post"]) @ CGC @ _S(d, d["pre"])


@pytest.mark.parametrize("cycle", ["V", "W", "F"])
@pytest.mark.parametrize("cpl", [1, 2])
def test_cycle_matches_error_propagator(cycle, cpl):
    """x_new = x* - E (x* - x) with E from dense matrices; different pre/post
    smoothers make a pre/post swap visible."""
    h = sa_hierarchy((12, 12))
    pre = ("jacobi", {"iterations": 2, "omega": 0.6})
    post = ("jacobi", {"iterations": 1, "omega": 0.9})
    om = OracleMG(h, pre, post)
    assert len(om.levels) >= 3
    lv = _dense_levels(om)
    E = _error_propagator(lv, 0, cycle, cpl, lambda Ac: np.zeros_like(Ac))
    rng = np.random.default_rng(0)
    A = lv[0]["A"]
    xs = rng.standard_normal(A.shape[0])
    b = A @ xs
    x0 = rng.standard_normal(A.shape[0])
    x1 = om.cycle(x0, b, cycle, cpl)
    assert rel(xs - x1, E @ (xs - x0)) < 1e-10


def test_cycle_with_jacobi_coarse_solver_matches_propagator():
    h = sa_hierarchy((12, 12))
    spec = ("jacobi", {"iterations": 1, "omega": 0.7})
    om = OracleMG(h, spec, spec, coarse_solver="jacobi", coarse_solver_kwargs={"iterations": 3, "omega": 0.5})
    lv = _dense_levels(om)

    def coarse(Ac):
        D = np.diag(1.0 / np.diag(Ac))
        # zero initial guess: e = x* - x_3, x_3 = (I - S^3) Ac^-1 b, so E_c = S^3
        return np.linalg.matrix_power(np.eye(Ac.shape[0]) - 0.5 * D @ Ac, 3)

    E = _error_propagator(lv, 0, "V", 1, coarse)
    A = lv[0]["A"]
    rng = np.random.default_rng(1)
    xs = rng.standard_normal(A.shape[0])
    x0 = rng.standard_normal(A.shape[0])
    assert rel(xs - om.cycle(x0, A @ xs, "V"), E @ (xs - x0)) < 1e-10


def test_symmetric_cycle_is_symmetric_operator():
    h = sa_hierarchy((12, 12))
    spec = ("jacobi", {"iterations": 1, "withrho": True})
    om = OracleMG(h, spec, spec)
    M = om.aspreconditioner("V")
    n = om.levels[0].A.shape[0]
    B = np.stack([M(e) for e in np.eye(n)], axis=1)
    assert np.abs(B - B.T).max() < 1e-12 * np.abs(B).max()
    assert np.all(np.linalg.eigvalsh((B + B.T) / 2) > 0)


def test_solve_loop_count_and_iterates():
    """Iterates follow x_{t+1} = x_t + (I - E) A^-1 (b - A x_t); the count is the
    first t with ||b - A x_t|| < tol ||b|| (amjax/multilevel.py:414-416)."""
    h = sa_hierarchy((12, 12))
    spec = ("jacobi", {"iterations": 1, "withrho": True})
    om = OracleMG(h, spec, spec)
    lv = _dense_levels(om)
    E = _error_propagator(lv, 0, "V", 1, lambda Ac: np.zeros_like(Ac))
    A = lv[0]["A"]
    Binv = (np.eye(A.shape[0]) - E) @ np.linalg.inv(A)
    b = np.random.default_rng(2).standard_normal(A.shape[0])
    for tol in (1e-3, 1e-8, 1e-12):
        x = np.zeros_like(b)
        t = 0
        while t < 100 and np.linalg.norm(b - A @ x) >= tol * np.linalg.norm(b):
            x = x + Binv @ (b - A @ x)
            t += 1
        xo, it = om.solve(b, tol=tol, maxiter=100, return_info=True)
        assert it == t
        assert rel(xo, x) < 1e-10
    # normb == 0 guard: zero right-hand side converges at once
    x0, it0 = om.solve(np.zeros_like(b), tol=1e-8, return_info=True)
    assert it0 == 0 and np.all(x0 == 0)


def test_solve_converges_to_direct_solution():
    h = sa_hierarchy((32, 32))
    spec = ("jacobi", {"iterations": 1, "withrho": True})
    om = OracleMG(h, spec, spec)
    A = om.levels[0].A
    b = np.random.default_rng(0).standard_normal(A.shape[0])
    xs = spla.spsolve(A.tocsc(), b)
    for cycle in "VWF":
        x = om.solve(b, tol=1e-12, maxiter=200, cycle=cycle)
        assert rel(x, xs) < 1e-9


def test_pcg_matches_scipy_cg():
    """jax-cg restatement against scipy.sparse.linalg.cg (same Hestenes-Stiefel
    recurrences, stopping on the recurrence residual)."""
    h = sa_hierarchy((32, 32))
    spec = ("jacobi", {"iterations": 1, "withrho": True})
    om = OracleMG(h, spec, spec)
    A = om.levels[0].A
    b = np.random.default_rng(0).standard_normal(A.shape[0])
    M = om.aspreconditioner()
    count = [0]
    xs, info = spla.cg(A, b, rtol=1e-8, atol=0.0, maxiter=500,
                       M=spla.LinearOperator(A.shape, matvec=M),
                       callback=lambda xk: count.__setitem__(0, count[0] + 1))
    x, it = pcg(A, b, M, tol=1e-8, maxiter=500, return_info=True)
    assert info == 0
    assert it == count[0]
    assert rel(x, xs) < 1e-10
    assert relres(A, b, x) < 1e-8
    # unpreconditioned
    count[0] = 0
    xs, _ = spla.cg(A, b, rtol=1e-8, atol=0.0, maxiter=2000,
                    callback=lambda xk: count.__setitem__(0, count[0] + 1))
    x, it = pcg(A, b, None, tol=1e-8, maxiter=2000, return_info=True)
    assert it == count[0] and rel(x, xs) < 1e-8


def test_multi_rhs_is_columnwise():
    h = sa_hierarchy((16, 16))
    om = OracleMG(h)
    B = np.random.default_rng(0).standard_normal((256, 3))
    X, its = om.solve(B, tol=1e-6, maxiter=400, return_info=True)
    for c in range(3):
        x, it = om.solve(B[:, c], tol=1e-6, maxiter=400, return_info=True)
        assert it == its[c] and np.array_equal(x, X[:, c])


# ---------------------------------------------------------------- polynomial smoothers

def test_chebyshev_coefficients_closed_forms():
    """Degree 1 is Richardson with 2/(a+b); C(0) = 1; C vanishes at the Chebyshev points
    of [a, b]; max |C| on [a, b] is 1/T_d((b+a)/(b-a)) (the minimax property)."""
    from oracle.solve import chebyshev_polynomial_coefficients
    a, b = 0.3, 8.5
    c1 = chebyshev_polynomial_coefficients(a, b, 1)
    np.testing.assert_allclose(-c1[:-1], [2.0 / (a + b)], rtol=1e-14)
    for d in (2, 3, 5):
        C = chebyshev_polynomial_coefficients(a, b, d)
        assert np.polyval(C, 0.0) == pytest.approx(1.0, abs=1e-14)
        pts = 0.5 * (b - a) * (1 + np.cos(np.pi * (np.arange(d) + 0.5) / d)) + a
        assert np.abs(np.polyval(C, pts)).max() < 1e-12
        lam = np.linspace(a, b, 20001)
        t = (b + a) / (b - a)
        Td = np.cosh(d * np.arccosh(t))
        assert np.abs(np.polyval(C, lam)).max() == pytest.approx(1.0 / Td, rel=1e-6)
    with pytest.raises(ValueError):
        chebyshev_polynomial_coefficients(1.0, 0.5, 3)


def test_polynomial_sweep_error_propagator():
    """One sweep with coefficients -C[:-1] maps the error e to C(A) e = prod (I - A/r_i) e
    (r_i the scaled Chebyshev roots), from a nonzero and from a zero guess; one
    coefficient gives x + c (b - A x)."""
    from oracle.solve import chebyshev_polynomial_coefficients, polynomial, poly_coefficients
    rng = np.random.default_rng(5)
    n = 30
    Q = np.linalg.qr(rng.standard_normal((n, n)))[0]
    A = (Q * rng.uniform(0.5, 6.0, n)) @ Q.T
    A = 0.5 * (A + A.T)
    a, b = 0.2, 7.0
    for d in (1, 2, 3, 4):
        C = chebyshev_polynomial_coefficients(a, b, d)
        roots = 0.5 * (b - a) * (1 + np.cos(np.pi * (np.arange(d) + 0.5) / d)) + a
        E = np.eye(n)
        for r_ in roots:
            E = E @ (np.eye(n) - A / r_)
        xs = rng.standard_normal(n)
        rhs = A @ xs
        for x0 in (rng.standard_normal(n), np.zeros(n)):
            x1 = polynomial(A, x0, rhs, -C[:-1])
            np.testing.assert_allclose(x1 - xs, E @ (x0 - xs), rtol=1e-10, atol=1e-12)
    x0 = rng.standard_normal(n)
    rhs = rng.standard_normal(n)
    np.testing.assert_allclose(polynomial(A, x0, rhs, [0.1]), x0 + 0.1 * (rhs - A @ x0), rtol=1e-15)
    # setup: richardson omega / rho, chebyshev on [rho/30, 1.1 rho], degree 3
    assert poly_coefficients("richardson", 4.0, {"omega": 2.0}) == [0.5]
    cc = poly_coefficients("chebyshev", 4.0, {})
    np.testing.assert_allclose(cc, -chebyshev_polynomial_coefficients(4.0 / 30, 4.4, 3)[:-1])
    assert len(cc) == 3


def test_bicgstab_matches_scipy_iterates():
    """The oracle's BiCGSTAB equals scipy.sparse.linalg.bicgstab (the same right-
    preconditioned algorithm, rhat = r0) after fixed iteration counts on a nonsymmetric
    system, with and without a preconditioner, and converges to spsolve."""
    import scipy.sparse.linalg as spla
    from oracle.solve import bicgstab
    n = 200
    rng = np.random.default_rng(0)
    A = (sp.diags([4.0 * np.ones(n), -1.3 * np.ones(n - 1), -0.7 * np.ones(n - 1)], [0, 1, -1])
         + sp.random(n, n, density=0.01, random_state=1) * 0.3).tocsr()
    b = rng.standard_normal(n)
    Dinv = 1.0 / A.diagonal()
    for M in (None, lambda v: Dinv * v):
        Mop = None if M is None else spla.LinearOperator((n, n), matvec=M)
        for its in (1, 2, 5):
            xo, k = bicgstab(A, b, M=M, tol=0.0, maxiter=its, return_info=True)
            xs, _ = spla.bicgstab(A, b, rtol=0.0, atol=0.0, maxiter=its, M=Mop)
            assert k == its
            np.testing.assert_allclose(xo, xs, rtol=1e-10, atol=1e-12)
        x = bicgstab(A, b, M=M, tol=1e-12, maxiter=500)
        np.testing.assert_allclose(x, spla.spsolve(A.tocsc(), b), rtol=1e-9)
    # a start x0: the same iterates as scipy started from x0
    x0 = rng.standard_normal(n)
    for its in (1, 3):
        xo, k = bicgstab(A, b, M=None, tol=0.0, maxiter=its, return_info=True, x0=x0)
        xs, _ = spla.bicgstab(A, b, x0=x0, rtol=0.0, atol=0.0, maxiter=its)
        assert k == its
        np.testing.assert_allclose(xo, xs, rtol=1e-10, atol=1e-12)


def test_pcg_ir_oracle_reduces_to_pcg_and_converges():
    """Refinement around an inner CG: with float64 vectors, one step and the
    Hestenes-Stiefel beta it is the oracle's PCG; in float32 it still reaches a
    float64 tolerance far below float32 precision, in a few steps."""
    from oracle.solve import OracleMG, pcg, pcg_ir, relres
    from conftest import sa_hierarchy
    h = sa_hierarchy((24, 24))
    om = OracleMG(h, ("jacobi", {"iterations": 1, "withrho": True}), ("jacobi", {"iterations": 1, "withrho": True}))
    A = om.levels[0].A
    b = np.random.default_rng(8).standard_normal(A.shape[0])
    M = om.aspreconditioner()
    x0, it0 = pcg(A, b, M, tol=1e-9, return_info=True)
    x1, it1, st1 = pcg_ir(A, b, M, tol=1e-9, inner_tol=1e-9, maxiter=1, dtype=np.float64, flexible=False,
                          return_info=True)
    assert (it1, st1) == (it0, 1)
    assert np.linalg.norm(x1 - x0) <= 1e-14 * np.linalg.norm(x0)
    om32 = OracleMG(h, ("jacobi", {"iterations": 1, "withrho": True}), ("jacobi", {"iterations": 1, "withrho": True}),
                    dtype=np.float32)
    x2, it2, st2 = pcg_ir(A, b, om32.aspreconditioner(), tol=1e-11, inner_tol=1e-4, maxiter=10, return_info=True)
    assert relres(A, b, x2) < 1e-11
    assert 2 <= st2 <= 4
    assert np.linalg.norm(x2 - pcg(A, b, M, tol=1e-14)) < 1e-9 * np.linalg.norm(x0)


def test_gmres_oracle_closed_forms():
    """The GMRES restatement: a full-length restart solves a small nonsymmetric
    system in one cycle (exact in exact arithmetic); the Givens least squares equals
    numpy's lstsq; restarted runs reach the tolerance on the preconditioned residual
    and agree with the direct solution; a zero right-hand side returns 0."""
    import scipy.sparse as sps
    from oracle.solve import _givens_lstsq, gmres
    rng = np.random.default_rng(4)
    n = 30
    A = sps.csr_matrix(np.eye(n) * 4 + rng.standard_normal((n, n)) * 0.3)
    b = rng.standard_normal(n)
    x, k = gmres(A, b, tol=1e-12, restart=n, return_info=True)
    assert k == 1 and np.linalg.norm(A @ x - b) < 1e-10 * np.linalg.norm(b)
    H = np.triu(rng.standard_normal((9, 8)), -1)
    y = _givens_lstsq(H, 8, 2.5)
    e1 = np.zeros(9)
    e1[0] = 2.5
    assert np.allclose(y, np.linalg.lstsq(H, e1, rcond=None)[0], rtol=1e-12, atol=1e-12)
    Dinv = 1.0 / A.diagonal()
    M = lambda v: Dinv * v
    x2, k2 = gmres(A, b, M=M, tol=1e-10, restart=5, return_info=True)
    assert k2 > 1
    assert np.linalg.norm(M(b - A @ x2)) <= 1e-10 * np.linalg.norm(M(b)) * 1.000001
    assert np.linalg.norm(x2 - np.linalg.solve(A.toarray(), b)) < 1e-8 * np.linalg.norm(x2)
    assert np.all(gmres(A, np.zeros(n)) == 0.0)
    # solve_method='incremental': the least-squares residual it stops on equals the
    # preconditioned residual of the updated x (exact arithmetic), so a full-length
    # restart ends in one restart, and it never needs more restarts than 'batched'
    pt = np.linalg.norm(M(b)) * 1e-10
    x5, k5 = gmres(A, b, M=M, tol=1e-10, restart=n, return_info=True, solve_method="incremental")
    assert k5 == 1 and np.linalg.norm(M(b - A @ x5)) <= 1.0001 * pt
    for rs in (3, 5, 8):
        xi, ki = gmres(A, b, M=M, tol=1e-10, restart=rs, return_info=True, solve_method="incremental")
        xb, kb = gmres(A, b, M=M, tol=1e-10, restart=rs, return_info=True)
        assert ki <= kb and np.linalg.norm(xi - np.linalg.solve(A.toarray(), b)) < 1e-8 * np.linalg.norm(xi)
    # a start x0: the exact solution returns at once; another converges to the same
    # preconditioned tolerance (ptol from b) and the direct solution
    xe = np.linalg.solve(A.toarray(), b)
    x3, k3 = gmres(A, b, M=M, tol=1e-10, restart=5, return_info=True, x0=xe)
    assert k3 == 0 and np.array_equal(x3, xe)
    x4, k4 = gmres(A, b, M=M, tol=1e-10, restart=5, return_info=True, x0=rng.standard_normal(n))
    assert k4 > 1 and np.linalg.norm(M(b - A @ x4)) <= 1e-10 * np.linalg.norm(M(b)) * 1.000001
    assert np.linalg.norm(x4 - xe) < 1e-8 * np.linalg.norm(xe)
