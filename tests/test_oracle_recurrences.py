"""Pins the Chebyshev iteration and the PPCG outer recurrence by closed forms
of the mathematics, on 8x8 grids, driving the kernel entry points directly.

* Chebyshev (tea_leaf_cheby.cuknl:5-152, tea_leaf_kernel_cuda.cu:105-135, the
  coefficients of tea_leaf_cheby.f90:147-174): for a generalized eigenpair
  A x = lam M x (M the preconditioner as a dense matrix), the right-hand side
  u0 = M x and u = 0 make r0 = M x an eigenvector of A M^-1, so after
  cheby_init and k cheby_iterate calls the residual is exactly
  P_k(lam) r0 with P_k(lam) = T_k((theta - lam)/delta) / T_k(theta/delta).
  An off-by-one in the coefficient index (the reference passes
  cheby_calc_step-1, tea_leaf_kernel_cuda.cu:132) or a wrong coefficient
  changes P_k.
* PPCG outer iteration (tea_solve.f90:294-331, tea_leaf_ppcg.cuknl:249-420):
  it is flexible CG with the polynomial preconditioner Q of the inner steps,
  for which Q A = I - P_{S+1}(M^-1 A) in closed form.  With the initial error
  in the span of m eigenvectors of M^-1 A the iteration terminates after m
  outer iterations (finite termination of CG), and u reaches A^-1 b.  With
  the inner polynomial changed between outer iterations (a variable
  preconditioner) the flexible beta = (r - r_store).z / rro keeps each new
  direction A-conjugate to the previous one and the recurrence residual equal
  to u0 - A u; the Fletcher-Reeves beta r.z / rro, or a wrong alpha, does
  not.

The same tests run on the HIP engine (marked gpu), which must agree.
"""
from __future__ import annotations

import numpy as np
import pytest
from numpy.polynomial import chebyshev as npcheb
from scipy.linalg import eigh

from tealeaf_cuda_amd import Engine
from tealeaf_cuda_amd import runtime as rt
from test_oracle_math import dense_operator, interior, put_interior, setup_chunk

NX = NY = 8
N = NX * NY


@pytest.fixture(params=["oracle", pytest.param("hip", marks=pytest.mark.gpu)])
def eng(request, oracle_so, oracle_env):
    oracle_env("dot2")
    e = Engine(oracle_so) if request.param == "oracle" else Engine()
    yield e
    e.finalize()


def dense_preconditioner(A: np.ndarray, prec: str) -> np.ndarray:
    """M as a dense matrix: I, diag(A), or the 4x4 diagonal blocks of A on
    strips of 4 rows per column aligned at the first interior row
    (tea_block_jacobi.cuknl, cuda_kernel_header.hpp:11-16)."""
    if prec == "none":
        return np.eye(A.shape[0])
    if prec == "jac_diag":
        return np.diag(np.diag(A))
    M = np.zeros_like(A)
    for c in range(NX):
        for r0 in range(0, NY, 4):
            ids = [r * NX + c for r in range(r0, min(r0 + 4, NY))]
            M[np.ix_(ids, ids)] = A[np.ix_(ids, ids)]
    return M


def cheb_coefficients(lmin: float, lmax: float, n: int):
    """theta, delta and n (alpha, beta) pairs of the Chebyshev semi-iteration
    for the interval [lmin, lmax] (Saad, Iterative Methods, Alg. 12.1)."""
    theta, delta = (lmax + lmin) / 2, (lmax - lmin) / 2
    sigma = theta / delta
    rho_old = 1.0 / sigma
    alphas, betas = [], []
    for _ in range(n):
        rho_new = 1.0 / (2.0 * sigma - rho_old)
        alphas.append(rho_new * rho_old)
        betas.append(2.0 * rho_new / delta)
        rho_old = rho_new
    return theta, delta, alphas, betas


def residual_polynomial(k: int, lam, theta: float, delta: float):
    """P_k(lam) = T_k((theta - lam)/delta) / T_k(theta/delta)."""
    ek = np.zeros(k + 1)
    ek[k] = 1.0
    return npcheb.chebval((theta - np.asarray(lam)) / delta, ek) / npcheb.chebval(theta / delta, ek)


def operator_and_eigenpairs(eng, tmp_path, solver, prec):
    # dt 4 makes the operator stiff enough (condition 50-180) that the
    # polynomials damp slowly and every step's coefficients matter
    setup_chunk(eng, tmp_path, nx=NX, ny=NY, solver=solver, prec=prec, dt=4.0)
    A = dense_operator(eng)
    M = dense_preconditioner(A, prec)
    eng.calc_residual()
    eng.cg_init()                  # the preconditioner's set-up (Mi, or the Thomas factors)
    lam, X = eigh(A, M)            # A x = lam M x, X^T M X = I
    return A, M, lam, X


@pytest.mark.parametrize("prec", ["none", "jac_diag", "jac_block"])
def test_chebyshev_residual_is_the_chebyshev_polynomial(eng, tmp_path, prec):
    A, M, lam, X = operator_and_eigenpairs(eng, tmp_path, "chebyshev", prec)
    K = 12
    theta, delta, alphas, betas = cheb_coefficients(lam[0] * 0.95, lam[-1] * 1.05, K)
    for j in (0, 1, N // 3, N // 2, N - 1):
        r0 = M @ X[:, j]
        put_interior(eng, rt.FIELD_U, np.zeros((NY, NX)))
        put_interior(eng, rt.XFIELD_U0, r0.reshape(NY, NX))
        eng.cheby_init(alphas, betas, theta)
        scale = np.abs(r0).max()
        r = interior(eng, eng.get_field(rt.FIELD_R)).ravel()
        np.testing.assert_allclose(r, r0, rtol=0, atol=1e-14 * scale)
        for k in range(1, K + 1):
            eng.cheby_iterate(k)
            r = interior(eng, eng.get_field(rt.FIELD_R)).ravel()
            expect = residual_polynomial(k, lam[j], theta, delta) * r0
            np.testing.assert_allclose(r, expect, rtol=0, atol=1e-12 * scale, err_msg=f"eig {j} step {k}")
        eng.calc_residual()        # u after the last iterate: one degree more
        r = interior(eng, eng.get_field(rt.FIELD_R)).ravel()
        np.testing.assert_allclose(r, residual_polynomial(K + 1, lam[j], theta, delta) * r0, rtol=0,
                                   atol=1e-12 * scale)


@pytest.mark.parametrize("prec", ["none", "jac_diag", "jac_block"])
def test_chebyshev_residual_mixed_eigenvectors(eng, tmp_path, prec):
    """Linearity: r0 spanning several eigenvectors damps each by its own P_k."""
    A, M, lam, X = operator_and_eigenpairs(eng, tmp_path, "chebyshev", prec)
    K = 7
    theta, delta, alphas, betas = cheb_coefficients(lam[0] * 0.95, lam[-1] * 1.05, K)
    c = np.random.default_rng(3).standard_normal(N)
    r0 = M @ (X @ c)
    put_interior(eng, rt.FIELD_U, np.zeros((NY, NX)))
    put_interior(eng, rt.XFIELD_U0, r0.reshape(NY, NX))
    eng.cheby_init(alphas, betas, theta)
    for k in range(1, K + 1):
        eng.cheby_iterate(k)
    r = interior(eng, eng.get_field(rt.FIELD_R)).ravel()
    expect = M @ (X @ (residual_polynomial(K, lam, theta, delta) * c))
    np.testing.assert_allclose(r, expect, rtol=0, atol=1e-12 * np.abs(r0).max())


def _inner_steps(eng, theta, S):
    """tea_leaf_run_ppcg_inner_steps (tea_leaf_ppcg.f90:174-230) at one chunk."""
    eng.ppcg_init_sd_new(theta)
    for k in range(1, S + 1):
        eng.ppcg_inner(k)
    eng.ppcg_update_z()


def _outer_iteration(eng, theta, S):
    """One PPCG outer iteration in the reference's call order
    (tea_solve.f90:294-331); returns (alpha, beta, rro, rrn)."""
    pw = eng.cg_calc_w()
    eng.ppcg_store_r()
    rro = eng.ppcg_calc_2norm()
    alpha = rro / pw
    eng.cg_calc_ur(alpha)
    _inner_steps(eng, theta, S)
    rrn = eng.ppcg_calc_rrn()
    beta = rrn / rro
    eng.ppcg_calc_p(beta)
    return alpha, beta, rro, rrn


def _start(eng, b, theta, S):
    """u = 0, r = b; the re-initialisation at the switch to PPCG
    (tea_solve.f90:223-240): ppcg_init step 2, inner steps, step 3."""
    put_interior(eng, rt.FIELD_U, np.zeros((NY, NX)))
    put_interior(eng, rt.XFIELD_U0, b.reshape(NY, NX))
    eng.calc_residual()
    eng.ppcg_init(2)
    _inner_steps(eng, theta, S)
    return eng.ppcg_init(3)


@pytest.mark.parametrize("prec", ["none", "jac_diag", "jac_block"])
@pytest.mark.parametrize("m", [1, 2, 3])
def test_ppcg_outer_finite_termination(eng, tmp_path, prec, m):
    A, M, lam, X = operator_and_eigenpairs(eng, tmp_path, "ppcg", prec)
    S = 4
    theta, delta, alphas, betas = cheb_coefficients(lam[0] * 0.95, lam[-1] * 1.05, S)
    eng.ppcg_init_constants(alphas, betas)
    # Q A = I - P_{S+1}(M^-1 A): eigenvalues mu = 1 - P_{S+1}(lam) on the
    # eigenvectors of M^-1 A; take m of them with well separated mu
    mu = 1.0 - residual_polynomial(S + 1, lam, theta, delta)
    assert np.all((mu > 0) & (mu < 2))
    picks = [int(np.argmin(np.abs(mu - t))) for t in np.linspace(mu.min(), mu.max(), m + 2)[1:-1]]
    assert len(set(np.round(mu[picks], 6))) == m
    # residual components of similar size: e0 = sum c_j x_j / lam_j, b = A e0 = sum c_j M x_j
    e0 = X[:, picks] @ (np.array([1.0, -0.7, 0.4][:m]) / lam[picks])
    b = A @ e0
    _start(eng, b, theta, S)
    rnorm0 = np.linalg.norm(b)
    history = []
    for _ in range(m):
        _outer_iteration(eng, theta, S)
        history.append(np.linalg.norm(interior(eng, eng.get_field(rt.FIELD_R))) / rnorm0)
    assert history[-1] <= 1e-12, history
    assert all(h > 1e-3 for h in history[:-1]), history   # not earlier: each iteration takes one eigenvalue
    u = interior(eng, eng.get_field(rt.FIELD_U)).ravel()
    np.testing.assert_allclose(u, e0, rtol=0, atol=1e-11 * np.abs(e0).max())
    eng.calc_residual()            # the recurrence residual is the true one
    assert np.linalg.norm(interior(eng, eng.get_field(rt.FIELD_R))) / rnorm0 <= 1e-12


@pytest.mark.parametrize("prec", ["none", "jac_diag", "jac_block"])
def test_ppcg_outer_flexible_recurrence(eng, tmp_path, prec):
    """The inner polynomial changes every outer iteration (a variable
    preconditioner): each new direction stays A-conjugate to the previous one,
    the new residual stays orthogonal to the previous direction, and the
    recurrence residual stays u0 - A u."""
    A, M, lam, X = operator_and_eigenpairs(eng, tmp_path, "ppcg", prec)
    S = 3
    b = np.random.default_rng(11).standard_normal(N)
    theta0, _, a0, b0 = cheb_coefficients(lam[0] * 0.95, lam[-1] * 1.05, S)
    eng.ppcg_init_constants(a0, b0)
    _start(eng, b, theta0, S)
    p_old = interior(eng, eng.get_field(rt.FIELD_P)).ravel()
    Anorm = lambda v: np.sqrt(v @ (A @ v))
    for it, (fl, fh) in enumerate([(0.3, 1.0), (0.95, 0.6), (0.5, 1.4), (0.9, 1.05), (0.2, 0.8)]):
        theta, _, al, be = cheb_coefficients(lam[0] * fl, lam[-1] * fh, S)
        # the coefficients of the inner steps of this iteration: uploaded before cg_calc_w, whose
        # results they do not affect
        eng.ppcg_init_constants(al, be)
        alpha, beta, rro, rrn = _outer_iteration(eng, theta, S)
        p = interior(eng, eng.get_field(rt.FIELD_P)).ravel()
        r = interior(eng, eng.get_field(rt.FIELD_R)).ravel()
        u = interior(eng, eng.get_field(rt.FIELD_U)).ravel()
        assert abs(p @ (A @ p_old)) <= 1e-12 * Anorm(p) * Anorm(p_old), f"iteration {it}: directions not conjugate"
        assert abs(r @ p_old) <= 1e-12 * np.linalg.norm(r) * np.linalg.norm(p_old), f"iteration {it}"
        np.testing.assert_allclose(r, b - A @ u, rtol=0, atol=1e-12 * np.abs(b).max())
        p_old = p
