"""Pins the oracle's kernels by what the mathematics fixes, on grids small
enough for dense linear algebra:

* the operator assembled column by column through calc_residual equals the
  textbook 5-point matrix built from Kx/Ky, is symmetric positive definite and
  has unit row sums (zero-flux boundaries), and Kx/Ky have closed forms for
  uniform density under both coefficient options;
* CG reaches the dense solve; Jacobi converges to it;
* jac_diag is division by the diagonal; jac_block equals dense solves of the
  4x4 diagonal blocks of A on strips aligned at the first interior row
  (truncated to 2 rows for a ragged top strip);
* the Lanczos eigenvalue estimate from CG alpha/beta (tqli in the driver)
  equals numpy's eigvalsh of the same tridiagonal, lies inside the dense
  spectrum, and is scaled by the single-precision 0.95 / 1.05 of tea_solve.f90;
* PPCG inner steps keep rtemp = r - A utemp exactly and reduce the residual by
  at least the Chebyshev bound 1/T_S(theta/delta) in the M^-1 norm;
* field_summary has closed forms on a uniform state, and conserves energy;
* the reflective halo update and the pack/unpack buffers follow the mirror
  rule and layout of pack_kernel.cuknl.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import GOLDEN
from tealeaf_cuda_amd import Deck, Engine, StateSpec, TeaError, run_deck
from tealeaf_cuda_amd import runtime as rt


def setup_chunk(eng: Engine, tmp_path, nx=12, ny=8, solver="cg", prec="none", halo=1, states=None,
                xmax=10.0, ymax=10.0, dt=0.004, coefficient=rt.CONDUCTIVITY, extra=()):
    deck = Deck(x_cells=nx, y_cells=ny, solver=solver, preconditioner=prec, halo_depth=halo,
                states=states or [StateSpec(100.0, 0.0001), StateSpec(0.1, 25.0, "rectangle", 0.0, 4.0, 2.0, 6.0),
                                  StateSpec(3.0, 1.0, "rectangle", 5.0, 9.0, 1.0, 9.0)],
                xmax=xmax, ymax=ymax, extra=list(extra))
    path = deck.write(tmp_path / "tea.in")
    eng.initialise(path, nx, ny)
    dx, dy = xmax / nx, ymax / ny
    eng.initialise_chunk(0.0, 0.0, dx, dy)
    st = deck.states
    sh = [(s.xmin + dx / 100, s.xmax - dx / 100, s.ymin + dy / 100, s.ymax - dy / 100) if i else (0, 0, 0, 0)
          for i, s in enumerate(st)]
    eng.generate_chunk([s.density for s in st], [s.energy for s in st], [b[0] for b in sh], [b[1] for b in sh],
                       [b[2] for b in sh], [b[3] for b in sh], [s.radius for s in st],
                       [{"rectangle": 1, "circle": 2, "point": 3}.get(s.geometry or "", 0) for s in st])
    eng.set_field_kernel()
    rx, ry = eng.common_init(coefficient, dt)
    return rx, ry


def interior(eng, a):
    h = eng.h
    return a[h:h + eng.ny, h:h + eng.nx]


def put_interior(eng, field, vals):
    a = np.zeros(eng.shape())
    h = eng.h
    a[h:h + eng.ny, h:h + eng.nx] = vals
    eng.set_field(field, a)


def dense_operator(eng):
    """Column j of A from r = u0 - A u with u = e_j, u0 = 0 (calc_residual)."""
    n = eng.nx * eng.ny
    A = np.zeros((n, n))
    put_interior(eng, rt.XFIELD_U0, np.zeros((eng.ny, eng.nx)))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        put_interior(eng, rt.FIELD_U, e.reshape(eng.ny, eng.nx))
        eng.calc_residual()
        A[:, j] = -interior(eng, eng.get_field(rt.FIELD_R)).ravel()
    return A


def five_point(eng):
    """The textbook 5-point matrix from the face coefficients Kx, Ky."""
    h, nx, ny = eng.h, eng.nx, eng.ny
    Kx, Ky = eng.get_field(rt.XFIELD_KX), eng.get_field(rt.XFIELD_KY)
    n = nx * ny
    A = np.zeros((n, n))
    idx = lambda c, r: r * nx + c
    for r in range(ny):
        for c in range(nx):
            R, C = r + h, c + h
            i = idx(c, r)
            A[i, i] = 1.0 + Kx[R, C] + Kx[R, C + 1] + Ky[R, C] + Ky[R + 1, C]
            if c > 0:
                A[i, idx(c - 1, r)] = -Kx[R, C]
            if c < nx - 1:
                A[i, idx(c + 1, r)] = -Kx[R, C + 1]
            if r > 0:
                A[i, idx(c, r - 1)] = -Ky[R, C]
            if r < ny - 1:
                A[i, idx(c, r + 1)] = -Ky[R + 1, C]
    return A


@pytest.fixture
def eng(oracle_so, oracle_env):
    oracle_env("dot2")
    e = Engine(oracle_so)
    yield e
    e.finalize()


def test_operator_is_the_five_point_matrix(eng, tmp_path):
    setup_chunk(eng, tmp_path)
    A = dense_operator(eng)
    B = five_point(eng)
    np.testing.assert_allclose(A, B, rtol=1e-14, atol=1e-14 * np.abs(B).max())
    assert np.array_equal(A, A.T)                                   # symmetric, bit for bit
    assert np.linalg.eigvalsh(A).min() > 0.99                       # I + SPD Laplacian
    np.testing.assert_allclose(A.sum(axis=1), 1.0, rtol=0, atol=1e-12 * np.abs(B).max())  # zero flux


@pytest.mark.parametrize("coef", [rt.CONDUCTIVITY, rt.RECIP_CONDUCTIVITY])
def test_face_coefficients_uniform_density(eng, tmp_path, coef):
    rho = 2.5
    rx, ry = setup_chunk(eng, tmp_path, nx=10, ny=6, xmax=5.0, ymax=3.0, states=[StateSpec(rho, 1.0)],
                         coefficient=coef)
    assert rx == pytest.approx(0.004 / 0.5 ** 2, rel=1e-15) and ry == pytest.approx(0.004 / 0.5 ** 2, rel=1e-15)
    Kx, Ky = eng.get_field(rt.XFIELD_KX), eng.get_field(rt.XFIELD_KY)
    h = eng.h
    w = rho if coef == rt.CONDUCTIVITY else 1.0 / rho
    # interior faces: K = r * (w + w) / (2 w w) = r / w ; boundary faces zeroed
    np.testing.assert_allclose(Kx[h:h + 6, h + 1:h + 10], rx / w, rtol=1e-15)
    np.testing.assert_allclose(Ky[h + 1:h + 6, h:h + 10], ry / w, rtol=1e-15)
    assert np.all(Kx[h:h + 6, h] == 0) and np.all(Kx[h:h + 6, h + 10] == 0)
    assert np.all(Ky[h, h:h + 10] == 0) and np.all(Ky[h + 6, h:h + 10] == 0)


def test_cg_reaches_dense_solve(oracle_so, oracle_env, tmp_path):
    oracle_env("dot2")
    eng = Engine(oracle_so)
    setup_chunk(eng, tmp_path, nx=10, ny=7)
    A = dense_operator(eng)
    eng.finalize()
    deck = Deck(x_cells=10, y_cells=7, end_step=1, solver="cg", eps=1e-15,
                states=[StateSpec(100.0, 0.0001), StateSpec(0.1, 25.0, "rectangle", 0.0, 4.0, 2.0, 6.0),
                        StateSpec(3.0, 1.0, "rectangle", 5.0, 9.0, 1.0, 9.0)])
    res = run_deck(deck.write(tmp_path / "cg.in"), backend=oracle_so)
    eng = Engine(oracle_so)  # same library instance: the chunk the driver left behind
    u = interior(eng, eng.get_field(rt.FIELD_U)).ravel()
    u0 = interior(eng, eng.get_field(rt.XFIELD_U0)).ravel()
    np.testing.assert_allclose(u, np.linalg.solve(A, u0), rtol=1e-10)
    assert res.iterations[0] <= 70


def test_jacobi_converges_to_dense_solve(oracle_so, oracle_env, tmp_path):
    oracle_env("dot2")
    eng = Engine(oracle_so)
    setup_chunk(eng, tmp_path, nx=8, ny=6, solver="jacobi")
    A = dense_operator(eng)
    put_interior(eng, rt.FIELD_U, np.zeros((6, 8)))
    u0 = interior(eng, eng.get_field(rt.XFIELD_U0)).ravel()
    put_interior(eng, rt.FIELD_U, u0.reshape(6, 8))
    for _ in range(3000):
        err = eng.jacobi_solve()
        if err < 1e-13:
            break
    u = interior(eng, eng.get_field(rt.FIELD_U)).ravel()
    np.testing.assert_allclose(u, np.linalg.solve(A, u0), rtol=1e-9)
    eng.finalize()


@pytest.mark.parametrize("ny", [8, 10])
def test_block_jacobi_is_dense_strip_solve(eng, tmp_path, ny):
    setup_chunk(eng, tmp_path, nx=6, ny=ny, prec="jac_block")
    A = dense_operator(eng)
    rng = np.random.default_rng(ny)
    r = rng.standard_normal((ny, 6))
    put_interior(eng, rt.FIELD_R, r)
    rro = eng.cg_init()
    z = interior(eng, eng.get_field(rt.XFIELD_Z))
    expect = np.zeros_like(r)
    for c in range(6):
        for r0 in range(0, ny, 4):
            rows = list(range(r0, min(r0 + 4, ny)))
            ids = [rr * 6 + c for rr in rows]
            expect[rows, c] = np.linalg.solve(A[np.ix_(ids, ids)], r[rows, c])
    np.testing.assert_allclose(z, expect, rtol=1e-13, atol=1e-15)
    assert rro == pytest.approx(float(np.sum(r * z)), rel=1e-14)


def test_jac_diag_divides_by_diagonal(eng, tmp_path):
    setup_chunk(eng, tmp_path, nx=7, ny=5, prec="jac_diag")
    A = dense_operator(eng)
    r = np.random.default_rng(1).standard_normal((5, 7))
    put_interior(eng, rt.FIELD_R, r)
    eng.cg_init()
    z = interior(eng, eng.get_field(rt.XFIELD_Z))
    np.testing.assert_allclose(z.ravel(), r.ravel() / np.diag(A), rtol=1e-15)


def _cg_coefficients(eng, nsteps):
    rro = eng.cg_init()
    alphas, betas = [], []
    for _ in range(nsteps):
        pw = eng.cg_calc_w()
        a = rro / pw
        rrn = eng.cg_calc_ur(a)
        b = rrn / rro
        eng.cg_calc_p(b)
        alphas.append(a)
        betas.append(b)
        rro = rrn
    return alphas, betas


def test_lanczos_eigenvalues(oracle_so, oracle_env, tmp_path):
    oracle_env("dot2")
    nsteps = 15
    states = [StateSpec(100.0, 0.0001), StateSpec(0.1, 25.0, "rectangle", 0.0, 4.0, 2.0, 6.0)]
    eng = Engine(oracle_so)
    setup_chunk(eng, tmp_path, nx=12, ny=10, solver="ppcg", states=states)
    A = dense_operator(eng)
    setup_chunk(eng, tmp_path, nx=12, ny=10, solver="ppcg", states=states)
    eng.calc_residual()
    alphas, betas = _cg_coefficients(eng, nsteps)
    eng.finalize()
    # Lanczos tridiagonal of CG (tea_leaf_cheby.f90:118-125)
    d = np.array([1.0 / alphas[0]] + [1.0 / alphas[k] + betas[k - 1] / alphas[k - 1] for k in range(1, nsteps)])
    e = np.array([np.sqrt(betas[k]) / alphas[k] for k in range(nsteps - 1)])
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    ritz = np.linalg.eigvalsh(T)
    lam = np.linalg.eigvalsh(A)
    assert lam.min() - 1e-9 <= ritz.min() and ritz.max() <= lam.max() * (1 + 1e-9)
    assert ritz.max() >= 0.9 * lam.max()

    # the driver switches to PPCG after nsteps+1 CG steps and reports the
    # estimate scaled by the single-precision 0.95 / 1.05 literals
    deck = Deck(x_cells=12, y_cells=10, end_step=1, solver="ppcg", preconditioner="none", eps=1e-30,
                max_iters=nsteps + 3, ch_cg_presteps=nsteps + 1, ch_cg_epslim=1e300, ppcg_inner_steps=4,
                states=states)
    res = run_deck(deck.write(tmp_path / "pp.in"), backend=oracle_so)
    assert res.eigmin == pytest.approx(ritz.min() * float(np.float32(0.95)), rel=1e-12)
    assert res.eigmax == pytest.approx(ritz.max() * float(np.float32(1.05)), rel=1e-12)


def _cheb_T(k, x):
    return np.cosh(k * np.arccosh(x))


@pytest.mark.parametrize("prec", ["none", "jac_diag", "jac_block"])
def test_ppcg_inner_steps_chebyshev_bound(eng, tmp_path, prec):
    nx, ny, S = 8, 8, 6
    setup_chunk(eng, tmp_path, nx=nx, ny=ny, solver="ppcg", prec=prec)
    A = dense_operator(eng)
    eng.cg_init()  # builds Mi or the block factors
    n = nx * ny
    Minv = np.zeros((n, n))
    for j in range(n):                  # M^-1 column by column through ppcg_init step 2
        e = np.zeros(n)
        e[j] = 1.0
        put_interior(eng, rt.FIELD_R, e.reshape(ny, nx))
        eng.ppcg_init(2)
        Minv[:, j] = interior(eng, eng.get_field(rt.XFIELD_Z)).ravel() if prec != "none" else e
    lam = np.linalg.eigvals(Minv @ A).real
    lmin, lmax = lam.min() * 0.99, lam.max() * 1.01
    theta, delta = (lmax + lmin) / 2, (lmax - lmin) / 2
    sigma = theta / delta
    rho_old = 1.0 / sigma
    alphas, betas = [], []
    for _ in range(S):
        rho_new = 1.0 / (2.0 * sigma - rho_old)
        alphas.append(rho_new * rho_old)
        betas.append(2.0 * rho_new / delta)
        rho_old = rho_new
    eng.ppcg_init_constants(alphas, betas)
    r = np.random.default_rng(5).standard_normal((ny, nx))
    put_interior(eng, rt.FIELD_R, r)
    eng.ppcg_init(2)
    eng.ppcg_init_sd_new(theta)
    for k in range(1, S + 1):
        eng.ppcg_inner(k)
    rtemp = interior(eng, eng.get_field(rt.XFIELD_RTEMP)).ravel()
    utemp = interior(eng, eng.get_field(rt.XFIELD_UTEMP)).ravel()
    sd = interior(eng, eng.get_field(rt.FIELD_SD)).ravel()
    r = r.ravel()
    # invariant of the recurrence: rtemp holds the residual of utemp before the last update
    np.testing.assert_allclose(rtemp, r - A @ (utemp - sd), rtol=0, atol=1e-12 * np.abs(r).max())
    # utemp is the (S+1)-step Chebyshev iterate: residual bound 1/T_{S+1}(theta/delta) in the M^-1 norm
    norm = lambda v: np.sqrt(v @ (Minv @ v))
    assert norm(r - A @ utemp) <= norm(r) / _cheb_T(S + 1, sigma) * (1 + 1e-9)
    assert norm(rtemp) <= norm(r) / _cheb_T(S, sigma) * (1 + 1e-9)


def test_field_summary_closed_form(eng, tmp_path):
    setup_chunk(eng, tmp_path, nx=12, ny=8, xmax=3.0, ymax=2.0, states=[StateSpec(2.0, 3.0)])
    vol, mass, ie, temp = eng.field_summary()
    assert vol == pytest.approx(6.0, rel=1e-14)
    assert mass == pytest.approx(12.0, rel=1e-14)
    assert ie == pytest.approx(36.0, rel=1e-14)
    assert temp == pytest.approx(72.0, rel=1e-14)     # u = density * energy = 6


def test_energy_conserved_by_solve(oracle_so, oracle_env):
    oracle_env("dot2")
    res = run_deck(GOLDEN / "tea_bm_2.in", backend=oracle_so)
    ie0, ie1 = res.summaries[0].energy, res.summaries[-1].energy
    assert ie1 == pytest.approx(ie0, rel=1e-12)
    assert res.summaries[-1].mass == res.summaries[0].mass


@pytest.mark.parametrize("depth", [1, 2])
def test_reflective_halo_mirror(eng, tmp_path, depth):
    setup_chunk(eng, tmp_path, nx=7, ny=5, halo=2)
    h = eng.h
    a = np.random.default_rng(depth).standard_normal(eng.shape())
    eng.set_field(rt.FIELD_ENERGY0, a)
    fields = [0] * 7
    fields[rt.FIELD_ENERGY0 - 1] = 1
    eng.update_halo(fields, depth)
    b = eng.get_field(rt.FIELD_ENERGY0)
    exp = a.copy()
    y0, y1, x0, x1 = h, h + 5, h, h + 7       # interior [y0, y1) x [x0, x1)
    for k in range(1, depth + 1):             # bottom / top over the widened columns
        exp[y0 - k, x0 - depth:x1 + depth] = exp[y0 + k - 1, x0 - depth:x1 + depth]
        exp[y1 - 1 + k, x0 - depth:x1 + depth] = exp[y1 - k, x0 - depth:x1 + depth]
    for k in range(1, depth + 1):             # then left / right over the widened rows
        exp[y0 - depth:y1 + depth, x0 - k] = exp[y0 - depth:y1 + depth, x0 + k - 1]
        exp[y0 - depth:y1 + depth, x1 - 1 + k] = exp[y0 - depth:y1 + depth, x1 - k]
    np.testing.assert_array_equal(b, exp)


@pytest.mark.parametrize("depth", [1, 2])
def test_pack_unpack_layout(eng, tmp_path, depth):
    nx, ny = 7, 5
    setup_chunk(eng, tmp_path, nx=nx, ny=ny, halo=2)
    h = eng.h
    a = np.random.default_rng(10 + depth).standard_normal(eng.shape())
    eng.set_field(rt.FIELD_U, a)
    fields = [0] * 7
    fields[rt.FIELD_U - 1] = 1
    off = [0] * 7
    lr, bt = depth * (ny + 2 * depth), depth * (nx + 2 * depth)
    rows = slice(h - depth, h + ny + depth)
    cols = slice(h - depth, h + nx + depth)
    left = eng.pack(fields, off, depth, rt.CHUNK_LEFT, lr).reshape(ny + 2 * depth, depth)
    right = eng.pack(fields, off, depth, rt.CHUNK_RIGHT, lr).reshape(ny + 2 * depth, depth)
    bottom = eng.pack(fields, off, depth, rt.CHUNK_BOTTOM, bt).reshape(depth, nx + 2 * depth)
    top = eng.pack(fields, off, depth, rt.CHUNK_TOP, bt).reshape(depth, nx + 2 * depth)
    # left and bottom strips hold the face cells mirrored, right and top in natural order
    np.testing.assert_array_equal(left, a[rows, h:h + depth][:, ::-1])
    np.testing.assert_array_equal(right, a[rows, h + nx - depth:h + nx])
    np.testing.assert_array_equal(bottom, a[h:h + depth, cols][::-1, :])
    np.testing.assert_array_equal(top, a[h + ny - depth:h + ny, cols])
    # a neighbour's right strip unpacked into our left halo lands in natural order
    eng.unpack(fields, off, depth, rt.CHUNK_LEFT, right.ravel())
    b = eng.get_field(rt.FIELD_U)
    np.testing.assert_array_equal(b[rows, h - depth:h], a[rows, h + nx - depth:h + nx])
    # our own left strip unpacked into our left halo is the reflective update
    eng.set_field(rt.FIELD_U, a)
    eng.unpack(fields, off, depth, rt.CHUNK_LEFT, left.ravel())
    b = eng.get_field(rt.FIELD_U)
    for k in range(1, depth + 1):
        np.testing.assert_array_equal(b[rows, h - k], a[rows, h + k - 1])


def test_errors_are_reported(eng, tmp_path):
    with pytest.raises(TeaError, match="before initialise"):
        eng.cg_calc_w()
    setup_chunk(eng, tmp_path, solver="cg")
    with pytest.raises(TeaError, match="norm_array"):
        eng.calc_2norm(3)
    with pytest.raises(TeaError, match="not registered"):
        eng.ppcg_store_r()
    with pytest.raises(TeaError, match="not registered"):
        eng.jacobi_solve()
    with pytest.raises(TeaError, match="face"):
        eng.pack([1] + [0] * 6, [0] * 7, 1, 7, 100)


@pytest.mark.parametrize("bad,match", [
    (dict(states=[]), "No states"),
    (dict(halo_depth=0), "halo"),
    (dict(solver="ppcg", preconditioner="jac_block", halo_depth=2), "block jacobi"),
    (dict(preconditioner="jacobi_block"), "preconditioner"),
])
def test_invalid_decks(oracle_so, driver_built, tmp_path, bad, match):
    kw = dict(x_cells=8, y_cells=8, end_step=1)
    kw.update(bad)
    d = Deck(**kw)
    with pytest.raises(TeaError, match=match):
        run_deck(d.write(tmp_path / "bad.in"), backend=oracle_so)
