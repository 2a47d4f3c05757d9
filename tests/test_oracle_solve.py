"""Pins of the oracle's PoissonDirect restatement (oracle/orc_solve.c, solver.js:105-165):
scipy's DST-II, closed-form eigenvalues, a manufactured solution and a sparse-LU solve."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla
from scipy.fft import dst

import _oracle
from _ref_solve import apply_operator, solve_scipy
from _util import rel_l2


@pytest.mark.parametrize("N", [8, 16, 37, 144])
def test_dst_matrix_is_scipy_orthonormal_dst2(N):
    q, qi = _oracle.dst_matrix(N)
    x = np.random.default_rng(N).standard_normal(N)
    assert np.allclose(qi @ x, dst(x, type=2, norm="ortho"), atol=1e-13)
    assert np.allclose(q @ qi, np.eye(N), atol=1e-13)
    assert np.array_equal(q.T, qi)


@pytest.mark.parametrize("N", [8, 30])
def test_dst_eigenvalues_closed_form(N):
    """v_l is an eigenvector of the 1-D Laplacian with phi = 0 on the outer faces (antisymmetric ghosts)."""
    dx = 750.0
    q, _ = _oracle.dst_matrix(N)
    lam = _oracle.dst_eig(N, dx)
    L = (np.diag(-2 * np.ones(N)) + np.diag(np.ones(N - 1), 1) + np.diag(np.ones(N - 1), -1))
    L[0, 0] = L[-1, -1] = -3.0
    L /= dx * dx
    for l in range(N):
        v = q[:, l]
        assert np.allclose(L @ v, lam[l] * v, atol=1e-12 / dx ** 2)


def _density(nz):
    rc = np.linspace(1.15, 0.2, nz)
    rw = np.concatenate([[1.2], 0.5 * (rc[1:] + rc[:-1]), [0.94 * rc[-1]]])
    return rc, rw


def test_oracle_solve_matches_scipy_and_manufactured_solution():
    nz, ny, nx = 12, 20, 24
    dx, dz = 1000.0, 400.0
    rc, rw = _density(nz)
    rng = np.random.default_rng(1)
    phi_true = rng.standard_normal((nz, ny, nx))
    rhs = apply_operator(phi_true, dx, dz, rc, rw)
    phi = _oracle.solve(nx, ny, nz, dx, dz, rc, rw, rhs, "f64")
    assert rel_l2(phi, phi_true) < 1e-9
    assert rel_l2(phi, solve_scipy(rhs, dx, dz, rc, rw)) < 1e-12


def test_oracle_solve_matches_sparse_lu():
    """The assembled 7-point operator solved by sparse LU (scipy.sparse.linalg.spsolve)."""
    nz, ny, nx = 6, 10, 12
    dx, dz = 2000.0, 500.0
    rc, rw = _density(nz)
    n = nx * ny * nz
    idx = np.arange(n).reshape(nz, ny, nx)
    rows, cols, vals = [], [], []

    def add(r, c, v):
        rows.append(r); cols.append(c); vals.append(v)

    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                r = idx[k, j, i]
                diag = 0.0
                for di, dj in ((1, 0), (-1, 0), (0, 1), (0, -1)):
                    ii, jj = i + di, j + dj
                    if 0 <= ii < nx and 0 <= jj < ny:
                        add(r, idx[k, jj, ii], rc[k] / dx ** 2)
                        diag -= rc[k] / dx ** 2
                    else:  # antisymmetric ghost: phi_ghost = -phi
                        diag -= 2 * rc[k] / dx ** 2
                if k > 0:
                    add(r, idx[k - 1, j, i], rw[k] / dz ** 2)
                    diag -= rw[k] / dz ** 2
                if k < nz - 1:
                    add(r, idx[k + 1, j, i], rw[k + 1] / dz ** 2)
                    diag -= rw[k + 1] / dz ** 2
                add(r, r, diag)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    rhs = np.random.default_rng(2).standard_normal(n)
    x = spla.spsolve(A.tocsc(), rhs).reshape(nz, ny, nx)
    phi = _oracle.solve(nx, ny, nz, dx, dz, rc, rw, rhs.reshape(nz, ny, nx), "f64")
    assert rel_l2(phi, x) < 1e-10


def test_oracle_f32ref_is_fp32_accurate_on_well_conditioned_modes():
    """The reference arithmetic: FP32 everywhere; its error is the FP32 Thomas on near-singular
    low modes (cond ~ 4/(|lambda| dz^2)) -- well below that bound, not zero."""
    nz, ny, nx = 10, 32, 32
    dx, dz = 2000.0, 500.0
    rc, rw = _density(nz)
    rhs = np.random.default_rng(3).standard_normal((nz, ny, nx)).astype(np.float32)
    r64 = _oracle.solve(nx, ny, nz, dx, dz, rc, rw, rhs.astype(np.float64), "f64")
    r32 = _oracle.solve(nx, ny, nz, dx, dz, rc, rw, rhs, "f32ref")
    lam0 = 2 * 4 / dx ** 2 * np.sin(np.pi / (2 * nx)) ** 2
    cond = 4 / (lam0 * dz ** 2)
    e = rel_l2(r32, r64)
    assert 1e-9 < e < cond * 6e-8 * 10
