"""Pins the oracle's restarted GMRES (gmres.cpp:10-164) on small dense
systems against LAPACK, and its edge behaviour."""
import numpy as np
import pytest


def test_matches_lapack(oracle):
    rng = np.random.default_rng(0)
    n = 60
    A = np.eye(n) + 0.3 * rng.normal(size=(n, n)) / np.sqrt(n)
    b = rng.normal(size=n)
    x, rc, it, res = oracle.gmres_dense(A, b, restart=15, maxit=200, tol=1e-12)
    assert rc == 0 and res < 1e-12
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=0, atol=1e-9)


def test_full_krylov_space_converges_in_n_steps(oracle):
    rng = np.random.default_rng(1)
    n = 12
    A = rng.normal(size=(n, n)) + 4 * np.eye(n)
    b = rng.normal(size=n)
    x, rc, it, res = oracle.gmres_dense(A, b, restart=n, maxit=n, tol=1e-10)
    assert rc == 0 and it <= n
    np.testing.assert_allclose(A @ x, b, atol=1e-8)


def test_identity_one_step(oracle):
    b = np.arange(1.0, 8.0)
    x, rc, it, res = oracle.gmres_dense(np.eye(7), b, restart=5, maxit=10, tol=1e-12)
    assert rc == 0 and it == 1
    np.testing.assert_allclose(x, b, rtol=1e-15)


def test_zero_rhs_and_iteration_limit(oracle):
    rng = np.random.default_rng(2)
    A = np.eye(30) + rng.normal(size=(30, 30))
    x, rc, it, res = oracle.gmres_dense(A, np.zeros(30), restart=5, maxit=10, tol=1e-10)
    assert rc == 0 and it == 0 and np.all(x == 0)  # |b| = 0 is taken as 1
    x, rc, it, res = oracle.gmres_dense(A, rng.normal(size=30), restart=5, maxit=1, tol=1e-14)
    assert rc == 1 and it == 1


def test_restarts_reduce_residual(oracle):
    rng = np.random.default_rng(3)
    n = 80
    A = np.eye(n) + 0.5 * rng.normal(size=(n, n)) / np.sqrt(n)
    b = rng.normal(size=n)
    _, _, _, r1 = oracle.gmres_dense(A, b, restart=4, maxit=4, tol=1e-14)
    _, _, _, r2 = oracle.gmres_dense(A, b, restart=4, maxit=16, tol=1e-14)
    assert r2 < r1 < 1.0
