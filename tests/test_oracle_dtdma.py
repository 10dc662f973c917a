"""Pins the oracle's distributed tridiagonal solve (gaussel_dtdma, solver.f90:309-517).

The slab-partitioned elimination must give the same solution as the plain Thomas
sweep (gaussel) and as dense solves of the same systems, for 1 to 8 slabs,
including slab counts that do not divide the row count (at least three rows a slab); only the
rounding differs.
"""
import numpy as np
import pytest

from oracle import oracle as O


def systems(n, nx, ny, rng):
    a = rng.uniform(0.5, 1.5, n)
    c = rng.uniform(0.5, 1.5, n)
    b = -(a + c)
    lam = -rng.uniform(0.1, 4.0, (nx, ny))  # strictly diagonally dominant lines (no zero mode)
    return a, b, c, np.asfortranarray(lam)


@pytest.mark.parametrize("n,nslab", [(24, 1), (24, 2), (24, 3), (25, 4), (48, 8), (25, 8)])
def test_dtdma_matches_gaussel_and_dense(n, nslab):
    rng = np.random.default_rng(n * 10 + nslab)
    nx, ny = 5, 3
    a, b, c, lam = systems(n, nx, ny, rng)
    norm = 0.7
    rhs = np.asfortranarray(rng.uniform(-1, 1, (nx, ny, n)))
    p1 = rhs.copy(order="F")
    p2 = rhs.copy(order="F")
    O.lib().orc_gaussel(nx, ny, n, O.P(a), O.P(b), O.P(c), 0, norm, O.P(p1), O.VP(lam))
    rc = O.lib().orc_gaussel_dtdma(nx, ny, n, nslab, O.P(a), O.P(b), O.P(c), norm, O.P(p2), O.VP(lam), 0)
    assert rc == 0
    np.testing.assert_allclose(p2, p1, rtol=0, atol=1e-13 * np.abs(p1).max())
    for i in range(nx):
        for j in range(ny):
            A = np.diag(b + lam[i, j]) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
            x = np.linalg.solve(A, rhs[i, j, :] * norm)
            np.testing.assert_allclose(p2[i, j, :], x, rtol=0, atol=1e-12 * np.abs(x).max())


def test_dtdma_without_eigenvalues():
    rng = np.random.default_rng(5)
    n, nx, ny = 30, 4, 2
    a = rng.uniform(0.5, 1.0, n)
    c = rng.uniform(0.5, 1.0, n)
    b = -(a + c) - rng.uniform(1.0, 2.0, n)
    rhs = np.asfortranarray(rng.uniform(-1, 1, (nx, ny, n)))
    p1 = rhs.copy(order="F")
    p2 = rhs.copy(order="F")
    O.lib().orc_gaussel(nx, ny, n, O.P(a), O.P(b), O.P(c), 0, 1.0, O.P(p1), None)
    assert O.lib().orc_gaussel_dtdma(nx, ny, n, 4, O.P(a), O.P(b), O.P(c), 1.0, O.P(p2), None, 0) == 0
    np.testing.assert_allclose(p2, p1, rtol=0, atol=1e-13 * np.abs(p1).max())


def test_dtdma_rejects_thin_slabs():
    """Each slab needs three rows: with two, its boundary-row step would couple the neighbours' unknowns."""
    a = b = c = np.ones(17)
    p = np.zeros((1, 1, 17), order="F")
    assert O.lib().orc_gaussel_dtdma(1, 1, 17, 8, O.P(a), O.P(b), O.P(c), 1.0, O.P(p), None, 0) != 0
    assert O.lib().orc_gaussel_dtdma(1, 1, 17, 1, O.P(a), O.P(b), O.P(c), 1.0, O.P(p), None, 1) != 0


@pytest.mark.parametrize("n,nslab", [(24, 2), (24, 3), (25, 4), (48, 8), (25, 8), (9, 3)])
def test_dtdma_periodic_matches_gaussel_and_dense(n, nslab):
    """Periodic z: a(1) couples row 1 to row n and c(n) row n to row 1 (solver.f90:469-498).  The
    cyclic reduced system must give the periodic gaussel's solution and the dense cyclic solve's."""
    rng = np.random.default_rng(n * 100 + nslab)
    nx, ny = 4, 3
    a, b, c, lam = systems(n, nx, ny, rng)
    norm = 1.3
    rhs = np.asfortranarray(rng.uniform(-1, 1, (nx, ny, n)))
    p1 = rhs.copy(order="F")
    p2 = rhs.copy(order="F")
    O.lib().orc_gaussel(nx, ny, n, O.P(a), O.P(b), O.P(c), 1, norm, O.P(p1), O.VP(lam))
    assert O.lib().orc_gaussel_dtdma(nx, ny, n, nslab, O.P(a), O.P(b), O.P(c), norm, O.P(p2), O.VP(lam), 1) == 0
    np.testing.assert_allclose(p2, p1, rtol=0, atol=1e-13 * np.abs(p1).max())
    for i in range(nx):
        for j in range(ny):
            A = np.diag(b + lam[i, j]) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
            A[0, n - 1] = a[0]
            A[n - 1, 0] = c[n - 1]
            x = np.linalg.solve(A, rhs[i, j, :] * norm)
            np.testing.assert_allclose(p2[i, j, :], x, rtol=0, atol=1e-12 * np.abs(x).max())


def test_dtdma_periodic_dropped_wrap_term_fails():
    """The dense check above is sharp: the same systems without the wrap couplings differ at O(1)."""
    rng = np.random.default_rng(3)
    n, nx, ny = 24, 2, 2
    a, b, c, lam = systems(n, nx, ny, rng)
    rhs = np.asfortranarray(rng.uniform(-1, 1, (nx, ny, n)))
    p2 = rhs.copy(order="F")
    assert O.lib().orc_gaussel_dtdma(nx, ny, n, 3, O.P(a), O.P(b), O.P(c), 1.0, O.P(p2), O.VP(lam), 1) == 0
    A = np.diag(b + lam[0, 0]) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    x = np.linalg.solve(A, rhs[0, 0, :])
    assert np.abs(p2[0, 0, :] - x).max() > 1e-3 * np.abs(x).max()
