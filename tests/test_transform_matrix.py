"""The dense transform of a direction whose length the FFT passes do not take
(cans_transform_matrix; the matrix k_dense applies).

Pinned three ways on the CPU, for lengths the FFT passes refuse (primes, m < 3, many odd
factors, odd n1) and for FFT lengths:
  * against the oracle's FFTW-definition r2r of the kind fft.f90:260-313 names, with
    R2HC/HC2R in the packed order of the reference's GPU build (fft.f90:402-451);
  * backward after forward is the normalisation factor of fft.f90:75,140 times the
    identity (the face-centred DD boundary face passes through);
  * the forward matrix diagonalises the second-difference operator of the boundary pair
    with the eigenvalues cans_eigenvalues returns, in the same order (initsolver.f90:85-144),
    so the z solve pairs each output with its own eigenvalue.  A transposed index or a
    swapped slot in the packed order fails this check.
"""
import ctypes as C

import numpy as np
import pytest

from cans_amd import _lib as LB
from oracle import oracle as O

LENGTHS = [1, 2, 3, 5, 7, 12, 14, 16, 45, 49]


def matrix(n, bc, c_or_f, direction):
    m = np.zeros(n * n)
    rc = LB.lib().cans_transform_matrix(n, bc.encode(), c_or_f.encode(), direction, m.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == LB.CANS_OK, LB.lib().cans_last_error()
    return m.reshape(n, n)


def packed(hc):
    """FFTW halfcomplex (r0, r1, .., r_{n/2}, i_{(n+1)/2-1}, .., i1) -> the GPU build's packed order."""
    n = hc.shape[0]
    out = np.zeros_like(hc)
    h = n // 2
    out[0] = hc[0]
    if n == 1:
        return out
    if n % 2 == 0:
        out[1] = hc[h]
        for k in range(1, h):
            out[2 * k], out[2 * k + 1] = hc[k], hc[n - k]
    else:
        out[1] = hc[n - h]
        for k in range(1, h):
            out[2 * k], out[2 * k + 1] = hc[k], hc[n - k]
        out[2 * h] = hc[h]
    return out


def unpacked(pk):
    n = pk.shape[0]
    hc = np.zeros_like(pk)
    h = n // 2
    hc[0] = pk[0]
    if n == 1:
        return hc
    if n % 2 == 0:
        hc[h] = pk[1]
        for k in range(1, h):
            hc[k], hc[n - k] = pk[2 * k], pk[2 * k + 1]
    else:
        hc[n - h] = pk[1]
        for k in range(1, h):
            hc[k], hc[n - k] = pk[2 * k], pk[2 * k + 1]
        hc[h] = pk[2 * h]
    return hc


# (bc, c_or_f) -> (forward kind, backward kind, normalisation factor as a function of n)
KINDS = {("PP", "c"): (O.R2HC, O.HC2R, lambda n: n), ("NN", "c"): (O.REDFT10, O.REDFT01, lambda n: 2 * n),
         ("DD", "c"): (O.RODFT10, O.RODFT01, lambda n: 2 * n), ("ND", "c"): (O.REDFT11, O.REDFT11, lambda n: 2 * n),
         ("DN", "c"): (O.RODFT11, O.RODFT11, lambda n: 2 * n), ("PP", "f"): (O.R2HC, O.HC2R, lambda n: n),
         ("DD", "f"): (O.RODFT00, O.RODFT00, lambda n: 2 * n), ("ND", "f"): (O.REDFT10, O.REDFT01, lambda n: 2 * n),
         ("DN", "f"): (O.RODFT01, O.RODFT10, lambda n: 2 * n)}


def oracle_apply(kind, x, bc, c_or_f):
    if (bc, c_or_f) == ("DD", "f"):  # RODFT00 on the n-1 interior faces (fft.f90:80-82)
        y = x.copy()
        if x.shape[0] > 1:
            y[:-1] = O.r2r(kind, x[:-1], axis=0)
        return y
    if kind == O.HC2R:
        return O.r2r(kind, unpacked(x), axis=0)
    y = O.r2r(kind, x, axis=0)
    return packed(y) if kind == O.R2HC else y


@pytest.mark.parametrize("n", LENGTHS)
@pytest.mark.parametrize("bc,c_or_f", sorted(KINDS), ids=["%s-%s" % k for k in sorted(KINDS)])
def test_matrix_is_the_oracle_transform(n, bc, c_or_f):
    if (bc, c_or_f) == ("DD", "f") and n < 2:
        pytest.skip("face-centred DD needs 2 points")
    kf, kb, nfac = KINDS[(bc, c_or_f)]
    rng = np.random.default_rng(n)
    x = rng.uniform(-1, 1, (n, 3))
    F, B = matrix(n, bc, c_or_f, 1), matrix(n, bc, c_or_f, -1)
    yo = oracle_apply(kf, x, bc, c_or_f)
    np.testing.assert_allclose(F @ x, yo, rtol=0, atol=1e-13 * max(1, n))
    np.testing.assert_allclose(B @ yo, oracle_apply(kb, yo, bc, c_or_f), rtol=0, atol=1e-12 * max(1, n))
    eye = nfac(n) * np.eye(n)
    if (bc, c_or_f) == ("DD", "f"):
        eye[-1, -1] = 1.0
    np.testing.assert_allclose(B @ F, eye, rtol=0, atol=1e-12 * n * n)


def second_difference(n, bc, c_or_f):
    """The unit-spacing 1-D Laplacian the reference's eigenvalues belong to (initsolver.f90:85-144)."""
    if (bc, c_or_f) == ("DD", "f"):  # interior faces 1..n-1, both ends Dirichlet; the boundary face apart
        m = n - 1
        A = np.diag(-2.0 * np.ones(m)) + np.diag(np.ones(m - 1), 1) + np.diag(np.ones(m - 1), -1)
        out = np.zeros((n, n))
        out[:m, :m] = A
        return out
    A = np.diag(-2.0 * np.ones(n)) + np.diag(np.ones(n - 1), 1) + np.diag(np.ones(n - 1), -1)
    if bc == "PP":
        A[0, -1] += 1.0
        A[-1, 0] += 1.0
        return A
    ghost = {"N": 1.0, "D": -1.0}  # cell-centred: the ghost mirrors (N) or negates (D) the first cell
    A[0, 0] += ghost[bc[0]]
    A[-1, -1] += ghost[bc[1]]
    return A


EIGEN_KINDS = [("PP", "c"), ("NN", "c"), ("DD", "c"), ("ND", "c"), ("DN", "c"), ("PP", "f"), ("DD", "f")]


@pytest.mark.parametrize("n", [2, 3, 5, 7, 12, 14, 45])
@pytest.mark.parametrize("bc,c_or_f", EIGEN_KINDS, ids=["%s-%s" % k for k in EIGEN_KINDS])
def test_matrix_diagonalises_the_laplacian_in_eigenvalue_order(n, bc, c_or_f):
    F = matrix(n, bc, c_or_f, 1)
    lam = np.zeros(n)
    assert LB.lib().cans_eigenvalues(n, bc.encode(), c_or_f.encode(),
                                     lam.ctypes.data_as(C.POINTER(C.c_double))) == LB.CANS_OK
    A = second_difference(n, bc, c_or_f)
    lhs, rhs = F @ A, lam[:, None] * F
    if (bc, c_or_f) == ("DD", "f"):  # the boundary face row is the identity, outside the operator
        lhs, rhs = lhs[:-1], rhs[:-1]
    np.testing.assert_allclose(lhs, rhs, rtol=0, atol=1e-12 * n * n)


def test_swapped_packed_slots_fail_the_eigenvalue_check():
    """Negative control: the PP matrix with the slots of r1 and r2 swapped (or r_{n/2} moved off
    slot 1) no longer pairs each row with its eigenvalue."""
    n = 12
    F = matrix(n, "PP", "c", 1)
    F[[2, 4]] = F[[4, 2]]
    lam = np.zeros(n)
    LB.lib().cans_eigenvalues(n, b"PP", b"c", lam.ctypes.data_as(C.POINTER(C.c_double)))
    assert np.abs(F @ second_difference(n, "PP", "c") - lam[:, None] * F).max() > 1e-3
    F = matrix(n, "PP", "c", 1)
    F[[1, 2]] = F[[2, 1]]
    assert np.abs(F @ second_difference(n, "PP", "c") - lam[:, None] * F).max() > 1e-3


def test_matrix_arguments_checked():
    m = np.zeros(4)
    p = m.ctypes.data_as(C.POINTER(C.c_double))
    assert LB.lib().cans_transform_matrix(2, b"PN", b"c", 1, p) == LB.CANS_EINVAL
    assert LB.lib().cans_transform_matrix(2, b"NN", b"f", 1, p) == LB.CANS_EUNSUPPORTED
    assert LB.lib().cans_transform_matrix(2, b"PP", b"c", 0, p) == LB.CANS_EINVAL
