"""Pins the oracle's r2r transforms (FFTW kind definitions, src/fft.f90:260-313).

Two independent checks per kind: scipy.fft (whose unnormalised DCT/DST types
I-IV and rfft coincide with FFTW's REDFTab/RODFTab/R2HC by definition) and a
brute-force O(n^2) sum written straight from FFTW's published formulas.
"""
import numpy as np
import pytest
import scipy.fft as sf

from oracle import oracle as O

SIZES = [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 15, 16, 30, 32, 48, 64, 96, 100, 128, 1024]


def brute(kind, x):
    n = x.shape[-1]
    j = np.arange(n)[None, :]
    k = np.arange(n)[:, None]
    if kind == O.R2HC:
        X = np.exp(-2j * np.pi * j * k / n) @ x
        out = np.empty(n)
        h = n // 2
        out[: h + 1] = X[: h + 1].real
        for kk in range(1, (n + 1) // 2):
            out[n - kk] = X[kk].imag
        return out
    if kind == O.HC2R:
        # x_j = sum_k X_k e^{+2 pi i jk/n} with X Hermitian, from halfcomplex
        X = np.zeros(n, complex)
        X[0] = x[0]
        for kk in range(1, (n + 1) // 2):
            X[kk] = x[kk] + 1j * x[n - kk]
            X[n - kk] = x[kk] - 1j * x[n - kk]
        if n % 2 == 0:
            X[n // 2] = x[n // 2]
        return (np.exp(2j * np.pi * j * k / n) @ X).real
    if kind == O.REDFT00:
        if n == 1:
            return x.copy()
        M = np.cos(np.pi * j[:, 1:-1] * k / (n - 1)) * 2
        return x[0] + (-1.0) ** k[:, 0] * x[-1] + M @ x[1:-1]
    if kind == O.REDFT10:
        return 2 * np.cos(np.pi * (j + 0.5) * k / n) @ x
    if kind == O.REDFT01:
        return x[0] + 2 * np.cos(np.pi * j[:, 1:] * (k + 0.5) / n) @ x[1:]
    if kind == O.REDFT11:
        return 2 * np.cos(np.pi * (j + 0.5) * (k + 0.5) / n) @ x
    if kind == O.RODFT00:
        return 2 * np.sin(np.pi * (j + 1) * (k + 1) / (n + 1)) @ x
    if kind == O.RODFT10:
        return 2 * np.sin(np.pi * (j + 0.5) * (k + 1) / n) @ x
    if kind == O.RODFT01:
        return (-1.0) ** k[:, 0] * x[-1] + 2 * np.sin(np.pi * (j[:, :-1] + 1) * (k + 0.5) / n) @ x[:-1]
    if kind == O.RODFT11:
        return 2 * np.sin(np.pi * (j + 0.5) * (k + 0.5) / n) @ x
    raise ValueError(kind)


def scipy_ref(kind, x):
    n = x.shape[-1]
    if kind == O.R2HC:
        X = np.fft.rfft(x)
        out = np.empty(n)
        out[: n // 2 + 1] = X.real
        for kk in range(1, (n + 1) // 2):
            out[n - kk] = X[kk].imag
        return out
    if kind == O.HC2R:
        X = np.zeros(n // 2 + 1, complex)
        X[0] = x[0]
        for kk in range(1, (n + 1) // 2):
            X[kk] = x[kk] + 1j * x[n - kk]
        if n % 2 == 0:
            X[n // 2] = x[n // 2]
        return np.fft.irfft(X, n) * n
    t = {O.REDFT00: 1, O.REDFT10: 2, O.REDFT01: 3, O.REDFT11: 4,
         O.RODFT00: 1, O.RODFT10: 2, O.RODFT01: 3, O.RODFT11: 4}[kind]
    if kind in (O.REDFT00, O.REDFT10, O.REDFT01, O.REDFT11):
        return sf.dct(x, type=t)
    return sf.dst(x, type=t)


KINDS = [O.R2HC, O.HC2R, O.REDFT00, O.REDFT10, O.REDFT01, O.REDFT11,
         O.RODFT00, O.RODFT10, O.RODFT01, O.RODFT11]


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("kind", KINDS, ids=[O.KIND_NAMES[k] for k in KINDS])
def test_r2r_against_scipy_and_brute_force(kind, variant):
    O.fft_variant(variant)
    rng = np.random.default_rng(1234 + kind)
    for n in SIZES:
        if kind == O.REDFT00 and n == 1:
            continue
        x = rng.uniform(-1, 1, n)
        got = O.r2r(kind, x)
        scale = max(1.0, np.abs(x).sum())
        if n <= 128:
            np.testing.assert_allclose(got, brute(kind, x), rtol=0, atol=2e-14 * scale,
                                       err_msg="%s n=%d (brute)" % (O.KIND_NAMES[kind], n))
        np.testing.assert_allclose(got, scipy_ref(kind, x), rtol=0, atol=2e-14 * scale,
                                   err_msg="%s n=%d (scipy)" % (O.KIND_NAMES[kind], n))
    O.fft_variant(0)


PAIRS = [(O.R2HC, O.HC2R, lambda n: n), (O.REDFT10, O.REDFT01, lambda n: 2 * n),
         (O.RODFT10, O.RODFT01, lambda n: 2 * n), (O.REDFT11, O.REDFT11, lambda n: 2 * n),
         (O.RODFT11, O.RODFT11, lambda n: 2 * n), (O.REDFT00, O.REDFT00, lambda n: 2 * (n - 1)),
         (O.RODFT00, O.RODFT00, lambda n: 2 * (n + 1))]


@pytest.mark.parametrize("fwd,bwd,scale", PAIRS, ids=[O.KIND_NAMES[p[0]] for p in PAIRS])
def test_round_trip_scaling(fwd, bwd, scale):
    """bwd(fwd(x)) = norm * x: the factors fftini folds into normfft (fft.f90:75,140,202-208)."""
    rng = np.random.default_rng(7)
    for n in (2, 3, 8, 48, 64, 96, 128, 512):
        x = rng.uniform(-1, 1, (5, n))
        y = O.r2r(bwd, O.r2r(fwd, x))
        np.testing.assert_allclose(y, scale(n) * x, rtol=0, atol=1e-13 * scale(n))


def test_batched_strided_matches_lines():
    rng = np.random.default_rng(3)
    a = np.asfortranarray(rng.uniform(-1, 1, (12, 10, 4)))
    for kind in (O.R2HC, O.REDFT10, O.RODFT01):
        ref = np.stack([np.stack([O.r2r(kind, a[:, j, k]) for j in range(10)], axis=1)
                        for k in range(4)], axis=2)
        b = a.copy(order="F")
        # along x: stride 1, dist 12 (howmany 40)
        O.lib().orc_r2r(kind, 12, 40, 1, 12, O.P(b))
        np.testing.assert_array_equal(b, ref)
        # along y inside each z plane: stride 12, dist 1
        c = a.copy(order="F")
        refy = np.stack([np.stack([O.r2r(kind, a[i, :, k]) for i in range(12)], axis=0)
                         for k in range(4)], axis=2)
        for k in range(4):
            sub = c[:, :, k]
            O.lib().orc_r2r(kind, 10, 12, 12, 1, O.P(c[:, :, k:k + 1]) if False else
                            O.C.cast(c.ctypes.data + 8 * 12 * 10 * k, O._dp))
        np.testing.assert_array_equal(c, refy)
