"""Seeded initial conditions (host side), restating src/initflow.f90 shapes.

The reference seeds noise with the Fortran runtime's random_number stream
(initflow.f90:347-377), which no other runtime reproduces; identical seeded
inputs are therefore generated here (numpy Philox, keyed by seed) and handed
to both the oracle and the device path.
"""
import numpy as np


def rng(seed):
    return np.random.Generator(np.random.Philox(key=int(seed)))


def tgv(n, dl, l, zc, uref=1.0):
    """Taylor-Green vortex, initflow.f90:102-117 (interior values, Fortran order)."""
    n1, n2, n3 = n
    i = np.arange(1, n1 + 1, dtype=np.float64)
    j = np.arange(1, n2 + 1, dtype=np.float64)
    xc = (i - .5) * dl[0] / l[0] * 2. * np.pi
    xf = (i - .0) * dl[0] / l[0] * 2. * np.pi
    yc = (j - .5) * dl[1] / l[1] * 2. * np.pi
    yf = (j - .0) * dl[1] / l[1] * 2. * np.pi
    zcc = zc[1:n3 + 1] / l[2] * 2. * np.pi
    u = np.sin(xf)[:, None, None] * np.cos(yc)[None, :, None] * np.cos(zcc)[None, None, :] * uref
    v = -np.cos(xc)[:, None, None] * np.sin(yf)[None, :, None] * np.cos(zcc)[None, None, :] * uref
    w = np.zeros((n1, n2, n3))
    return np.asfortranarray(u), np.asfortranarray(v), np.asfortranarray(w)


def poiseuille(zc_over_l, norm):
    """initflow.f90:422-437: 6 z (1-z) norm."""
    z = zc_over_l
    return 6. * z * (1. - z) * norm


def add_noise(q, generator, norm):
    """q += 2 (rn - 1/2) norm with rn uniform in [0,1) (initflow.f90:347-377)."""
    q += 2. * (generator.random(q.shape) - .5) * norm
    return q


def set_mean(q, grid_vol_ratio_k, mean):
    """initflow.f90:379-403: rescale q so that its volume-weighted mean is `mean`."""
    meanold = float((q * grid_vol_ratio_k[None, None, :]).sum())
    if abs(meanold) > np.finfo(np.float64).eps:
        q[...] = q / meanold * mean
    return q


def halo(n, interior):
    """Embed an interior (n1,n2,n3) array in a zeroed halo'd Fortran array."""
    a = np.zeros((n[0] + 2, n[1] + 2, n[2] + 2), order="F")
    a[1:-1, 1:-1, 1:-1] = interior
    return a
