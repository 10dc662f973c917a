"""The C-ABI library loads without a GPU and exports every entry point of include/cans_hip.h."""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "cans_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cans_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from cans_amd import _lib as LB
    L = LB.lib()
    names = header_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
        assert n in LB.SIGNATURES, "no ctypes signature for " + n


def test_host_setup_matches_oracle():
    """The library's host helpers (initgrid, tridmatrix, eigenvalues) agree with the oracle."""
    from cans_amd import host as H
    from oracle import oracle as O
    from cans_amd import _lib as LB
    L = LB.lib()
    for gtype, gr, per in ((1, 1.8, 0), (2, 1.2, 0), (3, 0.7, 0), (4, 2.0, 0), (1, 0.0, 1)):
        a = [np.zeros(34) for _ in range(4)]
        L.cans_initgrid(gtype, 32, gr, 1.3, per, *[H._dp(x) for x in a])
        b = O.initgrid(gtype, 32, gr, 1.3, per)
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)
    dzc, dzf, zc, zf = O.initgrid(1, 24, 1.5, 1.0, False)
    for bc in ("NN", "DD", "PP", "ND"):
        for cf in "cf":
            np.testing.assert_array_equal(np.array(H.tridmatrix(bc, 24, 1 / dzc, 1 / dzf, cf)),
                                          np.array(O.tridmatrix(bc, 24, 1 / dzc, 1 / dzf, cf)))
    for n in (2, 7, 8, 48, 64):
        for bc in ("NN", "DD", "ND"):
            np.testing.assert_array_equal(H.eigenvalues(n, bc, "c"), O.eigenvalues(n, bc, "c"))
        # periodic: the device ordering is the packed real-FFT order (r0, r_{n/2}, r1, i1, ...)
        cpu = O.eigenvalues(n, "PP", "c")
        dev = H.eigenvalues(n, "PP", "c")
        if n % 2 == 0:
            ref = np.empty(n)
            ref[0], ref[1] = cpu[0], cpu[n // 2]
            for k in range(1, n // 2):
                ref[2 * k] = cpu[k]
                ref[2 * k + 1] = cpu[n - k]
            np.testing.assert_array_equal(dev, ref)


def test_errors_are_reported():
    from cans_amd import _lib as LB
    L = LB.lib()
    h = ctypes.c_void_p()
    rc = L.cans_ctx_create(LB.ivec((0, 4, 4)), 0, 0, None, ctypes.byref(h))
    assert rc == LB.CANS_EINVAL
    assert b"sizes" in L.cans_last_error()
    assert L.cans_eigenvalues(4, b"PN", b"c", None) == LB.CANS_EINVAL
