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


def test_fortran_module_binds_every_declared_symbol():
    """module cans_hip (cans_amd/fortran/cans_hip_iface.f90) has a bind(C) interface for each
    entry point of the header, under the C name."""
    txt = open(os.path.join(ROOT, "cans_amd", "fortran", "cans_hip_iface.f90")).read().lower()
    bound = set(re.findall(r"bind\(c,\s*name='(cans_[a-z0-9_]+)'\)", txt))
    missing = [n for n in header_functions() if n not in bound]
    assert not missing, missing


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
    for n, per in ((64, 0), (33, 0), (48, 1)):  # the natural channel grid (is_gridpoint_natural_channel)
        a = [np.zeros(n + 2) for _ in range(4)]
        L.cans_initgrid_natural(n, 1.7, per, *[H._dp(x) for x in a])
        for x, y in zip(a, O.initgrid_natural(n, 1.7, per)):
            np.testing.assert_allclose(x, y, rtol=1e-15, atol=0)
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
    # entry points added beside the reference's: argument checks before any device work
    box = (LB.PoissonBox * 1)()
    assert L.cans_poisson_batch_create(None, 1, box, None, ctypes.byref(h)) == LB.CANS_EINVAL
    assert b"cans_poisson_batch_create" in L.cans_last_error()
    assert L.cans_poisson_batch_solve(None) == LB.CANS_EINVAL
    assert L.cans_poisson_batch_destroy(None) == LB.CANS_OK
    assert L.cans_rk3_substep_adjoint(None, None, 1, 1e-3, *([None] * 16)) == LB.CANS_EINVAL
    assert b"cans_rk3_substep_adjoint" in L.cans_last_error()
    assert L.cans_mom_jt(None, None, None, 1.0, *([None] * 12)) == LB.CANS_EINVAL
    assert L.cans_solver_refined(None, None, None, None, 1.0, None, None, None, None, b"PPPPNN", b"ccc", None,
                                 None, None, 2, None) == LB.CANS_EINVAL


def test_bc_rhs_matches_oracle_and_ghost_cells():
    """cans_bc_rhs (initsolver.f90:189-232) against the oracle, and against the ghost-cell
    form of the boundary conditions: with x_0 / x_{n+1} set from the boundary values, the
    full 1-D second difference equals tridmatrix's operator applied to x minus the
    boundary rhs added by updt_rhs_b (bound.f90:514-598)."""
    import ctypes as C
    from cans_amd import host as H
    from oracle import oracle as O
    n = 12
    dzc, dzf, zc, zf = O.initgrid(1, n, 1.4, 1.0, False)
    dzci, dzfi = 1 / dzc, 1 / dzf
    rng = np.random.default_rng(3)
    for cbc in ("DD", "NN", "DN", "ND", "PP"):
        for cf in "cf":
            bv = rng.uniform(-2, 2, 2)
            if cf == "c":
                dlc, dlf = [dzc[0], dzc[n]], [dzf[1], dzf[n]]
            else:
                dlc, dlf = [dzc[1], dzc[n - 1]], [dzf[1], dzf[n]]
            got = H.bc_rhs(cbc, bv, dlc, dlf, cf)
            ref = np.zeros(2)
            O.lib().orc_bc_rhs(cbc.encode(), bv.ctypes.data_as(C.POINTER(C.c_double)),
                               np.asarray(dlc).ctypes.data_as(C.POINTER(C.c_double)),
                               np.asarray(dlf).ctypes.data_as(C.POINTER(C.c_double)), cf.encode(),
                               ref.ctypes.data_as(C.POINTER(C.c_double)))
            np.testing.assert_array_equal(got, ref)
            if cbc == "PP":
                assert not got.any()
                continue
            a, b, c = (np.array(t) for t in H.tridmatrix(cbc, n, dzci, dzfi, cf))
            x = rng.uniform(-1, 1, n + 2)  # x[1..m] unknowns; x[0], x[m+1] ghosts or wall faces
            if cf == "c":
                m = n
                # centred ghosts: D: x0 = 2v - x1; N: x0 = x1 - dzc0*v, x_{n+1} = x_n + dzc_n*v
                x[0] = 2 * bv[0] - x[1] if cbc[0] == "D" else x[1] - dzc[0] * bv[0]
                x[n + 1] = 2 * bv[1] - x[n] if cbc[1] == "D" else x[n] + dzc[n] * bv[1]

                def full(k):
                    return ((x[k + 1] - x[k]) * dzci[k] - (x[k] - x[k - 1]) * dzci[k - 1]) * dzfi[k]
                ends = ((1, 0), (m, 1))
            else:
                # faces: a D wall face holds the boundary value (the last face is then not
                # solved, solver.f90:77); face-centred N ends are covered by the oracle equality
                m = n - 1 if cbc[1] == "D" else n
                if cbc[0] == "D":
                    x[0] = bv[0]
                if cbc[1] == "D":
                    x[n] = bv[1]

                def full(k):
                    return ((x[k + 1] - x[k]) * dzfi[k + 1] - (x[k] - x[k - 1]) * dzfi[k]) * dzci[k]
                ends = tuple(e for e in ((1, 0), (m, 1)) if cbc[e[1]] == "D")
            for k, ib in ends:
                t = b[k - 1] * x[k] + (a[k - 1] * x[k - 1] if k > 1 else 0.) + (c[k - 1] * x[k + 1] if k < m else 0.)
                assert abs(full(k) - (t - got[ib])) <= 1e-9 * max(1.0, abs(full(k))), (cbc, cf, k, full(k), t - got[ib])
