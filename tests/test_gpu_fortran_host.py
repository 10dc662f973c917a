"""The Fortran host (cans_amd/fortran/drive_step.f90: main.f90's call sequence
through module cans_hip) against the Python host's fused substep and the oracle.

State goes in and out through the reference's checkpoint format, so this also
exercises the drop-in path a CaNS build would take: namelist + restart files ->
Fortran host -> C ABI -> HIP kernels.
"""
import os
import subprocess

import numpy as np
import pytest

from cans_amd import initflow as IF
from cans_amd import restart
from tests import cases

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cans_amd", "fortran", "drive_step")


def fstr(chars):
    return ", ".join("'%s'" % c for c in chars)


def write_nml(path, p, dt, nstep):
    with open(path, "w") as f:
        f.write("&dns\n")
        f.write("ng = %d, %d, %d\n" % tuple(p["ng"]))
        f.write("l = %r, %r, %r\n" % tuple(float(x) for x in p["l"]))
        f.write("gtype = %d\ngr = %r\nvisci = %r\ndt = %r\nnstep = %d\n" % (p["gtype"], p["gr"], p["visci"], dt, nstep))
        f.write("cbcvel = %s\n" % fstr(p["cbcvel"]))
        f.write("cbcpre = %s\n" % fstr(p["cbcpre"]))
        f.write("bcvel = %s\n" % ", ".join(repr(float(x)) for x in p.get("bcvel", (0.0,) * 18)))
        f.write("is_forced = %s\n" % ", ".join("T" if x else "F" for x in p.get("is_forced", (0, 0, 0))))
        f.write("velf = %s\n" % ", ".join(repr(float(x)) for x in p.get("velf", (0, 0, 0))))
        f.write("/\n")


@pytest.mark.parametrize("name,ng,dt,nstep", [("channel", (64, 32, 48), 2e-4, 2), ("tgv", (32, 32, 32), 1e-3, 2)])
def test_fortran_host_matches_python_host_and_oracle(tmp_path, name, ng, dt, nstep):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    assert os.path.exists(EXE), "Fortran driver not built (run __graft_entry__.build())"
    from cans_amd.host import Simulation
    print("\n[case] fortran host %s ng=%s" % (name, ng), flush=True)
    p = cases.params(name, ng=ng)
    sim = Simulation(cases.host_case(p))
    u, v, w, pr = cases.initial_state(name, p, sim.grid, seed=3)
    for k, q in zip("uvwp", (u, v, w, pr)):
        restart.write_field(tmp_path / ("fld_%s.bin" % k), q)
    write_nml(tmp_path / "case.nml", p, dt, nstep)
    out = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0
    # Python host, fused substeps, same inputs
    sim.set_state(*(IF.halo(p["ng"], q) for q in (u, v, w, pr)))
    sim.init_bounds()
    for _ in range(nstep):
        sim.rk3_step(dt)
    sim.ctx.sync()
    # oracle
    orc = cases.oracle_case(p)
    st = orc.new_state()
    for k, q in zip("uvwp", (u, v, w, pr)):
        st[k][1:-1, 1:-1, 1:-1] = q
    orc.init_bounds(st)
    for _ in range(nstep):
        orc.rk3_step(st, dt)
    for k in "uvwp":
        q, t, i = restart.read_field(tmp_path / ("out_%s.bin" % k), p["ng"])
        assert i == nstep and abs(t - nstep * dt) < 1e-15
        mine = getattr(sim, k).interior()
        # the per-routine Fortran sequence and the fused substep run the same arithmetic
        np.testing.assert_array_equal(q, mine, err_msg="fortran vs fused " + k)
        ref = st[k][1:-1, 1:-1, 1:-1]
        if k != "p":
            assert np.abs(q - ref).max() <= 1e-12 * np.abs(ref).max()
