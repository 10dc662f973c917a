"""The reference's checkpoint format (src/load.f90:461-589) round-trips and is size-checked."""
import numpy as np
import pytest

from cans_amd import restart


def test_round_trip(tmp_path):
    q = np.asfortranarray(np.random.default_rng(0).uniform(-1, 1, (6, 5, 4)))
    f = tmp_path / "fld_u.bin"
    restart.write_field(f, q, time=1.25, istep=42)
    assert f.stat().st_size == (6 * 5 * 4 + 2) * 8
    raw = np.fromfile(f, dtype=np.float64)
    np.testing.assert_array_equal(raw[:120], q.ravel(order="F"))  # Fortran order, no halos
    assert raw[-2] == 1.25 and raw[-1] == 42.0
    r, t, i = restart.read_field(f, (6, 5, 4))
    np.testing.assert_array_equal(r, q)
    assert (t, i) == (1.25, 42)


def test_size_mismatch_is_rejected(tmp_path):
    f = tmp_path / "fld_v.bin"
    restart.write_field(f, np.zeros((4, 4, 4)))
    with pytest.raises(ValueError):
        restart.read_field(f, (4, 4, 5))
    with open(f, "ab") as fh:
        fh.write(b"\0" * 8)
    with pytest.raises(ValueError):
        restart.read_field(f, (4, 4, 4))
