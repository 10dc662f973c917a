"""The reference's single-field checkpoint format (src/load.f90:461-589, MPI-IO backend).

A file holds ng(1)*ng(2)*ng(3) reals of the interior in Fortran order, native
endianness, followed by two reals [time, istep].  Reading validates the size as
load.f90:494-502 does (the reference aborts on a mismatch; this raises).
"""
import numpy as np


def write_field(path, q, time=0.0, istep=0, dtype=np.float64):
    q = np.asarray(q)
    with open(path, "wb") as f:
        np.asfortranarray(q, dtype=dtype).ravel(order="F").tofile(f)
        np.array([time, float(istep)], dtype=dtype).tofile(f)


def read_field(path, ng, dtype=np.float64):
    ng = tuple(int(x) for x in ng)
    good = (int(np.prod(ng)) + 2) * np.dtype(dtype).itemsize
    data = np.fromfile(path, dtype=dtype)
    if data.nbytes != good:
        raise ValueError("checkpoint %s has %d bytes, expected %d for ng=%s" % (path, data.nbytes, good, ng))
    q = data[:-2].reshape(ng, order="F")
    return q, float(data[-2]), int(round(float(data[-1])))
