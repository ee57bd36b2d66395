"""Reader of the raw container written by umwm_hip_write_output and
umwm_hip_write_restart (layout in umwm_amd/csrc/rawio.hpp): the variables
of the reference's umwmout_*.nc / umwmrst_*.nc with the same names,
dimension names (fastest first) and units."""
import struct

import numpy as np

MAGIC = b"UMWMRAW1"
_TYPES = {1: np.float32, 2: np.float64, 3: np.int32}


class Var:
    def __init__(self, name, units, description, dims, data):
        self.name, self.units, self.description = name, units, description
        self.dims = dims          # fastest first, as the reference's NetCDF dims
        self.data = data          # numpy array, Fortran order, shape = sizes of dims

    def __repr__(self):
        return f"Var({self.name}, dims={self.dims}, shape={self.data.shape}, units={self.units!r})"


def read(path):
    """Returns {name: Var} in file order."""
    out = {}
    with open(path, "rb") as f:
        if f.read(8) != MAGIC:
            raise ValueError(f"{path} is not a UMWMRAW1 file")
        (nv,) = struct.unpack("<i", f.read(4))

        def s():
            (n,) = struct.unpack("<i", f.read(4))
            return f.read(n).decode()

        for _ in range(nv):
            name, units, desc = s(), s(), s()
            typ, nd = struct.unpack("<ii", f.read(8))
            dims, shape = [], []
            for _ in range(nd):
                dims.append(s())
                (n,) = struct.unpack("<q", f.read(8))
                shape.append(n)
            dt = np.dtype(_TYPES[typ]).newbyteorder("<")
            count = int(np.prod(shape)) if shape else 1
            a = np.frombuffer(f.read(count * dt.itemsize), dtype=dt, count=count)
            out[name] = Var(name, units, desc, dims, a.reshape(shape, order="F"))
    return out
