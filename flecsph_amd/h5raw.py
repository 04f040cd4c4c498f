"""The H5part step layout of the reference (doc/H5_STRUCTURE.md,
mpisph/io.h:632-1081) as the library stores it without HDF5: a directory
<prefix>.h5part holding header.json and one Step#<k> directory per output
with attributes.json and one raw little-endian array per dataset
(<name>.f8 double, .i8 int64, .i4 int32).  This module reads and writes
that layout with numpy, for post-processing and for preparing inputs that
fsph_read_bodies loads."""
import json
import os

import numpy as np

_DTYPES = {".f8": np.dtype("<f8"), ".i8": np.dtype("<i8"),
           ".i4": np.dtype("<i4")}


def steps(path):
    """Step numbers present in a .h5part directory, ascending."""
    out = []
    k = 0
    while os.path.isdir(os.path.join(path, f"Step#{k}")):
        out.append(k)
        k += 1
    return out


def read_step(path, step=0):
    """(attributes, {dataset: array}) of Step#step."""
    sd = os.path.join(path, f"Step#{step}")
    with open(os.path.join(sd, "attributes.json")) as f:
        attr = json.load(f)
    data = {}
    for name in sorted(os.listdir(sd)):
        base, ext = os.path.splitext(name)
        if ext in _DTYPES:
            data[base] = np.fromfile(os.path.join(sd, name), _DTYPES[ext])
    return attr, data


def write_step(path, data, time=0.0, iteration=0, timestep=0.0):
    """Appends a step: `data` maps dataset names to arrays (float arrays
    are stored as .f8, id as .i8, type as .i4)."""
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, "header.json"), "w") as f:
        json.dump({"format": "fsph-h5raw-1", "dimension": 3}, f)
    k = len(steps(path))
    sd = os.path.join(path, f"Step#{k}")
    os.makedirs(sd)
    n = None
    for name, arr in data.items():
        a = np.asarray(arr)
        ext = ".i8" if name == "id" else ".i4" if name == "type" else ".f8"
        a = a.astype(_DTYPES[ext])
        if n is None:
            n = len(a)
        elif len(a) != n:
            raise ValueError(f"dataset {name} holds {len(a)} values, not {n}")
        a.tofile(os.path.join(sd, name + ext))
    with open(os.path.join(sd, "attributes.json"), "w") as f:
        json.dump({"time": float(time), "iteration": int(iteration),
                   "timestep": float(timestep), "nparticles": int(n)}, f)
    return k
