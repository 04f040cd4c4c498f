"""read_bodies / write_bodies of the library (the H5part step layout of
doc/H5_STRUCTURE.md without HDF5, flecsph_amd/csrc/fsph_io.hip): what the
device writes is the state it holds, with the reference's dataset names;
a step written by numpy (flecsph_amd/h5raw.py) uploads exactly; a restart
from a step restores its time, dt and iteration."""
import os

import numpy as np
import pytest

from flecsph_amd import BodySystem, FsphError, _abi, generators, h5raw

pytestmark = pytest.mark.gpu


def test_write_bodies_matches_the_device_state(tmp_path):
    parts, prm = generators.sedov(n_side=12)
    prm = dict(prm, precision=_abi.PREC_FP64)
    g = BodySystem(_abi.default_params(**prm))
    g.upload(parts)
    prefix = str(tmp_path / "sedov")
    for _ in range(3):
        g.step()
        g.write_bodies(prefix)
    path = prefix + ".h5part"
    assert h5raw.steps(path) == [0, 1, 2]
    attr, d = h5raw.read_step(path, 2)
    it, dt, t = g.time()
    assert attr["iteration"] == it and attr["timestep"] == dt
    assert attr["time"] == t
    n_tot, n_real = g.count()
    assert attr["nparticles"] == n_tot  # WALL copies included
    for name, field in (("x", "x"), ("vy", "vy"), ("h", "h"), ("rho", "rho"),
                        ("u", "u"), ("P", "P"), ("m", "m"), ("ax", "ax")):
        assert np.array_equal(d[name], g.get(field)), name
    assert np.array_equal(d["id"], g.get("id"))
    assert np.array_equal(d["type"], (g.get("id") >= n_real).astype(np.int32))
    # one output per file
    g.write_bodies(prefix + "_sep", separate_iterations=True)
    assert os.path.isdir(prefix + "_sep_00003.h5part")


def test_read_bodies_uploads_a_step_and_restarts(tmp_path):
    parts, prm = generators.random_box(3000, seed=21, h_spread=0.5)
    path = str(tmp_path / "in.h5part")
    data = {k: parts[k] for k in ("x", "y", "z", "vx", "vy", "vz", "u", "h",
                                  "m", "id")}
    h5raw.write_step(path, data, time=0.0, iteration=0)
    h5raw.write_step(path, data, time=0.25, iteration=40, timestep=2e-4)
    p = _abi.default_params(**prm)
    a, b = BodySystem(p), BodySystem(p)
    a.upload(parts)
    b.read_bodies(str(tmp_path / "in"))
    for f in ("x", "vz", "u", "h", "m"):
        assert np.array_equal(a.get(f), b.get(f)), f
    assert np.array_equal(a.get("id"), b.get("id"))
    for s in (a, b):
        s.step()
    for f in ("x", "rho", "ax"):
        assert np.array_equal(a.get(f), b.get(f)), f
    c = BodySystem(p)
    c.read_bodies(str(tmp_path / "in"), start_iteration=40)
    it, dt, t = c.time()
    assert (it, dt, t) == (40, 2e-4, 0.25)
    c.step()  # the driver's initial iteration at iteration 40
    assert c.time()[0] == 41
    with pytest.raises(FsphError):
        c.read_bodies(str(tmp_path / "in"), start_iteration=41)
