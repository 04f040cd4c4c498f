"""The on-disk step layout (flecsph_amd/h5raw.py) on the host: a written
step reads back with the reference's dataset names, dtypes and step
attributes (doc/H5_STRUCTURE.md: id int64, everything else double)."""
import numpy as np
import pytest

from flecsph_amd import generators, h5raw


def test_write_read_roundtrip(tmp_path):
    parts, _ = generators.random_box(300, seed=4)
    path = str(tmp_path / "out.h5part")
    data = {k: parts[k] for k in ("x", "y", "z", "vx", "vy", "vz", "u", "h",
                                  "m", "id")}
    assert h5raw.write_step(path, data, time=0.5, iteration=7,
                            timestep=1e-3) == 0
    assert h5raw.write_step(path, data, time=0.6, iteration=8) == 1
    assert h5raw.steps(path) == [0, 1]
    attr, got = h5raw.read_step(path, 1)
    assert attr == {"time": 0.6, "iteration": 8, "timestep": 0.0,
                    "nparticles": 300}
    assert got["id"].dtype == np.int64 and got["x"].dtype == np.float64
    for k, v in data.items():
        assert np.array_equal(got[k], v), k
    with pytest.raises(ValueError):
        h5raw.write_step(path, {"x": np.zeros(3), "y": np.zeros(4)})
