"""The C-ABI library loads and exports every function include/fsph.h
declares (no device calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "fsph.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b(fsph_[a-z_0-9]+)\s*\(", src)
    return sorted(set(names))


def test_header_declares_the_pipeline():
    names = declared_functions()
    for required in ("fsph_create", "fsph_upload", "fsph_update_iteration",
                     "fsph_density_eos", "fsph_acceleration", "fsph_dudt",
                     "fsph_integrate", "fsph_update_h", "fsph_timestep",
                     "fsph_step", "fsph_neighbor_lists", "fsph_destroy"):
        assert required in names


def test_library_exports_every_declared_symbol():
    path = os.path.join(ROOT, "flecsph_amd", "libfsph.so")
    if not os.path.exists(path):
        pytest.fail("libfsph.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_default_params_match_reference_defaults():
    from flecsph_amd import _abi
    lib = ctypes.CDLL(os.path.join(ROOT, "flecsph_amd", "libfsph.so"))
    p = _abi.FsphParams()
    lib.fsph_default_params(ctypes.byref(p))
    q = _abi.default_params()
    for name, _ in _abi.FsphParams._fields_:
        a, b = getattr(p, name), getattr(q, name)
        if hasattr(a, "__len__"):
            assert list(a) == list(b), name
        else:
            assert a == b, name
    # params.h: initial_dt 0.001, sph_viscosity_epsilon 0.01,
    # gravitational_constant 1, sph_eta 1.5 (params.h:212),
    # sub_entities_ 128 (tree_topology.h:2121)
    assert p.initial_dt == 0.001 and p.visc_eps == 0.01 and p.G == 1.0
    assert p.eta == 1.5
    assert p.group_max == 128


def test_create_without_device_fails_cleanly():
    """With no HIP device visible, fsph_create returns an error code instead
    of crashing (there is no CPU fallback)."""
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a device is present")
    from flecsph_amd import _abi
    lib = ctypes.CDLL(os.path.join(ROOT, "flecsph_amd", "libfsph.so"))
    p = _abi.default_params()
    h = ctypes.c_void_p()
    lib.fsph_create.argtypes = [ctypes.POINTER(_abi.FsphParams), ctypes.c_int,
                                ctypes.POINTER(ctypes.c_void_p)]
    rc = lib.fsph_create(ctypes.byref(p), 0, ctypes.byref(h))
    assert rc == _abi.FSPH_E_NODEV
