"""ctypes driver for the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker of the HIP path and the CPU baseline
leg of bench.py.  It exposes the same call surface as
flecsph_amd.system.BodySystem so tests run identical call sequences on both.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from flecsph_amd import _abi
from flecsph_amd.system import FsphError, _System

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    """Compiles liboracle.so (and oracle/_ref/libref.so when the reference
    tree is present)."""
    subprocess.check_call(["make", "-s", "-C", _HERE, "-B", "liboracle.so"])
    if os.path.isdir("/root/reference/include/tree_topology"):
        subprocess.check_call(["make", "-s", "-C", _HERE, "-B", "ref"])


def load():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = C.CDLL(path)
        _LIB.fo_create.restype = C.c_void_p
        _LIB.fo_create.argtypes = [C.POINTER(_abi.FsphParams)]
        _LIB.fo_set_threads.argtypes = [C.c_int]
        _LIB.fo_morton_key.restype = C.c_uint64
        _LIB.fo_morton_key.argtypes = [C.POINTER(C.c_double)] * 2 + [
            C.c_double] * 3
        _LIB.fo_kernel.restype = C.c_double
        _LIB.fo_kernel.argtypes = [C.c_int, C.c_double, C.c_double]
        _LIB.fo_kernel_gradient.argtypes = [C.c_int, C.POINTER(C.c_double),
                                            C.c_double,
                                            C.POINTER(C.c_double)]
        _LIB.fo_direct_gravity.argtypes = [C.c_int64] + [
            C.POINTER(C.c_double)] * 4 + [C.c_double] + [
            C.POINTER(C.c_double)] * 4
        pd = C.POINTER(C.c_double)
        _LIB.fo_sphere_box.argtypes = [pd, pd, pd, C.c_double]
        _LIB.fo_box_box.argtypes = [pd, pd, pd, pd]
        _LIB.fo_within_distance2.argtypes = [pd, pd, C.c_double]
        _LIB.fo_gravity_engine.argtypes = [C.c_void_p]
        _LIB.fo_get_node_data.argtypes = [C.c_void_p] + [
            C.POINTER(C.c_double)] * 6
    return _LIB


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OracleSystem(_System):
    def __init__(self, params, threads=1):
        lib = load()
        lib.fo_set_threads(int(threads))
        h = lib.fo_create(C.byref(params))
        if not h:
            raise FsphError("fo_create", _abi.FSPH_E_OOM)
        super().__init__(_abi.bind(lib, "fo_"), C.c_void_p(h), params)

    def gravity_engine(self):
        rc = load().fo_gravity_engine(self._h)
        if rc:
            raise FsphError("fo_gravity_engine", rc)

    def node_data(self):
        t = self.tree()
        nn = len(t["key"])
        com = np.empty(3 * nn); mass = np.empty(nn); rad = np.empty(nn)
        lap = np.empty(nn); bmin = np.empty(3 * nn); bmax = np.empty(3 * nn)
        rc = load().fo_get_node_data(self._h, _d(com), _d(mass), _d(rad),
                                     _d(lap), _d(bmin), _d(bmax))
        if rc:
            raise FsphError("fo_get_node_data", rc)
        return dict(com=com.reshape(-1, 3), mass=mass, radius=rad, lap=lap,
                    bmin=bmin.reshape(-1, 3), bmax=bmax.reshape(-1, 3))


def set_threads(n):
    load().fo_set_threads(int(n))


def morton_key(lo, hi, x, y, z):
    lo = np.ascontiguousarray(lo, np.float64)
    hi = np.ascontiguousarray(hi, np.float64)
    return int(load().fo_morton_key(_d(lo), _d(hi), x, y, z))


def kernel(kind, r, h):
    return load().fo_kernel(int(kind), float(r), float(h))


def kernel_gradient(kind, v, h):
    v = np.ascontiguousarray(v, np.float64)
    out = np.zeros(3)
    load().fo_kernel_gradient(int(kind), _d(v), float(h), _d(out))
    return out


def direct_gravity(x, y, z, m, G=1.0):
    n = len(x)
    x, y, z, m = (np.ascontiguousarray(a, np.float64) for a in (x, y, z, m))
    gx, gy, gz, phi = (np.empty(n) for _ in range(4))
    load().fo_direct_gravity(n, _d(x), _d(y), _d(z), _d(m), float(G), _d(gx),
                             _d(gy), _d(gz), _d(phi))
    return gx, gy, gz, phi


def _dp(v):
    return np.ascontiguousarray(v, np.float64).ctypes.data_as(
        C.POINTER(C.c_double))


def sphere_box(mn, mx, c, r):
    a, b, cc = (np.ascontiguousarray(v, np.float64) for v in (mn, mx, c))
    return bool(load().fo_sphere_box(_d(a), _d(b), _d(cc), float(r)))


def box_box(mn1, mx1, mn2, mx2):
    arrs = [np.ascontiguousarray(v, np.float64) for v in (mn1, mx1, mn2, mx2)]
    return bool(load().fo_box_box(*[_d(a) for a in arrs]))


def within_distance2(p1, p2, r):
    a, b = (np.ascontiguousarray(v, np.float64) for v in (p1, p2))
    return bool(load().fo_within_distance2(_d(a), _d(b), float(r)))
