"""ctypes access to oracle/_ref/libref.so: the reference's own Morton curve
and geometric predicates compiled from its headers (oracle/ref_shim.cc).
TEST INFRASTRUCTURE ONLY; present only where the reference tree was
available at build time."""
import ctypes as C
import os

import numpy as np

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref",
                     "libref.so")
_LIB = None


def available():
    return os.path.exists(_PATH)


def load():
    global _LIB
    if _LIB is None:
        _LIB = C.CDLL(_PATH)
        pd = C.POINTER(C.c_double)
        _LIB.ref_morton_keys.argtypes = [C.c_int64, pd, pd, pd, pd, pd,
                                         C.POINTER(C.c_uint64)]
        _LIB.ref_key_coordinates.argtypes = [C.c_uint64, pd, pd, pd]
        for f in ("ref_key_root", "ref_key_min", "ref_key_max"):
            getattr(_LIB, f).restype = C.c_uint64
        _LIB.ref_within_distance2.argtypes = [pd, pd, C.c_double]
        _LIB.ref_intersects_box_box.argtypes = [pd, pd, pd, pd]
        _LIB.ref_intersects_sphere_box.argtypes = [pd, pd, pd, C.c_double]
        _LIB.ref_mac.argtypes = [pd, C.c_double, pd, C.c_double, C.c_double]
    return _LIB


def _d(a):
    return np.ascontiguousarray(a, np.float64).ctypes.data_as(
        C.POINTER(C.c_double))


def morton_keys(lo, hi, x, y, z):
    x, y, z = (np.ascontiguousarray(a, np.float64) for a in (x, y, z))
    lo = np.ascontiguousarray(lo, np.float64)
    hi = np.ascontiguousarray(hi, np.float64)
    out = np.empty(len(x), np.uint64)
    load().ref_morton_keys(len(x), _d(lo), _d(hi), _d(x), _d(y), _d(z),
                           out.ctypes.data_as(C.POINTER(C.c_uint64)))
    return out


def key_coordinates(key, lo, hi):
    out = np.zeros(3)
    lo = np.ascontiguousarray(lo, np.float64)
    hi = np.ascontiguousarray(hi, np.float64)
    load().ref_key_coordinates(C.c_uint64(int(key)), _d(lo), _d(hi),
                               out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def sphere_box(mn, mx, c, r):
    a, b, cc = (np.ascontiguousarray(v, np.float64) for v in (mn, mx, c))
    return bool(load().ref_intersects_sphere_box(
        a.ctypes.data_as(C.POINTER(C.c_double)),
        b.ctypes.data_as(C.POINTER(C.c_double)),
        cc.ctypes.data_as(C.POINTER(C.c_double)), float(r)))


def box_box(mn1, mx1, mn2, mx2):
    arrs = [np.ascontiguousarray(v, np.float64) for v in (mn1, mx1, mn2, mx2)]
    return bool(load().ref_intersects_box_box(
        *[a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs]))


def within_distance2(p1, p2, r):
    a, b = (np.ascontiguousarray(v, np.float64) for v in (p1, p2))
    return bool(load().ref_within_distance2(
        a.ctypes.data_as(C.POINTER(C.c_double)),
        b.ctypes.data_as(C.POINTER(C.c_double)), float(r)))
