"""Python mirror of the reference's body_system<T,D> call surface
(mpisph/bodies_system.h:38-379) over any library exporting the C ABI of
include/fsph.h.

`BodySystem` (the product) drives libfsph.so on a HIP device.  The same
class can drive another implementation of the ABI given its bound functions
(the tests use that to run the CPU oracle through identical calls).
"""
import ctypes as C

import numpy as np

from . import _abi


class FsphError(RuntimeError):
    def __init__(self, where, code, msg=""):
        name = _abi.ERROR_NAMES.get(code, str(code))
        super().__init__(f"{where} failed: {name} {msg}".strip())
        self.code = code


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class _System:
    """Common driver over bound ABI functions `fns` and an opaque handle."""

    def __init__(self, fns, handle, params, last_error=None):
        self._f = fns
        self._h = handle
        self.params = params
        self._last_error = last_error
        self._keep = None

    # -- plumbing -------------------------------------------------------
    def _call(self, name, *args):
        fn = self._f.get(name)
        if fn is None:
            raise NotImplementedError(name)
        rc = fn(self._h, *args)
        if rc != 0:
            msg = ""
            if self._last_error is not None:
                m = self._last_error(self._h)
                msg = m.decode() if m else ""
            raise FsphError(name, rc, msg)
        return rc

    def close(self):
        if self._h:
            self._f["destroy"](self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- data -----------------------------------------------------------
    def upload(self, parts):
        """parts: dict with x,y,z,vx,vy,vz,u,h,m (float64), id (int64),
        optional rho,P,c (the reference's read_bodies fields)."""
        n = len(parts["x"])
        keep = {}
        for k in ("x", "y", "z", "vx", "vy", "vz", "u", "h", "m", "rho", "P",
                  "c"):
            if k in parts and parts[k] is not None:
                keep[k] = np.ascontiguousarray(parts[k], dtype=np.float64)
        keep["id"] = np.ascontiguousarray(
            parts.get("id", np.arange(n)), dtype=np.int64)
        s = _abi.FsphParticles()
        s.n = n
        for k in ("x", "y", "z", "vx", "vy", "vz", "u", "h", "m", "rho", "P",
                  "c"):
            setattr(s, k, _dptr(keep.get(k)))
        s.id = keep["id"].ctypes.data_as(C.POINTER(C.c_int64))
        self._keep = keep
        self._call("upload", C.byref(s))
        self._keep = None

    def count(self):
        nt, nr = C.c_int64(), C.c_int64()
        self._call("count", C.byref(nt), C.byref(nr))
        return nt.value, nr.value

    def get(self, field):
        n, _ = self.count()
        idx = _abi.FIELD_INDEX[field]
        if field == "id":
            out = np.empty(n, dtype=np.int64)
        elif field == "key":
            out = np.empty(n, dtype=np.uint64)
        else:
            out = np.empty(n, dtype=np.float64)
        self._call("download", idx, out.ctypes.data_as(C.c_void_p))
        return out

    def set(self, field, values):
        """Overwrites one field of every particle (current order)."""
        v = np.ascontiguousarray(values, dtype=np.float64)
        n, _ = self.count()
        if v.shape != (n,):
            raise ValueError(f"{field}: expected {n} values")
        self._call("set_field", _abi.FIELD_INDEX[field], _dptr(v))

    def get_all(self, fields=None):
        fields = fields or [f for f in _abi.FIELDS]
        return {f: self.get(f) for f in fields}

    # -- the per-step pipeline -------------------------------------------
    def update_iteration(self):
        self._call("update_iteration")

    def eos_init(self):
        self._call("eos_init")

    def density_eos(self):
        self._call("density_eos")

    def acceleration(self):
        self._call("acceleration")

    def gravity(self):
        self._call("gravity")

    def dudt(self):
        self._call("dudt")

    def dedt(self):
        self._call("dedt")

    def write_bodies(self, prefix, separate_iterations=False):
        """body_system::write_bodies: appends a Step#k to <prefix>.h5part
        (a directory of raw arrays with JSON attributes)."""
        self._call("write_bodies", str(prefix).encode(),
                   int(bool(separate_iterations)))

    def read_bodies(self, prefix, start_iteration=0):
        """body_system::read_bodies from <prefix>.h5part."""
        self._call("read_bodies", str(prefix).encode(), int(start_iteration))

    def compute_alpha(self):
        self._call("compute_alpha")

    def initialize_alpha(self):
        self._call("initialize_alpha")

    def set_total_energy(self):
        self._call("set_total_energy")

    def integrate(self, what):
        self._call("integrate", int(what))

    def update_h(self):
        self._call("update_h")

    def timestep(self):
        d = C.c_double()
        self._call("timestep", C.byref(d))
        return d.value

    def advance_time(self):
        self._call("advance_time")

    def step(self):
        self._call("step")

    def check_state(self):
        """check_nans + check_negativity (default_physics.h:918-979)."""
        self._call("check_state")

    def set_checks(self, on=True):
        self._call("set_checks", int(bool(on)))

    def time(self):
        it, dt, tt = C.c_int64(), C.c_double(), C.c_double()
        self._call("time", C.byref(it), C.byref(dt), C.byref(tt))
        return it.value, dt.value, tt.value

    def set_dt(self, dt):
        self._call("set_dt", float(dt))

    def reductions(self):
        out = np.zeros(10)
        self._call("reductions", _dptr(out))
        return out

    def gravity_stats(self):
        out = np.zeros(8, np.int64)
        self._call("gravity_stats",
                   out.ctypes.data_as(C.POINTER(C.c_int64)))
        keys = ("cells", "buckets", "m2l_pairs", "p2p_pairs",
                "p2p_interactions", "pair_levels", "builder",
                "pool_chunks")
        return dict(zip(keys, (int(v) for v in out)))

    def set_fmm_builder(self, builder, stack_limit=0):
        """builder: _abi.FMM_WALK or _abi.FMM_PAIRS."""
        self._call("set_fmm_builder", int(builder), int(stack_limit))

    def set_fmm_capacity(self, cap_q, cap_m2l, cap_p2p):
        self._call("set_fmm_capacity", int(cap_q), int(cap_m2l), int(cap_p2p))

    def reductions_noalloc(self):
        out = np.zeros(10)
        self._call("reductions_noalloc", _dptr(out))
        return out

    def range(self):
        lo, hi = np.zeros(3), np.zeros(3)
        self._call("get_range", _dptr(lo), _dptr(hi))
        return lo, hi

    def tree(self):
        nn, ng = C.c_int64(), C.c_int64()
        z = None
        self._call("get_tree", C.byref(nn), z, z, z, z, C.byref(ng), z, z)
        key = np.empty(nn.value, np.uint64)
        st = np.empty(nn.value, np.int32)
        cnt = np.empty(nn.value, np.int32)
        par = np.empty(nn.value, np.int32)
        gs = np.empty(ng.value, np.int32)
        gc = np.empty(ng.value, np.int32)

        def p(a, t):
            return a.ctypes.data_as(C.POINTER(t))

        self._call("get_tree", C.byref(nn), p(key, C.c_uint64),
                   p(st, C.c_int32), p(cnt, C.c_int32), p(par, C.c_int32),
                   C.byref(ng), p(gs, C.c_int32), p(gc, C.c_int32))
        return dict(key=key, start=st, count=cnt, parent=par, group_start=gs,
                    group_count=gc)

    def neighbor_lists(self):
        """(off, idx): neighbours of sorted particle i are idx[off[i]:off[i+1]]
        as sorted particle indices."""
        n, _ = self.count()
        off = np.empty(n + 1, np.int64)
        pi64 = C.POINTER(C.c_int64)
        self._call("neighbor_lists", off.ctypes.data_as(pi64), None, 0)
        tot = int(off[n])
        idx = np.empty(max(tot, 1), np.int64)
        self._call("neighbor_lists", off.ctypes.data_as(pi64),
                   idx.ctypes.data_as(pi64), tot)
        return off, idx[:tot]


class BodySystem(_System):
    """One rank's share of particles on a HIP device (libfsph.so)."""

    def __init__(self, params, device=0):
        from . import _lib
        lib = _lib.load()
        fns = _abi.bind(lib, "fsph_")
        h = C.c_void_p()
        lib.fsph_create.restype = C.c_int
        lib.fsph_create.argtypes = [C.POINTER(_abi.FsphParams), C.c_int,
                                    C.POINTER(C.c_void_p)]
        rc = lib.fsph_create(C.byref(params), int(device), C.byref(h))
        if rc != 0:
            raise FsphError("fsph_create", rc)
        lib.fsph_last_error.restype = C.c_char_p
        lib.fsph_last_error.argtypes = [C.c_void_p]
        super().__init__(fns, h, params, lib.fsph_last_error)
        self._lib = lib

    def enable_timing(self, on=True):
        self._lib.fsph_enable_timing.argtypes = [C.c_void_p, C.c_int]
        rc = self._lib.fsph_enable_timing(self._h, int(bool(on)))
        if rc:
            raise FsphError("fsph_enable_timing", rc)

    def phase_times(self):
        out = (C.c_float * len(_abi.TIMERS))()
        self._lib.fsph_phase_times.argtypes = [C.c_void_p, C.c_void_p]
        rc = self._lib.fsph_phase_times(self._h, out)
        if rc:
            raise FsphError("fsph_phase_times", rc)
        return dict(zip(_abi.TIMERS, list(out)))

    # -- learned smoothing-length predictor (libfsph.so only) ------------
    def mlp_num_params(self):
        self._lib.fsph_mlp_num_params.restype = C.c_int
        self._lib.fsph_mlp_num_params.argtypes = []
        return int(self._lib.fsph_mlp_num_params())

    def mlp_params(self):
        p = np.zeros(self.mlp_num_params(), np.float32)
        self._call("mlp_get_params", _fptr(p), p.size)
        return p

    def set_mlp_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self._call("mlp_set_params", _fptr(p), p.size)

    def mlp_forward(self, x):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 6)
        y = np.zeros(len(x), np.float32)
        self._call("mlp_forward", _fptr(x), len(x), _fptr(y))
        return y

    def mlp_gradient(self, x, t):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 6)
        t = np.ascontiguousarray(t, np.float32).reshape(-1)
        g = np.zeros(self.mlp_num_params(), np.float32)
        loss = C.c_float()
        self._call("mlp_gradient", _fptr(x), _fptr(t), len(x), _fptr(g),
                   C.byref(loss))
        return g, float(loss.value)

    def mlp_train(self, x, t, iters, lr):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 6)
        t = np.ascontiguousarray(t, np.float32).reshape(-1)
        loss = C.c_float()
        self._call("mlp_train", _fptr(x), _fptr(t), len(x), int(iters),
                   float(lr), C.byref(loss))
        return float(loss.value)

    def h_gather(self):
        """(gather radius factor, gather mask words) of the last h update."""
        k, w = C.c_double(), C.c_int64()
        self._call("h_gather", C.byref(k), C.byref(w))
        return k.value, w.value

    def h_stats(self):
        fb, cl = C.c_int64(), C.c_int64()
        ts, loss = C.c_int32(), C.c_float()
        self._call("h_stats", C.byref(fb), C.byref(cl), C.byref(ts),
                   C.byref(loss))
        return {"fallback": fb.value, "clamped": cl.value,
                "trained_steps": ts.value, "last_loss": loss.value}


class BatchSystem(BodySystem):
    """n_inst independent simulations of one parameter set in one context
    (fsph_create_batch): one launch per kernel per step for all of them."""

    def __init__(self, params, n_inst, device=0):
        from . import _lib
        lib = _lib.load()
        fns = _abi.bind(lib, "fsph_")
        h = C.c_void_p()
        lib.fsph_create_batch.restype = C.c_int
        lib.fsph_create_batch.argtypes = [C.POINTER(_abi.FsphParams), C.c_int,
                                          C.c_int, C.POINTER(C.c_void_p)]
        rc = lib.fsph_create_batch(C.byref(params), int(device), int(n_inst),
                                   C.byref(h))
        if rc != 0:
            raise FsphError("fsph_create_batch", rc)
        lib.fsph_last_error.restype = C.c_char_p
        lib.fsph_last_error.argtypes = [C.c_void_p]
        _System.__init__(self, fns, h, params, lib.fsph_last_error)
        self._lib = lib
        self.n_inst = int(n_inst)

    def upload(self, parts_list):
        """parts_list: one particle dict per instance (ids in [0, n_i))."""
        if len(parts_list) != self.n_inst:
            raise ValueError("need one particle set per instance")
        arr = (_abi.FsphParticles * self.n_inst)()
        keep = []
        for i, parts in enumerate(parts_list):
            n = len(parts["x"])
            d = {}
            for k in ("x", "y", "z", "vx", "vy", "vz", "u", "h", "m", "rho",
                      "P", "c"):
                if k in parts and parts[k] is not None:
                    d[k] = np.ascontiguousarray(parts[k], dtype=np.float64)
            d["id"] = np.ascontiguousarray(parts.get("id", np.arange(n)),
                                           dtype=np.int64)
            keep.append(d)
            arr[i].n = n
            for k in ("x", "y", "z", "vx", "vy", "vz", "u", "h", "m", "rho",
                      "P", "c"):
                setattr(arr[i], k, _dptr(d.get(k)))
            arr[i].id = d["id"].ctypes.data_as(C.POINTER(C.c_int64))
        self._call("upload_batch", arr, self.n_inst)

    def instance_count(self, i):
        nt, nr = C.c_int64(), C.c_int64()
        self._call("instance_count", int(i), C.byref(nt), C.byref(nr))
        return nt.value, nr.value

    def instance_get(self, i, field):
        n, _ = self.instance_count(i)
        if field == "id":
            out = np.empty(n, dtype=np.int64)
        elif field == "key":
            out = np.empty(n, dtype=np.uint64)
        else:
            out = np.empty(n, dtype=np.float64)
        self._call("instance_download", int(i), _abi.FIELD_INDEX[field],
                   out.ctypes.data_as(C.c_void_p))
        return out

    def enable_graph(self, on=True):
        self._call("enable_graph", int(bool(on)))

    def instance_set_periodic(self, i, periodic):
        self._call("instance_set_periodic", int(i), int(bool(periodic)))

    def instance_time(self, i):
        it, dt, t = C.c_int64(), C.c_double(), C.c_double()
        self._call("instance_time", int(i), C.byref(it), C.byref(dt),
                   C.byref(t))
        return it.value, dt.value, t.value


def set_tuning(system, tile_rows, chunk):
    """Applies (tile_rows, chunk), e.g. autotune's result, to a system:
    tile_rows (64 or 128) rows of one group per pair-kernel workgroup, chunk
    candidates staged in LDS per round.  The neighbour lists do not depend on
    either (they follow the system's group_max, the reference's grouping)."""
    system._call("set_tuning", int(tile_rows), int(chunk))


def autotune(system, steps=2):
    """Runs the library's autotuner on a BodySystem or BatchSystem; returns
    (tile_rows, chunk)."""
    g, ch = C.c_int32(), C.c_int32()
    system._call("autotune", int(steps), C.byref(g), C.byref(ch))
    return g.value, ch.value
