"""ctypes mirror of include/fsph.h (structs, constants, prototypes).

The C ABI is the product boundary; this module only describes it so Python
callers (tests, bench, the BodySystem wrapper) can bind any library that
exports it with a given symbol prefix.
"""
import ctypes as C

FSPH_OK = 0
FSPH_E_ARG = -1
FSPH_E_HIP = -2
FSPH_E_OOM = -3
FSPH_E_DUPKEY = -4
FSPH_E_NAN = -5
FSPH_E_NEG = -6
FSPH_E_CAP = -7
FSPH_E_NODEV = -8

ERROR_NAMES = {
    FSPH_E_ARG: "FSPH_E_ARG", FSPH_E_HIP: "FSPH_E_HIP", FSPH_E_OOM: "FSPH_E_OOM",
    FSPH_E_DUPKEY: "FSPH_E_DUPKEY", FSPH_E_NAN: "FSPH_E_NAN",
    FSPH_E_NEG: "FSPH_E_NEG", FSPH_E_CAP: "FSPH_E_CAP",
    FSPH_E_NODEV: "FSPH_E_NODEV",
}

KERNEL_CUBIC_SPLINE = 0
KERNEL_WENDLAND_C6 = 1
EOS_IDEAL = 0
EOS_POLYTROPIC = 1
PREC_FP32 = 0
PREC_FP64 = 1
GRAV_NONE = 0
GRAV_FMM = 1
FMM_WALK = 0
FMM_PAIRS = 1

H_CLOSED = 0
H_NEWTON = 1
H_LEARNED = 2

KICK_V = 1
KICK_U = 2
SAVE_VHALF = 4
DRIFT = 8
KICK_E = 16

VISC_CONSTANT = 0
VISC_CULLEN = 1

FIELDS = ["x", "y", "z", "vx", "vy", "vz", "vhx", "vhy", "vhz", "u", "h", "m",
          "rho", "P", "c", "vsig", "ax", "ay", "az", "dudt", "gx", "gy", "gz",
          "phi", "id", "key", "nngb", "hngb", "hiter", "e", "dedt", "alpha",
          "divv"]
FIELD_INDEX = {name: i for i, name in enumerate(FIELDS)}

TIMERS = ["integrate", "boundary", "range_keys", "sort", "tree", "traverse",
          "masks", "density", "vsig", "accel", "gravity", "dudt", "h", "dt",
          "graph", "grav_pairs", "m2l", "p2p"]


class FsphParams(C.Structure):
    _fields_ = [
        ("kernel", C.c_int32), ("eos", C.c_int32), ("precision", C.c_int32),
        ("adaptive_dt", C.c_int32), ("gravity", C.c_int32),
        ("group_max", C.c_int32), ("leaf_size", C.c_int32),
        ("periodic", C.c_int32 * 3), ("n_target", C.c_int32),
        ("fmm_order", C.c_int32),
        ("gamma", C.c_double), ("poly_K", C.c_double), ("eta", C.c_double),
        ("visc_alpha", C.c_double), ("visc_beta", C.c_double),
        ("visc_eps", C.c_double), ("cfl", C.c_double),
        ("initial_dt", C.c_double), ("G", C.c_double), ("theta", C.c_double),
        ("box", C.c_double * 3),
        ("h_mode", C.c_int32), ("newton_maxit", C.c_int32),
        ("newton_tol", C.c_double), ("h_search", C.c_double),
        ("learn_steps", C.c_int32), ("n_tol", C.c_int32),
        ("thermokinetic", C.c_int32), ("viscosity", C.c_int32),
        ("visc_alpha_max", C.c_double), ("visc_l", C.c_double),
        ("visc_delta", C.c_double),
        ("evolve_internal_energy", C.c_int32),
    ]


class FsphParticles(C.Structure):
    _fields_ = [("n", C.c_int64)] + [
        (f, C.POINTER(C.c_double)) for f in
        ("x", "y", "z", "vx", "vy", "vz", "u", "h", "m")
    ] + [("id", C.POINTER(C.c_int64))] + [
        (f, C.POINTER(C.c_double)) for f in ("rho", "P", "c")
    ]


def default_params(**kw):
    """The reference's defaults (include/params.h DECLARE_PARAM values)."""
    p = FsphParams()
    p.kernel = KERNEL_CUBIC_SPLINE
    p.eos = EOS_IDEAL
    p.precision = PREC_FP32
    p.adaptive_dt = 0
    p.gravity = GRAV_NONE
    p.group_max = 128
    p.leaf_size = 32
    p.fmm_order = 2
    p.n_target = 100
    p.gamma = 1.4
    p.poly_K = 1.0
    p.eta = 1.5
    p.visc_alpha = 1.0
    p.visc_beta = 2.0
    p.visc_eps = 0.01
    p.cfl = 0.25
    p.initial_dt = 0.001
    p.G = 1.0
    p.theta = 0.0
    p.h_mode = H_CLOSED
    p.newton_maxit = 20
    p.newton_tol = 1e-4
    p.h_search = 1.3
    p.learn_steps = 50
    p.n_tol = 5
    p.thermokinetic = 1
    p.viscosity = VISC_CONSTANT
    p.visc_alpha_max = 1.0
    p.visc_l = 0.05
    p.visc_delta = 1.0
    p.evolve_internal_energy = 1
    for k in range(3):
        p.periodic[k] = 0
        p.box[k] = 1.0
    for k, v in kw.items():
        if k in ("periodic", "box"):
            for i in range(3):
                getattr(p, k)[i] = v[i]
        else:
            setattr(p, k, v)
    return p


_P = C.c_void_p
_PD = C.POINTER(C.c_double)
_PI64 = C.POINTER(C.c_int64)
_PU64 = C.POINTER(C.c_uint64)
_PI32 = C.POINTER(C.c_int32)
_PF = C.POINTER(C.c_float)

# name -> (restype, argtypes); "ctx" functions take the handle first.
PROTOTYPES = {
    "destroy": (None, [_P]),
    "upload": (C.c_int, [_P, C.POINTER(FsphParticles)]),
    "count": (C.c_int, [_P, _PI64, _PI64]),
    "download": (C.c_int, [_P, C.c_int, _P]),
    "update_iteration": (C.c_int, [_P]),
    "eos_init": (C.c_int, [_P]),
    "density_eos": (C.c_int, [_P]),
    "acceleration": (C.c_int, [_P]),
    "gravity": (C.c_int, [_P]),
    "dudt": (C.c_int, [_P]),
    "integrate": (C.c_int, [_P, C.c_uint32]),
    "update_h": (C.c_int, [_P]),
    "timestep": (C.c_int, [_P, _PD]),
    "advance_time": (C.c_int, [_P]),
    "step": (C.c_int, [_P]),
    "check_state": (C.c_int, [_P]),
    "set_checks": (C.c_int, [_P, C.c_int]),
    "time": (C.c_int, [_P, _PI64, _PD, _PD]),
    "set_dt": (C.c_int, [_P, C.c_double]),
    "reductions": (C.c_int, [_P, _PD]),
    "reductions_noalloc": (C.c_int, [_P, _PD]),
    "dedt": (C.c_int, [_P]),
    "h_gather": (C.c_int, [_P, _PD, C.POINTER(C.c_int64)]),
    "instance_set_periodic": (C.c_int, [_P, C.c_int, C.c_int]),
    "write_bodies": (C.c_int, [_P, C.c_char_p, C.c_int]),
    "read_bodies": (C.c_int, [_P, C.c_char_p, C.c_int64]),
    "compute_alpha": (C.c_int, [_P]),
    "initialize_alpha": (C.c_int, [_P]),
    "set_field": (C.c_int, [_P, C.c_int, _PD]),
    "set_total_energy": (C.c_int, [_P]),
    "gravity_stats": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "set_fmm_capacity": (C.c_int, [_P, C.c_int64, C.c_int64, C.c_int64]),
    "set_fmm_builder": (C.c_int, [_P, C.c_int, C.c_int]),
    "get_range": (C.c_int, [_P, _PD, _PD]),
    "get_tree": (C.c_int, [_P, _PI64, _PU64, _PI32, _PI32, _PI32, _PI64,
                           _PI32, _PI32]),
    "neighbor_lists": (C.c_int, [_P, _PI64, _PI64, C.c_int64]),
    "mlp_get_params": (C.c_int, [_P, _PF, C.c_int64]),
    "mlp_set_params": (C.c_int, [_P, _PF, C.c_int64]),
    "mlp_forward": (C.c_int, [_P, _PF, C.c_int64, _PF]),
    "mlp_gradient": (C.c_int, [_P, _PF, _PF, C.c_int64, _PF, _PF]),
    "mlp_train": (C.c_int, [_P, _PF, _PF, C.c_int64, C.c_int, C.c_float,
                            _PF]),
    "h_stats": (C.c_int, [_P, _PI64, _PI64, _PI32, _PF]),
    "upload_batch": (C.c_int, [_P, C.POINTER(FsphParticles), C.c_int]),
    "instance_count": (C.c_int, [_P, C.c_int, _PI64, _PI64]),
    "instance_download": (C.c_int, [_P, C.c_int, C.c_int, _P]),
    "instance_time": (C.c_int, [_P, C.c_int, _PI64, _PD, _PD]),
    "autotune": (C.c_int, [_P, C.c_int, _PI32, _PI32]),
    "set_tuning": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "enable_graph": (C.c_int, [_P, C.c_int]),
}


def bind(lib, prefix, names=None):
    """Returns {short_name: c_function} for the prototypes the library has."""
    out = {}
    for name, (res, args) in PROTOTYPES.items():
        if names is not None and name not in names:
            continue
        fn = getattr(lib, prefix + name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
        out[name] = fn
    return out
