"""Seeded initial-data generators for the benchmark problems.

The reference ships one generator per problem under app/id_generators/
(sodtube/main.cc, sedov/main.cc, collapse/main.cc); the problems here follow
the same physical setups at the sizes the benchmark names, on cubic lattices
or seeded uniform sampling, so every run sees identical inputs.  Each
returns (particles dict, params dict) ready for BodySystem.upload and
_abi.default_params(**params).
"""
import math

import numpy as np

from . import _abi

SEED_BASE = 0x5EEDF1EC5


def _eta_for(n_ngb, kernel_width=2.0):
    """eta such that (4 pi / 3) (eta kw)^3 = n_ngb: h = eta kw (m/rho)^(1/3)
    (default_physics.h:826-848) then holds n_ngb particles in its sphere."""
    return (3.0 * n_ngb / (4.0 * math.pi)) ** (1.0 / 3.0) / kernel_width


def _pack(x, y, z, vx, vy, vz, u, h, m, rho=None, P=None, c=None):
    n = len(x)
    d = dict(x=x, y=y, z=z, vx=vx, vy=vy, vz=vz, u=u, h=h, m=m,
             id=np.arange(n, dtype=np.int64))
    if rho is not None:
        d["rho"] = rho
    if P is not None:
        d["P"] = P
    if c is not None:
        d["c"] = c
    return {k: (np.ascontiguousarray(v, np.float64) if k != "id" else v)
            for k, v in d.items()}


def _lattice(nx, ny, nz, dx, x0, y0, z0):
    i = (np.arange(nx) + 0.5) * dx + x0
    j = (np.arange(ny) + 0.5) * dx + y0
    k = (np.arange(nz) + 0.5) * dx + z0
    X, Y, Z = np.meshgrid(i, j, k, indexing="ij")
    return X.ravel(), Y.ravel(), Z.ravel()


def sod(scale=1):
    """3D Sod shock tube in [0,1]x[0,0.2]^2, interface at x = 0.5: left
    rho=1, p=1 on a 60x24x24 cell-centred lattice (dx = 1/120), right
    rho=0.125, p=0.1 on 30x12x12 at 2dx (equal masses, spacing scaled by
    rho^-1/3), v = 0, gamma = 1.4, cubic spline, 60 neighbours, Monaghan
    alpha=1 beta=2, CFL 0.25.  N = 38,880 at scale 1 (`scale` multiplies
    the lattice counts per axis)."""
    gamma = 1.4
    nl = 60 * scale
    dx = 0.5 / nl
    xl, yl, zl = _lattice(nl, 24 * scale, 24 * scale, dx, 0.0, 0.0, 0.0)
    xr, yr, zr = _lattice(nl // 2, 12 * scale, 12 * scale, 2 * dx, 0.5, 0.0,
                          0.0)
    x = np.concatenate([xl, xr]); y = np.concatenate([yl, yr])
    z = np.concatenate([zl, zr])
    left = x < 0.5
    rho = np.where(left, 1.0, 0.125)
    P = np.where(left, 1.0, 0.1)
    u = P / ((gamma - 1.0) * rho)
    m = np.full(len(x), dx ** 3)
    eta = _eta_for(60)
    h = eta * 2.0 * (m / rho) ** (1.0 / 3.0)
    c = np.sqrt(gamma * (gamma - 1.0) * u)
    zero = np.zeros(len(x))
    params = dict(thermokinetic=0, kernel=_abi.KERNEL_CUBIC_SPLINE, eos=_abi.EOS_IDEAL,
                  precision=_abi.PREC_FP64, adaptive_dt=1, gamma=gamma,
                  eta=eta, visc_alpha=1.0, visc_beta=2.0, visc_eps=0.01,
                  cfl=0.25, initial_dt=1e-4, n_target=60)
    return _pack(x, y, z, zero, zero, zero, u, h, m, rho, P, c), params


def sedov(n_side=64, seed=SEED_BASE):
    """3D Sedov blast in the periodic box [-1/2,1/2]^3: n_side^3 particles
    on a cell-centred lattice, rho = 1, u = 1e-6, gamma = 5/3, Wendland C6,
    100 neighbours.  E = 1 is deposited with Gaussian weights
    w = exp(-r^2/h0^2) over particles with r < 2 h0."""
    del seed  # the lattice is deterministic
    gamma = 5.0 / 3.0
    dx = 1.0 / n_side
    x, y, z = _lattice(n_side, n_side, n_side, dx, -0.5, -0.5, -0.5)
    n = len(x)
    m = np.full(n, dx ** 3)
    eta = _eta_for(100)
    h0 = eta * 2.0 * dx
    h = np.full(n, h0)
    u = np.full(n, 1e-6)
    r2 = x * x + y * y + z * z
    inside = r2 < (2.0 * h0) ** 2
    w = np.where(inside, np.exp(-r2 / h0 ** 2), 0.0)
    u = u + 1.0 * w / (m * w.sum())
    rho = np.ones(n)
    P = (gamma - 1.0) * rho * u
    c = np.sqrt(gamma * (gamma - 1.0) * u)
    zero = np.zeros(n)
    params = dict(thermokinetic=0, kernel=_abi.KERNEL_WENDLAND_C6, eos=_abi.EOS_IDEAL,
                  precision=_abi.PREC_FP32, adaptive_dt=1, gamma=gamma,
                  eta=eta, visc_alpha=1.0, visc_beta=2.0, visc_eps=0.01,
                  cfl=0.25, initial_dt=1e-6, periodic=(1, 1, 1),
                  box=(1.0, 1.0, 1.0), n_target=100)
    return _pack(x, y, z, zero, zero, zero, u, h, m, rho, P, c), params


def sedov_batch(n_inst=64, lo=8_000, hi=200_000, seed=SEED_BASE + 7):
    """Batched workload: n_inst Sedov blasts whose particle counts are drawn
    log-uniform in [lo, hi] (seeded) and rounded to the nearest cube n_side^3.
    Returns (list of particle sets, common parameters)."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(np.log(lo), np.log(hi), n_inst))
    sides = np.clip(np.rint(np.cbrt(counts)).astype(int),
                    int(np.ceil(np.cbrt(lo))), int(np.floor(np.cbrt(hi))))
    sets, prm = [], None
    for ns in sides:
        parts, prm = sedov(n_side=int(ns))
        sets.append(parts)
    return sets, prm


def sod_tube(j=6, gamma=5.0 / 3.0, kernel=_abi.KERNEL_WENDLAND_C6,
             n_ngb=100):
    """A Sod tube for a batch of the Sedov parameter set: [0,1]x[0,0.2]^2,
    interface at x = 0.5, left rho=1, p=1 on a 10j x 4j x 4j cell-centred
    lattice (dx = 0.05/j), right rho=0.125, p=0.1 on 5j x 2j x 2j at 2dx
    (equal masses), v = 0; 180 j^3 particles, open boundaries.  Returns the
    particle set only (the batch holds the parameters)."""
    dx = 0.05 / j
    xl, yl, zl = _lattice(10 * j, 4 * j, 4 * j, dx, 0.0, 0.0, 0.0)
    xr, yr, zr = _lattice(5 * j, 2 * j, 2 * j, 2 * dx, 0.5, 0.0, 0.0)
    x = np.concatenate([xl, xr]); y = np.concatenate([yl, yr])
    z = np.concatenate([zl, zr])
    left = x < 0.5
    rho = np.where(left, 1.0, 0.125)
    P = np.where(left, 1.0, 0.1)
    u = P / ((gamma - 1.0) * rho)
    m = np.full(len(x), dx ** 3)
    h = _eta_for(n_ngb) * 2.0 * (m / rho) ** (1.0 / 3.0)
    c = np.sqrt(gamma * (gamma - 1.0) * u)
    zero = np.zeros(len(x))
    return _pack(x, y, z, zero, zero, zero, u, h, m, rho, P, c)


def mixed_batch(n_inst=64, lo=8_000, hi=200_000, seed=SEED_BASE + 7):
    """Batched workload of Sedov blasts and Sod tubes: per instance a seeded
    coin picks the kind and a seeded log-uniform draw in [lo, hi] the
    particle count, rounded to the nearest lattice of that kind (n_side^3
    for a blast, 180 j^3 for a tube).  All share the Sedov parameter set
    (Wendland C6, gamma 5/3, 100 neighbours, FP32); blasts are periodic in
    the unit box, tubes have open boundaries.  Returns (list of particle
    sets, common parameters, per-instance periodic flags)."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(np.log(lo), np.log(hi), n_inst))
    kinds = rng.random(n_inst) < 0.5  # True: Sod tube
    sets, periodic, prm = [], [], None
    for cnt, tube in zip(counts, kinds):
        if tube:
            j = int(np.clip(np.rint(np.cbrt(cnt / 180.0)),
                            np.ceil(np.cbrt(lo / 180.0)),
                            np.floor(np.cbrt(hi / 180.0))))
            sets.append(sod_tube(j))
            periodic.append(False)
        else:
            ns = int(np.clip(np.rint(np.cbrt(cnt)), np.ceil(np.cbrt(lo)),
                             np.floor(np.cbrt(hi))))
            parts, prm = sedov(n_side=ns)
            sets.append(parts)
            periodic.append(True)
    if prm is None:
        prm = sedov(n_side=8)[1]
    return sets, prm, periodic


def _polytrope_mass_inverse(q):
    """r with M(r) = q for M(r) = [sin(pi r) - pi r cos(pi r)] / pi."""
    lo = np.zeros_like(q)
    hi = np.ones_like(q)
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        Mm = (np.sin(np.pi * mid) - np.pi * mid * np.cos(np.pi * mid)) / np.pi
        below = Mm < q
        lo = np.where(below, mid, lo)
        hi = np.where(below, hi, mid)
    return 0.5 * (lo + hi)


def _uniform_ball(n, rng):
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    r = rng.random(n) ** (1.0 / 3.0)
    return v, r


def polytrope(n=1_000_000, seed=SEED_BASE + 3):
    """Self-gravitating n=1 polytrope, R = 1, M = 1, G = 1: seeded uniform
    points in the unit ball stretched radially to rho(r) = sin(pi r)/(4 r)
    (M(r) = [sin(pi r) - pi r cos(pi r)]/pi), K = 1, gamma = 2, u = K rho,
    v = 0, Wendland C6 with 100 neighbours, FMM theta = 0.5, leaf 32."""
    rng = np.random.default_rng(seed)
    dirs, r_u = _uniform_ball(n, rng)
    r = _polytrope_mass_inverse(r_u ** 3)
    pos = dirs * r[:, None]
    m = np.full(n, 1.0 / n)
    rho = np.sin(np.pi * r) / (4.0 * np.maximum(r, 1e-12))
    rho_floor = np.maximum(rho, 0.05 * np.pi / 4.0)
    eta = _eta_for(100)
    h = eta * 2.0 * (m / rho_floor) ** (1.0 / 3.0)
    K, gamma = 1.0, 2.0
    u = K * rho_floor
    P = K * rho_floor ** gamma
    c = np.sqrt(K * gamma * rho_floor ** (gamma - 1.0))
    zero = np.zeros(n)
    # the internal energy is not evolved: with a polytropic EOS it does not
    # enter the pressure, and the reference's polytropic-star setups turn it
    # off (data/mesa_nx20.par:52, data/implosion_nx20.par:25)
    params = dict(thermokinetic=0, evolve_internal_energy=0,
                  kernel=_abi.KERNEL_WENDLAND_C6, eos=_abi.EOS_POLYTROPIC,
                  precision=_abi.PREC_FP32, adaptive_dt=1, gamma=gamma,
                  poly_K=K, eta=eta, visc_alpha=1.0, visc_beta=2.0,
                  visc_eps=0.01, cfl=0.25, initial_dt=1e-4,
                  gravity=_abi.GRAV_FMM, G=1.0, theta=0.5, leaf_size=32,
                  n_target=100)
    return _pack(pos[:, 0], pos[:, 1], pos[:, 2], zero, zero, zero, u, h, m,
                 rho_floor, P, c), params


def uniform_sphere(n=1_000_000, seed=SEED_BASE + 5):
    """n seeded uniform random points in the unit ball, rho = 1,
    u = 0.05, gamma = 5/3, Wendland C6 with 100 neighbours (Poisson
    positions: a test input for the smoothing-length solvers)."""
    rng = np.random.default_rng(seed)
    dirs, r = _uniform_ball(n, rng)
    return _cold_sphere(dirs * r[:, None])


def _fcc_ball(n, rng):
    """Points of a face-centred cubic lattice inside the unit ball, with
    about n points: lattice constant a from 4 / a^3 = n / (4 pi / 3), the
    lattice origin shifted by a seeded offset in the unit cell."""
    a = (4.0 * (4.0 * math.pi / 3.0) / n) ** (1.0 / 3.0)
    m = int(math.ceil(1.0 / a)) + 1
    g = np.arange(-m, m + 1) * a
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    corner = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    basis = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]]) * a
    shift = (rng.random(3) - 0.5) * a
    pts = (corner[:, None, :] + basis[None, :, :]).reshape(-1, 3) + shift
    return pts[np.einsum("ij,ij->i", pts, pts) < 1.0]


def _cold_sphere(pos):
    n = len(pos)
    m = np.full(n, (4.0 * math.pi / 3.0) / n)
    rho = np.ones(n)
    eta = _eta_for(100)
    h = eta * 2.0 * (m / rho) ** (1.0 / 3.0)
    gamma = 5.0 / 3.0
    u = np.full(n, 0.05)
    P = (gamma - 1.0) * rho * u
    c = np.sqrt(gamma * (gamma - 1.0) * u)
    zero = np.zeros(n)
    params = dict(thermokinetic=0, kernel=_abi.KERNEL_WENDLAND_C6, eos=_abi.EOS_IDEAL,
                  precision=_abi.PREC_FP32, adaptive_dt=1, gamma=gamma,
                  eta=eta, visc_alpha=1.0, visc_beta=2.0, visc_eps=0.01,
                  cfl=0.25, initial_dt=1e-4, gravity=_abi.GRAV_FMM, G=1.0,
                  theta=0.5, leaf_size=32, n_target=100)
    return _pack(pos[:, 0], pos[:, 1], pos[:, 2], zero, zero, zero, u, h, m,
                 rho, P, c), params


def collapse(n=4_000_000, seed=SEED_BASE + 5, jitter=0.15):
    """Cold collapse of a uniform sphere: about n points of a seeded FCC
    lattice in the unit ball (the lattice of data/collapse_nx10.par; random
    points would leave close pairs that unsoftened gravity merges within a
    few steps), each moved by a seeded uniform offset of up to `jitter`
    nearest-neighbour spacings per axis (so hard neighbour counts are not
    locked to lattice shells, while no two points come closer than
    1 - 2 sqrt(3) jitter spacings), rho = 1, u = 0.05, gamma = 5/3, G = 1,
    theta = 0.5, Wendland C6 with 100 neighbours."""
    rng = np.random.default_rng(seed)
    pos = _fcc_ball(n, rng)
    if jitter:
        a = (4.0 * (4.0 * math.pi / 3.0) / n) ** (1.0 / 3.0)
        d_nn = a / math.sqrt(2.0)
        pos = pos + (rng.random(pos.shape) * 2.0 - 1.0) * (jitter * d_nn)
    return _cold_sphere(pos)


def random_box(n=2000, seed=1, h_spread=0.5, n_ngb=40, v_scale=0.1,
               kernel=_abi.KERNEL_CUBIC_SPLINE):
    """Test input: n seeded uniform points in the unit cube with smoothing
    lengths varied by a factor up to (1 + h_spread) around the value for
    n_ngb neighbours, random velocities and energies.  Variable h exercises
    the traversal's node pruning (tree_topology.h:323-355)."""
    rng = np.random.default_rng(seed)
    x, y, z = rng.random(n), rng.random(n), rng.random(n)
    m = np.full(n, 1.0 / n)
    eta = _eta_for(n_ngb)
    h0 = eta * 2.0 * (1.0 / n) ** (1.0 / 3.0)
    h = h0 * (1.0 + h_spread * rng.random(n))
    vx, vy, vz = (v_scale * rng.standard_normal(n) for _ in range(3))
    u = 1.0 + 0.5 * rng.random(n)
    gamma = 1.4
    c = np.sqrt(gamma * (gamma - 1.0) * u)
    params = dict(thermokinetic=0, kernel=kernel, eos=_abi.EOS_IDEAL, gamma=gamma, eta=eta,
                  precision=_abi.PREC_FP64, adaptive_dt=1, initial_dt=1e-3,
                  n_target=n_ngb)
    return _pack(x, y, z, vx, vy, vz, u, h, m, None, None, c), params
