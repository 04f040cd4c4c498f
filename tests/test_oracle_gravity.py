"""Pins the oracle's gravity (CPU): the literal restatement of the
reference's monopole traversal (tree_topology.h:449-795, fmm.h) and the
bucketed quadrupole engine the HIP path follows, against direct N^2
summation and against each other."""
import numpy as np
import pytest

from flecsph_amd import _abi, generators


def system(oracle_mod, n, theta, leaf, order, seed=5, G=1.0):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    P = v * (rng.random(n) ** (1 / 3))[:, None]
    z = np.zeros(n)
    parts = dict(x=P[:, 0], y=P[:, 1], z=P[:, 2], vx=z, vy=z, vz=z,
                 u=np.ones(n), h=np.full(n, 0.05), m=np.full(n, 1.0 / n),
                 id=np.arange(n, dtype=np.int64))
    prm = dict(gravity=_abi.GRAV_FMM, theta=theta, leaf_size=leaf,
               fmm_order=order, G=G)
    s = oracle_mod.OracleSystem(_abi.default_params(**prm), threads=4)
    s.upload(parts)
    s.update_iteration()
    s.acceleration()  # zeroes g and phi as compute_acceleration does
    return s


def fields(s):
    return (np.stack([s.get("gx"), s.get("gy"), s.get("gz")], 1),
            s.get("phi"))


def direct(oracle_mod, s, G=1.0):
    gx, gy, gz, phi = oracle_mod.direct_gravity(s.get("x"), s.get("y"),
                                                s.get("z"), s.get("m"), G)
    return np.stack([gx, gy, gz], 1), phi


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("leaf,order", [(1, 1), (1, 2), (32, 2), (8, 1)])
def test_tiny_opening_angle_is_direct_summation(oracle_mod, leaf, order):
    """theta -> 0 accepts no cell pair, so every particle pair is summed
    directly exactly once: the traversal covers all pairs."""
    s = system(oracle_mod, 700, 1e-9, leaf, order)
    s.gravity()
    g, phi = fields(s)
    gd, pd = direct(oracle_mod, s)
    assert rel(g, gd) < 1e-13 and rel(phi, pd) < 1e-13


def test_engine_reproduces_reference_traversal(oracle_mod):
    """At leaf size 1 and order 1 the bucketed engine computes the
    reference's scheme (constant field per affected node == first-order
    local shifts)."""
    s = system(oracle_mod, 3000, 0.5, 1, 1)
    s.gravity()  # literal reference traversal
    g1, p1 = fields(s)
    s.acceleration()
    s.gravity_engine()
    g2, p2 = fields(s)
    assert rel(g2, g1) < 1e-12 and rel(p2, p1) < 1e-12


def test_quadrupole_converges_faster_than_monopole(oracle_mod):
    errs = {}
    for order in (1, 2):
        for theta in (0.6, 0.3):
            s = system(oracle_mod, 4000, theta, 16, order)
            s.gravity()
            g, _ = fields(s)
            gd, _ = direct(oracle_mod, s)
            err = np.sqrt(((g - gd) ** 2).sum(1) / (gd ** 2).sum(1))
            errs[(order, theta)] = np.median(err)
    # first-order local field: error ~ theta^2; quadrupole + second-order
    # local: ~ theta^3
    assert errs[(2, 0.6)] < errs[(1, 0.6)]
    assert errs[(2, 0.3)] < errs[(2, 0.6)] / 4
    assert errs[(1, 0.3)] < errs[(1, 0.6)] / 2
    assert errs[(2, 0.6)] < 1e-2


def test_uniform_sphere_interior_field(oracle_mod):
    """Inside a uniform sphere g = -G M r / R^3 (to sampling noise)."""
    s = system(oracle_mod, 20000, 0.5, 32, 2)
    s.gravity()
    g, _ = fields(s)
    P = np.stack([s.get("x"), s.get("y"), s.get("z")], 1)
    r = np.linalg.norm(P, axis=1)
    sel = (r > 0.2) & (r < 0.7)
    radial = (g[sel] * P[sel]).sum(1) / r[sel]
    assert np.median(radial / (-r[sel])) == pytest.approx(1.0, abs=0.03)
