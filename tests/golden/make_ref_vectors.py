"""Writes tests/golden/ref_geometry.npz: input/output vectors of the
reference's own Morton curve (filling_curve.h:503-525, :528-549) and
geometric predicates (tree_geometry.h:265-350), evaluated by the
reference's headers compiled into oracle/_ref/libref.so (oracle/ref_shim.cc,
`make -C oracle ref`).  Run in a tree where /root/reference exists; the
fixture then travels without the reference.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(
    os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import ref  # noqa: E402


def main():
    assert ref.available(), "build oracle/_ref/libref.so first"
    rng = np.random.default_rng(20240601)
    out = {}
    # --- keys: random points in a random range
    lo = np.array([-0.7, 0.1, -2.0])
    hi = np.array([0.9, 0.35, 3.0])
    n = 4000
    pts = lo + (hi - lo) * rng.random((n, 3))
    # lattice points of a cell-centred 16^3 lattice in [-1/2,1/2]^3 with the
    # range widened by h (bodies_system.h:240-250), including points exactly
    # on the range bounds
    g = (np.arange(16) + 0.5) / 16 - 0.5
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    lat = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    h = 2.8794 / 16
    lo2 = lat.min(0) - h
    hi2 = lat.max(0) + h
    edge = np.array([lo2, hi2, (lo2 + hi2) / 2, lo2 + (hi2 - lo2) / 3])
    out["key_lo_a"], out["key_hi_a"], out["key_pts_a"] = lo, hi, pts
    out["key_out_a"] = ref.morton_keys(lo, hi, pts[:, 0], pts[:, 1],
                                       pts[:, 2])
    pb = np.concatenate([lat, edge])
    out["key_lo_b"], out["key_hi_b"], out["key_pts_b"] = lo2, hi2, pb
    out["key_out_b"] = ref.morton_keys(lo2, hi2, pb[:, 0], pb[:, 1],
                                       pb[:, 2])
    out["key_root_min_max"] = np.array(
        [ref.load().ref_key_root(), ref.load().ref_key_min(),
         ref.load().ref_key_max()], dtype=np.uint64)
    # --- round trip of a few keys (filling_curves.cc:202-241 pattern)
    rt = out["key_out_a"][:64]
    out["rt_keys"] = rt
    out["rt_coords"] = np.array([ref.key_coordinates(k, lo, hi) for k in rt])
    # --- predicates on random and tangent configurations
    m = 3000
    c = rng.normal(size=(m, 3))
    mn = rng.normal(size=(m, 3))
    mx = mn + rng.random((m, 3))
    r = rng.random(m) * 1.5
    # tangent cases: sphere touching a box face exactly
    t = 200
    mn[:t] = 0.0
    mx[:t] = 1.0
    c[:t] = np.stack([1.0 + r[:t], rng.random(t), rng.random(t)], 1)
    out["sb_min"], out["sb_max"], out["sb_c"], out["sb_r"] = mn, mx, c, r
    out["sb_out"] = np.array([ref.sphere_box(mn[i], mx[i], c[i], r[i])
                              for i in range(m)])
    mn2 = rng.normal(size=(m, 3))
    mx2 = mn2 + rng.random((m, 3))
    mn2[:t] = mx[:t]  # boxes touching at a face
    mx2[:t] = mx[:t] + 1.0
    out["bb_min2"], out["bb_max2"] = mn2, mx2
    out["bb_out"] = np.array([ref.box_box(mn[i], mx[i], mn2[i], mx2[i])
                              for i in range(m)])
    p1 = rng.normal(size=(m, 3))
    p2 = p1 + rng.normal(size=(m, 3)) * 0.3
    rr = np.sqrt(((p1 - p2) ** 2).sum(1))
    rr[:t] = np.sqrt((p1[:t] - p2[:t]) ** 2 @ np.ones(3))  # boundary cases
    rr[t:] *= rng.random(m - t) * 2
    out["wd_p1"], out["wd_p2"], out["wd_r"] = p1, p2, rr
    out["wd_out"] = np.array([ref.within_distance2(p1[i], p2[i], rr[i])
                              for i in range(m)])
    np.savez_compressed(os.path.join(os.path.dirname(__file__),
                                     "ref_geometry.npz"), **out)
    print("wrote ref_geometry.npz")


if __name__ == "__main__":
    main()
