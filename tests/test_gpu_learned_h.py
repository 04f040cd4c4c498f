"""Learned smoothing length (MLP proposal + narrowed Newton solve) against
the FP64 oracle's plain Newton-h from the same state.

After the predictor's training steps, every learned update is checked by
re-solving it on the CPU: the oracle takes the positions the GPU solved at
and the smoothing lengths from before the update, and runs plain Newton-h
(gather radius h_search h, start at h).  Both stop within newton_tol h of
the same root, so h agrees to 3e-4 relative (FP32 sums) and the hard
neighbour counts agree except where a neighbour sits at distance h."""
import numpy as np
import pytest

from flecsph_amd import BodySystem, _abi, generators

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "z", "vx", "vy", "vz", "u", "m", "h")


def _by_id(s, names):
    ids = s.get("id")
    o = np.argsort(ids)
    return {k: s.get(k)[o] for k in names}, ids[o]


def _oracle_newton(oracle_mod, prm, state, h_old, ids):
    p = _abi.default_params(**dict(prm, h_mode=_abi.H_NEWTON,
                                   precision=_abi.PREC_FP64))
    o = oracle_mod.OracleSystem(p, threads=8)
    parts = {k: state[k] for k in FIELDS}
    parts["h"] = h_old
    parts["id"] = ids
    o.upload(parts)
    o.update_iteration()
    o.update_h()
    out, oid = _by_id(o, ("h", "hngb", "hiter"))
    o.close()
    real = oid < len(ids)          # drop the oracle's own periodic copies
    assert np.array_equal(oid[real], ids)
    return {k: v[real] for k, v in out.items()}


@pytest.mark.parametrize("case", ["collapse", "sedov", "sedov_fp64"])
def test_learned_h_is_the_newton_root(oracle_mod, case):
    if case == "collapse":
        # the collapse with its self-gravity (FMM, theta 0.5, quadrupole)
        parts, prm = generators.collapse(n=40000)
        assert prm["gravity"] == _abi.GRAV_FMM
    else:
        parts, prm = generators.sedov(n_side=24)
        if case == "sedov_fp64":  # the FP64 solver kernels
            prm = dict(prm, precision=_abi.PREC_FP64)
    prm.update(h_mode=_abi.H_LEARNED, learn_steps=5)
    g = BodySystem(_abi.default_params(**prm))
    g.upload(parts)
    n_real = g.count()[1]
    checked = 0
    it_learned, it_plain = [], []
    for k in range(10):
        hs, ids0 = _by_id(g, ("h",))
        h_old = hs["h"]
        g.step()
        st = g.h_stats()
        if st["trained_steps"] < 5 or k < 6:
            continue
        state, ids = _by_id(g, FIELDS + ("hngb", "hiter"))
        real = ids < n_real
        assert np.array_equal(ids0[ids0 < n_real], ids[real])
        ref = _oracle_newton(oracle_mod, prm,
                             {f: state[f][real] for f in FIELDS},
                             h_old[ids0 < n_real], ids[real])
        rel = np.abs(state["h"][real] - ref["h"]) / ref["h"]
        print("%s step %d: fallback %d clamped %d loss %.3g, max rel %.2e, "
              "median %.2e, iterations learned %.2f plain %.2f" % (
                  case, k, st["fallback"], st["clamped"], st["last_loss"],
                  rel.max(), np.median(rel), state["hiter"][real].mean(),
                  ref["hiter"].mean()))
        assert rel.max() < 3e-4 and np.median(rel) < 2e-5
        dn = state["hngb"][real] - ref["hngb"]
        assert np.mean(dn == 0) > 0.995 and np.abs(dn).max() <= 2
        assert st["fallback"] <= 0.1 * n_real
        it_learned.append(state["hiter"][real].mean())
        it_plain.append(ref["hiter"].mean())
        checked += 1
    assert checked >= 3
    print("%s: mean Newton iterations learned %.3f plain %.3f" % (
        case, np.mean(it_learned), np.mean(it_plain)))
    # the learned start never needs more Newton steps than the plain one
    assert np.mean(it_learned) <= np.mean(it_plain) + 0.05


def test_learned_h_gathers_fewer_candidates():
    """What the learned form gains over plain Newton-h: the same steps of
    the 40k collapse with self-gravity, run once with Newton-h and once with
    learned-h (5 training updates).  After training, each learned update
    solves first on the step's SPH tiles, inside a bracket fitted to the
    predictor's error instead of h_search h, and builds gather tiles only
    for the tiles it could not settle there (those narrowed the same way),
    so it builds far fewer candidate words (the pair tests of the gather
    tiles) for the same roots (the test above checks the roots against the
    oracle)."""
    parts, prm = generators.collapse(n=40000)
    out = {}
    for mode in (_abi.H_NEWTON, _abi.H_LEARNED):
        g = BodySystem(_abi.default_params(**dict(prm, h_mode=mode,
                                                   learn_steps=5)))
        g.upload(parts)
        rec = []
        for k in range(10):
            g.step()
            rec.append(g.h_gather())
        out[mode] = rec
        g.close()
    hs = _abi.default_params(**prm).h_search
    for k in range(7, 10):
        (kn, wn), (kl, wl) = out[_abi.H_NEWTON][k], out[_abi.H_LEARNED][k]
        print("step %d: gather radius %.3f -> %.3f h, mask words %d -> %d "
              "(%.2fx)" % (k, kn, kl, wn, wl, wn / max(wl, 1)), flush=True)
        assert kn == pytest.approx(hs)
        assert kl < kn
        assert wl < 0.7 * wn


def test_learned_h_statistics_before_and_after_training():
    parts, prm = generators.random_box(6000, seed=31, h_spread=0.3,
                                       n_ngb=60)
    prm.update(h_mode=_abi.H_LEARNED, learn_steps=3, precision=_abi.PREC_FP32)
    g = BodySystem(_abi.default_params(**prm))
    g.upload(parts)
    seen = []
    for _ in range(5):
        g.step()
        seen.append(g.h_stats())
    # the first update (relaxing the input h) is solved, not learned from
    assert [s["trained_steps"] for s in seen] == [0, 1, 2, 3, 3]
    assert all(s["fallback"] == 0 for s in seen[:4])
    assert all(np.isfinite(s["last_loss"]) for s in seen)
    h = g.get("h")
    assert np.all(np.isfinite(h)) and np.all(h > 0)
