"""Batched mode: every instance of a batch evolves exactly as it does in a
context of its own (same range, keys, tree, neighbour tiles in the same
order, periodic copies and time step), so the state of each instance after
several steps is bit-identical to a separate run with the same parameters.
"""
import numpy as np
import pytest

from flecsph_amd import (BatchSystem, BodySystem, FsphError, _abi, autotune,
                         generators)

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "z", "vx", "vy", "vz", "u", "h", "rho", "P", "c", "ax",
          "dudt", "nngb")


def _single(parts, prm, steps):
    s = BodySystem(_abi.default_params(**prm))
    s.upload(parts)
    for _ in range(steps):
        s.step()
    ids = s.get("id")
    o = np.argsort(ids)
    out = {f: s.get(f)[o] for f in FIELDS}
    out["id"] = ids[o]
    out["time"] = s.time()
    s.close()
    return out


def _batched(sets, prm, steps):
    b = BatchSystem(_abi.default_params(**prm), len(sets))
    b.upload(sets)
    for _ in range(steps):
        b.step()
    res = []
    for i in range(len(sets)):
        ids = b.instance_get(i, "id")
        o = np.argsort(ids)
        out = {f: b.instance_get(i, f)[o] for f in FIELDS}
        out["id"] = ids[o]
        out["time"] = b.instance_time(i)
        res.append(out)
    b.close()
    return res


def _compare(res, ref, i):
    assert np.array_equal(res["id"], ref["id"]), i
    for f in FIELDS:
        assert np.array_equal(res[f], ref[f]), (i, f)
    assert res["time"] == ref["time"], i


@pytest.mark.parametrize("precision", [_abi.PREC_FP64, _abi.PREC_FP32])
def test_batch_sedov_matches_separate_runs(precision):
    sets, prm = [], None
    for ns in (12, 15, 13, 16):
        parts, prm = generators.sedov(n_side=ns)
        sets.append(parts)
    prm = dict(prm, precision=precision)
    res = _batched(sets, prm, 4)
    for i, parts in enumerate(sets):
        _compare(res[i], _single(parts, prm, 4), i)


@pytest.mark.parametrize("h_mode", [_abi.H_CLOSED, _abi.H_NEWTON])
def test_batch_random_boxes_match_separate_runs(h_mode):
    sets, prm = [], None
    for n, seed in ((3000, 41), (5000, 42), (2000, 43)):
        parts, prm = generators.random_box(n, seed=seed, h_spread=0.6,
                                           n_ngb=40)
        sets.append(parts)
    prm = dict(prm, h_mode=h_mode, n_target=40)
    res = _batched(sets, prm, 3)
    for i, parts in enumerate(sets):
        _compare(res[i], _single(parts, prm, 3), i)


def test_batch_rejects_one_simulation_calls():
    parts, prm = generators.random_box(1000, seed=3)
    b = BatchSystem(_abi.default_params(**prm), 2)
    b.upload([parts, parts])
    b.step()
    with pytest.raises(Exception):
        b.time()
    with pytest.raises(Exception):
        b.reductions()
    b.close()
    with pytest.raises(Exception):
        BatchSystem(_abi.default_params(**dict(prm, gravity=_abi.GRAV_FMM)),
                    2)


def test_autotune_keeps_the_physics():
    """The autotuner times every candidate from the same saved state and
    leaves the simulation where it was: after the call and one more step the
    run is the untuned run's second step, up to the re-association of the
    pair sums."""
    parts, prm = generators.sedov(n_side=16)
    s = BodySystem(_abi.default_params(**prm))
    s.upload(parts)
    s.step()  # the initial iteration
    before = {f: s.get(f).copy() for f in ("x", "vx", "u", "h")}
    t_before = s.time()
    g, ch = autotune(s, steps=1)
    assert g in (64, 128) and ch in (64, 128, 256)
    assert s.time() == t_before
    for f, v in before.items():
        assert np.array_equal(s.get(f), v), f
    s.step()
    ref = _single(parts, prm, 2)
    ids = s.get("id")
    o = np.argsort(ids)
    assert s.time()[0] == ref["time"][0]
    assert np.isclose(s.time()[2], ref["time"][2], rtol=1e-6)
    for f in ("x", "h", "rho"):
        v = s.get(f)[o]
        assert np.allclose(v, ref[f], rtol=1e-5, atol=1e-7), f
    s.close()


@pytest.mark.parametrize("case", ["random_varh", "polytrope50k"])
def test_set_tuning_keeps_neighbour_sets(oracle_mod, case):
    """tile_rows and chunk cut the groups' rows into tiles and the
    candidates into LDS rounds; the groups stay the reference's (group_max
    128), so the neighbour lists and the signal speed are the oracle's at
    group_max 128 for every choice.  Variable-h inputs: with uniform h a
    different grouping would give the same lists and prove nothing."""
    from flecsph_amd import set_tuning
    if case == "random_varh":
        parts, prm = generators.random_box(6000, seed=101, h_spread=1.5)
    else:
        parts, prm = generators.polytrope(n=50_000)
        prm = dict(prm, gravity=_abi.GRAV_NONE)
    prm = dict(prm, precision=_abi.PREC_FP64)
    p = _abi.default_params(**prm)
    assert p.group_max == 128
    o = oracle_mod.OracleSystem(p, threads=8)
    o.upload(parts)
    o.update_iteration()
    o.density_eos()
    oo, io = o.neighbor_lists()
    io_sorted = np.concatenate([np.sort(io[oo[i]:oo[i + 1]])
                                for i in range(len(oo) - 1)])
    vs_o = o.get("vsig")
    for tr, ch in ((64, 64), (128, 128), (64, 256)):
        s = BodySystem(p)
        s.upload(parts)
        set_tuning(s, tr, ch)
        s.update_iteration()
        s.density_eos()
        off, idx = s.neighbor_lists()
        assert np.array_equal(off, oo), (tr, ch)
        assert np.array_equal(idx, io_sorted), (tr, ch)
        vs = s.get("vsig")
        assert np.abs(vs - vs_o).max() <= 1e-12 * np.abs(vs_o).max(), (tr, ch)
        with pytest.raises(FsphError):
            set_tuning(s, 96, 128)
        with pytest.raises(FsphError):
            set_tuning(s, 64, 100)
        s.close()


def test_tile_width_change_applies_from_the_next_update():
    """A tile width changed between update_iteration and the pair passes
    takes effect at the next update_iteration: the passes keep sizing their
    work for the tiles they were given (up to 128 rows here), and give the
    results of a context that never changed it."""
    from flecsph_amd import set_tuning
    parts, prm = generators.random_box(6000, seed=101, h_spread=1.5)
    out = []
    for late in (False, True):
        s = BodySystem(_abi.default_params(**prm))
        s.upload(parts)
        set_tuning(s, 128, 128)
        s.update_iteration()
        assert s.tree()["group_count"].max() > 64
        if late:
            set_tuning(s, 64, 128)
        s.density_eos()
        s.integrate(_abi.SAVE_VHALF)
        s.acceleration()
        out.append({f: s.get(f) for f in ("rho", "P", "ax", "ay", "az")})
        s.close()
    for f in out[0]:
        assert np.array_equal(out[0][f], out[1][f]), f


@pytest.mark.parametrize("tuning", [(128, 128), (64, 64)])
@pytest.mark.parametrize("precision", [_abi.PREC_FP32, _abi.PREC_FP64])
def test_graph_replay_matches_eager_every_step(precision, tuning):
    """The graph-replayed step tail gives the eager batched results, step
    after step (each replay follows eager tree building and host syncs), with
    two-wavefront (tile width 128) and one-wavefront (tile width 64) pair
    kernels."""
    from flecsph_amd import set_tuning
    sets, prm = [], None
    for ns in (10, 14, 12):
        parts, prm = generators.sedov(n_side=ns)
        sets.append(parts)
    prm = dict(prm, precision=precision)
    eager = BatchSystem(_abi.default_params(**prm), len(sets))
    graph = BatchSystem(_abi.default_params(**prm), len(sets))
    graph.enable_graph(True)
    for b in (eager, graph):
        set_tuning(b, *tuning)
        b.upload(sets)
    for step in range(6):
        eager.step()
        graph.step()
        for i in range(len(sets)):
            for f in ("x", "vx", "u", "h", "rho", "P", "dudt", "ax"):
                a, g = eager.instance_get(i, f), graph.instance_get(i, f)
                assert np.array_equal(a, g), (step, i, f)
            assert eager.instance_time(i) == graph.instance_time(i)
    eager.close()
    graph.close()


def test_loguniform_64_batch_against_oracle_and_invariants(oracle_mod):
    """The benchmark batch at its full size: 64 instances with seeded
    log-uniform particle counts in [8k, 200k] (about 3.2M particles), each a
    periodic Sedov blast or an open Sod tube by a seeded coin, graph replay
    of the step tail, initial iteration + 2 steps.  The smallest blast and
    the smallest tube are checked against the FP64 oracle run on that
    instance alone (FP32 bar of the single-simulation tests); every instance
    against invariants that hold at any size: ids a permutation of the
    instance's particles and copies (no copies for a tube), sorted keys,
    interior density 1 of a blast's unit-density lattice, and vanishing
    total momentum change (pairwise antisymmetric forces: sum m a = 0 over
    the instance)."""
    sets, prm, periodic = generators.mixed_batch()
    assert len(sets) == 64 and 0 < sum(periodic) < 64
    n_all = sum(len(p["x"]) for p in sets)
    assert 2_000_000 < n_all < 5_000_000
    print(f"64-instance batch, {n_all} particles, {sum(periodic)} blasts",
          flush=True)
    b = BatchSystem(_abi.default_params(**prm), 64)
    b.upload(sets)
    for i, per in enumerate(periodic):
        b.instance_set_periodic(i, per)
    b.step()
    b.enable_graph(True)
    for _ in range(2):
        b.step()
    for i, parts in enumerate(sets):
        n_tot, n_real = b.instance_count(i)
        assert n_real == len(parts["x"])
        if not periodic[i]:
            assert n_tot == n_real, i
        ids = b.instance_get(i, "id")
        assert np.array_equal(np.sort(ids), np.arange(n_tot)), i
        key = b.instance_get(i, "key")
        assert np.all(key[1:] > key[:-1]), i
        real = ids < n_real
        if periodic[i]:
            rho = b.instance_get(i, "rho")
            assert abs(np.median(rho[real]) - 1.0) < 0.02, i
        m = b.instance_get(i, "m")
        A = np.stack([b.instance_get(i, f) for f in ("ax", "ay", "az")], 1)
        mom = np.abs((m[:, None] * A).sum(0)).max()
        assert mom <= 1e-5 * (m[:, None] * np.abs(A)).sum(), i
    sizes = np.array([len(p["x"]) for p in sets])
    per = np.array(periodic)
    picks = [int(np.flatnonzero(per)[np.argmin(sizes[per])]),
             int(np.flatnonzero(~per)[np.argmin(sizes[~per])])]
    for i in picks:
        print(f"instance {i}: {sizes[i]} particles, periodic {periodic[i]}, "
              f"vs the oracle", flush=True)
        q = dict(prm)
        if not periodic[i]:
            q["periodic"] = (0, 0, 0)
        o = oracle_mod.OracleSystem(_abi.default_params(**q), threads=8)
        o.upload(sets[i])
        for _ in range(3):
            o.step()
        assert b.instance_count(i) == o.count()
        dtg, dto = b.instance_time(i)[1], o.time()[1]
        assert abs(dtg - dto) <= 1e-4 * dto
        # real particles by id: FP32 and FP64 positions differ by ~1e-10,
        # which can move a particle across a Morton cell boundary, so the
        # sorted order (and the numbering of the periodic copies that
        # follows it) need not agree
        n_real = len(sets[i]["x"])
        ig, io = b.instance_get(i, "id"), o.get("id")
        rg, ro = ig < n_real, io < n_real
        og, oo = np.argsort(ig[rg]), np.argsort(io[ro])
        assert np.array_equal(ig[rg][og], io[ro][oo])
        for f in ("x", "y", "z", "u", "rho"):
            g = b.instance_get(i, f)[rg][og]
            r = o.get(f)[ro][oo]
            tol = 1e-9 if f in "xyz" else 1e-4 * np.abs(r).max()
            assert np.abs(g - r).max() <= tol, (i, f)
        o.close()
    b.close()


def test_open_instance_in_a_periodic_batch_equals_its_own_run():
    """An instance switched to open boundaries in a periodic batch evolves
    exactly as a context of its own without periodic boundaries, and the
    periodic instance beside it as its own periodic context."""
    tube = generators.sod_tube(j=3)
    blast, prm = generators.sedov(n_side=12)
    b = BatchSystem(_abi.default_params(**prm), 2)
    b.upload([blast, tube])
    b.instance_set_periodic(1, False)
    for _ in range(3):
        b.step()
    for i, (parts, per) in enumerate(((blast, True), (tube, False))):
        q = dict(prm) if per else dict(prm, periodic=(0, 0, 0))
        ref = _single(parts, q, 3)
        ids = b.instance_get(i, "id")
        o = np.argsort(ids)
        assert np.array_equal(ids[o], ref["id"]), i
        for f in FIELDS:
            assert np.array_equal(b.instance_get(i, f)[o], ref[f]), (i, f)
        assert b.instance_time(i) == ref["time"], i
    b.close()
