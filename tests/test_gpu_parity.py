"""Parity of the HIP path (libfsph.so through the C ABI) with the FP64 CPU
oracle on identical seeded inputs.

Bar: keys, sort order, tree, groups and neighbour lists bit-exact; FP64-mode
physics within 1e-11 relative (only summation order differs); FP32-mode
physics (compensated FP32 sums) within the tolerances stated per test,
measured against the FP64 oracle."""
import numpy as np
import pytest

from flecsph_amd import BodySystem, _abi, generators

pytestmark = pytest.mark.gpu


def both(oracle_mod, parts, prm):
    p = _abi.default_params(**prm)
    g = BodySystem(p)
    o = oracle_mod.OracleSystem(p, threads=8)
    g.upload(parts)
    o.upload(parts)
    return g, o


def cases():
    return {
        "random_varh": generators.random_box(6000, seed=101, h_spread=1.5),
        "random_wendland": generators.random_box(
            4000, seed=102, h_spread=0.8, n_ngb=100,
            kernel=_abi.KERNEL_WENDLAND_C6),
        "sedov16_periodic": generators.sedov(n_side=16),
        "sod_small": generators.sod(scale=1),
        # about 1500 neighbours each: a group's traversal accepts most of
        # the tree, which overflows the small LDS stack (the traversal is
        # run again with the large one) and the default candidate capacity
        "random_dense": generators.random_box(
            6000, seed=103, h_spread=0.5, n_ngb=1000),
    }


CASES = list(cases().keys())


def get_case(name, precision):
    parts, prm = cases()[name]
    prm = dict(prm)
    prm["precision"] = precision
    return parts, prm


def tree_sets(t):
    nodes = {(int(k), int(s), int(c)) for k, s, c in
             zip(t["key"], t["start"], t["count"])}
    par = {(int(t["key"][i]), int(t["key"][p]) if p >= 0 else -1)
           for i, p in enumerate(t["parent"])}
    groups = set(zip(t["group_start"].tolist(), t["group_count"].tolist()))
    return nodes, par, groups


@pytest.mark.parametrize("name", CASES)
def test_keys_sort_tree_groups_bit_exact(oracle_mod, name):
    parts, prm = get_case(name, _abi.PREC_FP32)
    g, o = both(oracle_mod, parts, prm)
    print(f"case {name} n={len(parts['x'])}", flush=True)
    g.update_iteration()
    o.update_iteration()
    assert g.count() == o.count()
    lo_g, hi_g = g.range()
    lo_o, hi_o = o.range()
    assert np.array_equal(lo_g, lo_o) and np.array_equal(hi_g, hi_o)
    assert np.array_equal(g.get("key"), o.get("key"))
    assert np.array_equal(g.get("id"), o.get("id"))
    for f in ("x", "y", "z", "h", "m", "u", "vx"):
        assert np.array_equal(g.get(f), o.get(f)), f
    assert tree_sets(g.tree()) == tree_sets(o.tree())


def test_tree_and_neighbors_bit_exact_past_one_workgroup_scan(oracle_mod):
    """Sedov 32^3 with its periodic copies (about 60,000 particles): the
    scans of the copy counts, node counts and group flags run over many
    tiles (more than one workgroup's 16,384 items), so the multi-tile scan
    path is what this case checks against the oracle."""
    parts, prm = generators.sedov(n_side=32)
    prm = dict(prm, precision=_abi.PREC_FP32)
    g, o = both(oracle_mod, parts, prm)
    for step in range(2):
        print(f"sedov32 update {step}", flush=True)
        g.update_iteration()
        o.update_iteration()
        assert g.count() == o.count()
        assert g.count()[0] > 3 * 16384
        assert np.array_equal(g.get("key"), o.get("key"))
        assert np.array_equal(g.get("id"), o.get("id"))
        assert tree_sets(g.tree()) == tree_sets(o.tree())
    og, ig = g.neighbor_lists()
    oo, io = o.neighbor_lists()
    assert np.array_equal(og, oo)
    io_sorted = np.concatenate([np.sort(io[oo[i]:oo[i + 1]])
                                for i in range(len(oo) - 1)])
    assert np.array_equal(ig, io_sorted)


@pytest.mark.parametrize("tile_rows", [64, 128])
@pytest.mark.parametrize("name", CASES)
def test_neighbor_lists_bit_exact(oracle_mod, name, tile_rows):
    """Lists read back through the tiles the pair passes walk, at both tile
    widths (64: groups of more than 64 particles are cut into two tiles),
    equal the oracle's at the reference's grouping (group_max 128)."""
    from flecsph_amd import set_tuning
    parts, prm = get_case(name, _abi.PREC_FP32)
    g, o = both(oracle_mod, parts, prm)
    print(f"case {name} tile_rows {tile_rows}", flush=True)
    set_tuning(g, tile_rows, 128)
    g.update_iteration()
    o.update_iteration()
    og, ig = g.neighbor_lists()
    oo, io = o.neighbor_lists()
    assert np.array_equal(og, oo)
    n = len(og) - 1
    io_sorted = np.concatenate([np.sort(io[oo[i]:oo[i + 1]])
                                for i in range(n)])
    assert np.array_equal(ig, io_sorted)
    assert np.array_equal(g.get("nngb"), np.diff(oo).astype(float))


def test_neighbor_lists_with_one_box_wide_sphere(oracle_mod):
    """20,000 random particles, one of them with an h that spans the box:
    its group's traversal accepts every node, so its pending nodes outgrow
    the first LDS stack (512 entries; the same walk replayed on the CPU
    peaks at 532) and the traversal is run again with the next one; its
    candidate tile holds every particle, so the tile capacity (2048 per
    slot at first) regrows four times.  The lists still equal the
    oracle's."""
    from flecsph_amd import set_tuning
    parts, prm = generators.random_box(n=20000, seed=7, h_spread=1.0)
    parts["h"][123] = 2.0
    prm = dict(prm, precision=_abi.PREC_FP32)
    g, o = both(oracle_mod, parts, prm)
    set_tuning(g, 64, 128)
    g.update_iteration()
    o.update_iteration()
    og, ig = g.neighbor_lists()
    oo, io = o.neighbor_lists()
    assert np.array_equal(og, oo)
    assert np.diff(oo).max() == 20000  # the wide particle meets everyone
    n = len(og) - 1
    io_sorted = np.concatenate([np.sort(io[oo[i]:oo[i + 1]])
                                for i in range(n)])
    assert np.array_equal(ig, io_sorted)


def rel_err(a, b, scale=None):
    scale = np.abs(b).max() if scale is None else scale
    return np.abs(a - b).max() / max(scale, 1e-300)


TOL = {
    # FP64 mode: same arithmetic, different summation order
    _abi.PREC_FP64: dict(rho=1e-12, P=1e-12, c=1e-12, vsig=1e-12, a=1e-11,
                         dudt=1e-11),
    # FP32 mode: compensated FP32 sums vs the FP64 oracle; errors are
    # normalised by the largest magnitude of the field
    _abi.PREC_FP32: dict(rho=2e-6, P=3e-6, c=2e-6, vsig=3e-6, a=1e-5,
                         dudt=1e-5),
}


@pytest.mark.parametrize("precision", [_abi.PREC_FP64, _abi.PREC_FP32])
@pytest.mark.parametrize("name", CASES)
def test_sph_passes(oracle_mod, name, precision):
    parts, prm = get_case(name, precision)
    g, o = both(oracle_mod, parts, prm)
    print(f"case {name} precision {precision}", flush=True)
    tol = TOL[precision]
    for s in (g, o):
        s.update_iteration()
        s.density_eos()
    for f in ("rho", "P", "c", "vsig"):
        e = rel_err(g.get(f), o.get(f))
        assert e <= tol[f], (f, e)
    for s in (g, o):
        s.integrate(_abi.SAVE_VHALF)
        s.acceleration()
    A = [np.stack([s.get("ax"), s.get("ay"), s.get("az")], 1) for s in (g, o)]
    e = rel_err(A[0], A[1])
    assert e <= tol["a"], ("a", e)
    for s in (g, o):
        s.set_dt(1e-4)
        s.integrate(_abi.KICK_V)
        s.dudt()
    e = rel_err(g.get("dudt"), o.get("dudt"))
    assert e <= tol["dudt"], ("dudt", e)
    for f in ("P", "c"):  # recompute_pressure_soundspeed
        e = rel_err(g.get(f), o.get(f))
        assert e <= tol[f], (f, e)


@pytest.mark.parametrize("tuning", [(64, 64), (128, 256), (64, 128)])
@pytest.mark.parametrize("precision", [_abi.PREC_FP64, _abi.PREC_FP32])
@pytest.mark.parametrize("name", ["random_varh", "sedov16_periodic"])
def test_sph_passes_tile_widths(oracle_mod, name, precision, tuning):
    """Tile width 64 selects the one-wavefront pair kernels over tiles of up
    to 64 rows (groups of up to 128 particles cut in two), 128 the
    two-wavefront kernels over whole groups; chunk 64/128/256 the LDS
    rounds.  Same bars for every choice, against the oracle at the
    reference's grouping (group_max 128)."""
    from flecsph_amd import set_tuning
    parts, prm = get_case(name, precision)
    g, o = both(oracle_mod, parts, prm)
    set_tuning(g, *tuning)
    tol = TOL[precision]
    for s in (g, o):
        s.update_iteration()
        s.density_eos()
    if name == "random_varh":
        assert g.tree()["group_count"].max() > 64  # some groups are cut
    for f in ("rho", "P", "c", "vsig"):
        assert rel_err(g.get(f), o.get(f)) <= tol[f], f
    for s in (g, o):
        s.integrate(_abi.SAVE_VHALF)
        s.acceleration()
    A = [np.stack([s.get("ax"), s.get("ay"), s.get("az")], 1) for s in (g, o)]
    assert rel_err(A[0], A[1]) <= tol["a"]
    for s in (g, o):
        s.set_dt(1e-4)
        s.integrate(_abi.KICK_V)
        s.dudt()
    assert rel_err(g.get("dudt"), o.get("dudt")) <= tol["dudt"]


@pytest.mark.parametrize("tile_rows", [64, 128])
def test_polytrope_lists_and_vsig_at_reference_grouping(oracle_mod,
                                                        tile_rows):
    """A 50k-particle polytrope (strongly variable h, groups of 1 to 128
    particles): lists bit-equal to the oracle's at group_max 128 for both
    tile widths, and the signal speed, an unweighted maximum over the list
    (so sensitive to any list difference), to FP64 rounding."""
    from flecsph_amd import set_tuning
    parts, prm = generators.polytrope(n=50_000)
    prm = dict(prm, precision=_abi.PREC_FP64, gravity=_abi.GRAV_NONE)
    g, o = both(oracle_mod, parts, prm)
    set_tuning(g, tile_rows, 64)
    for s in (g, o):
        s.update_iteration()
        s.density_eos()
    og, ig = g.neighbor_lists()
    oo, io = o.neighbor_lists()
    assert np.array_equal(og, oo)
    io_sorted = np.concatenate([np.sort(io[oo[i]:oo[i + 1]])
                                for i in range(len(oo) - 1)])
    assert np.array_equal(ig, io_sorted)
    for f in ("rho", "vsig"):
        assert rel_err(g.get(f), o.get(f)) <= 1e-12, f


@pytest.mark.parametrize("name", ["random_varh", "sedov16_periodic",
                                  "sod_small"])
def test_three_steps_fp64(oracle_mod, name):
    """Whole driver iterations (initial + 2 KDK steps) in FP64 mode."""
    parts, prm = get_case(name, _abi.PREC_FP64)
    g, o = both(oracle_mod, parts, prm)
    print(f"case {name}", flush=True)
    for _ in range(3):
        g.step()
        o.step()
    itg, dtg, ttg = g.time()
    ito, dto, tto = o.time()
    assert itg == ito and abs(dtg - dto) <= 1e-10 * dto
    assert g.count() == o.count()
    assert np.array_equal(g.get("id"), o.get("id"))
    for f in ("x", "y", "z", "vx", "vy", "vz", "u", "h"):
        e = rel_err(g.get(f), o.get(f),
                    scale=max(np.abs(o.get(f)).max(), 1e-12))
        assert e <= 1e-9, (f, e)
    rg, ro = g.reductions(), o.reductions()
    assert np.allclose(rg, ro, rtol=1e-9, atol=1e-12)


def test_three_steps_fp32_sedov(oracle_mod):
    parts, prm = get_case("sedov16_periodic", _abi.PREC_FP32)
    g, o = both(oracle_mod, parts, prm)
    for _ in range(3):
        g.step()
        o.step()
    assert g.count() == o.count()
    assert np.array_equal(g.get("id"), o.get("id"))
    dtg, dto = g.time()[1], o.time()[1]
    assert abs(dtg - dto) <= 1e-4 * dto
    for f in ("x", "y", "z"):
        assert np.abs(g.get(f) - o.get(f)).max() <= 1e-9
    for f in ("u", "rho"):
        assert rel_err(g.get(f), o.get(f)) <= 1e-4, f


def test_duplicate_key_is_reported(oracle_mod):
    parts, prm = generators.random_box(500, seed=3)
    parts["x"][7], parts["y"][7], parts["z"][7] = (
        parts["x"][3], parts["y"][3], parts["z"][3])
    g = BodySystem(_abi.default_params(**prm))
    g.upload(parts)
    with pytest.raises(Exception) as e:
        g.update_iteration()
    assert "DUPKEY" in str(e.value)


def test_sedov64_full_size_properties():
    """The benchmark problem at its full size (64^3 + periodic copies):
    properties that fix the result without the oracle."""
    parts, prm = generators.sedov(n_side=64)
    g = BodySystem(_abi.default_params(**prm))
    g.upload(parts)
    g.update_iteration()
    key = g.get("key")
    assert np.all(key[1:] > key[:-1])
    ids = g.get("id")
    n_tot, n_real = g.count()
    assert np.array_equal(np.sort(ids), np.arange(n_tot))
    g.density_eos()
    g.integrate(_abi.SAVE_VHALF)
    g.acceleration()
    m = g.get("m")
    A = np.stack([g.get("ax"), g.get("ay"), g.get("az")], 1)
    nngb = g.get("nngb")
    real = ids < n_real
    # interior density of the unit-density lattice is 1 to kernel accuracy
    assert abs(np.median(g.get("rho")[real]) - 1.0) < 0.02
    assert 85 <= np.mean(nngb[real]) <= 140
    # pairwise-antisymmetric forces: sum m a over all particles vanishes
    mom = np.abs((m[:, None] * A).sum(0)).max()
    assert mom <= 1e-5 * (m[:, None] * np.abs(A)).sum()
