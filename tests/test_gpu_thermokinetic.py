"""Parity of the thermokinetic formulation on the GPU (the reference's
default energy equation, thermokinetic_formulation = true) with the FP64
oracle: set_total_energy, u recovered in the density pass, the de/dt pass
with its half-step P and c (compute_dedt, default_physics.h:503-566;
recompute_pressure_soundspeed_thermokinetic, 170-180) and whole driver
iterations that kick e instead of u (hydro/main_driver.cc:176-229).

Bar: FP64 mode 1e-11 relative to the field's largest magnitude (summation
order only); FP32 mode 1e-5 for de/dt, 3e-6 for P and c (compensated FP32
pair sums against the FP64 oracle); u is FP64 arithmetic in both modes."""
import numpy as np
import pytest

from flecsph_amd import BodySystem, _abi, generators

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def case(name, precision):
    if name == "random_varh":
        parts, prm = generators.random_box(6000, seed=101, h_spread=1.5)
    elif name == "sedov16_periodic":
        parts, prm = generators.sedov(n_side=16)
    elif name == "sod_small":
        parts, prm = generators.sod(scale=1)
    else:
        parts, prm = generators.polytrope(n=20000)
    return parts, dict(prm, precision=precision, thermokinetic=1)


def both(oracle_mod, parts, prm):
    p = _abi.default_params(**prm)
    g, o = BodySystem(p), oracle_mod.OracleSystem(p, threads=8)
    g.upload(parts)
    o.upload(parts)
    return g, o


TOL = {_abi.PREC_FP64: dict(dedt=1e-11, P=1e-11, c=1e-11),
       _abi.PREC_FP32: dict(dedt=1e-5, P=3e-6, c=3e-6)}


@pytest.mark.parametrize("precision", [_abi.PREC_FP64, _abi.PREC_FP32])
@pytest.mark.parametrize("name", ["random_varh", "sedov16_periodic",
                                  "polytrope20k"])
def test_dedt_pass(oracle_mod, name, precision):
    parts, prm = case(name, precision)
    g, o = both(oracle_mod, parts, prm)
    print(f"de/dt pass {name} precision {precision}", flush=True)
    for s in (g, o):
        s.update_iteration()
        s.set_total_energy()
        s.density_eos()
        s.integrate(_abi.SAVE_VHALF)
        s.acceleration()
        if prm.get("gravity"):
            s.gravity()
        s.set_dt(1e-4)
        s.integrate(_abi.KICK_V)
    assert np.array_equal(g.get("e"), o.get("e"))
    for s in (g, o):
        s.dedt()
    tol = TOL[precision]
    for f in ("dedt", "P", "c"):
        e = rel_err(g.get(f), o.get(f))
        assert e <= tol[f], (f, e)
    # u = e - v^2/2 in FP64 on both sides; in FP32 mode v differs by the
    # FP32 error of the kick (dt a)
    tol_u = 1e-15 if precision == _abi.PREC_FP64 else 1e-9
    assert rel_err(g.get("u"), o.get("u")) <= tol_u
    if prm.get("gravity"):
        # the g.v term is there: without it de/dt would differ by it
        G = np.stack([o.get(k) for k in ("gx", "gy", "gz")], 1)
        V = np.stack([o.get(k) for k in ("vx", "vy", "vz")], 1)
        assert np.abs((G * V).sum(1)).max() > 10 * tol["dedt"] * np.abs(
            o.get("dedt")).max()


@pytest.mark.parametrize("name", ["random_varh", "sedov16_periodic",
                                  "sod_small"])
def test_three_steps_fp64(oracle_mod, name):
    parts, prm = case(name, _abi.PREC_FP64)
    g, o = both(oracle_mod, parts, prm)
    print(f"thermokinetic steps {name}", flush=True)
    for _ in range(3):
        g.step()
        o.step()
    itg, dtg, _ = g.time()
    ito, dto, _ = o.time()
    assert itg == ito and abs(dtg - dto) <= 1e-10 * dto
    assert np.array_equal(g.get("id"), o.get("id"))
    for f in ("x", "vx", "u", "e", "h", "dedt"):
        e = rel_err(g.get(f), o.get(f))
        assert e <= 1e-9, (f, e)


def test_three_steps_fp32_sedov(oracle_mod):
    parts, prm = case("sedov16_periodic", _abi.PREC_FP32)
    g, o = both(oracle_mod, parts, prm)
    for _ in range(3):
        g.step()
        o.step()
    assert np.array_equal(g.get("id"), o.get("id"))
    dtg, dto = g.time()[1], o.time()[1]
    assert abs(dtg - dto) <= 1e-4 * dto
    for f in ("x", "y", "z"):
        assert np.abs(g.get(f) - o.get(f)).max() <= 1e-9
    for f in ("u", "e", "rho"):
        assert rel_err(g.get(f), o.get(f)) <= 1e-4, f


def test_graph_replay_kicks_e(oracle_mod):
    """Batched graph replay in the thermokinetic formulation equals the
    eager batched steps bit for bit (the replayed tail runs the de/dt pass
    and kick_e)."""
    from flecsph_amd import BatchSystem
    sets, prm = [], None
    for ns in (10, 13):
        parts, prm = generators.sedov(n_side=ns)
        sets.append(parts)
    prm = dict(prm, thermokinetic=1)
    eager = BatchSystem(_abi.default_params(**prm), len(sets))
    graph = BatchSystem(_abi.default_params(**prm), len(sets))
    graph.enable_graph(True)
    for b in (eager, graph):
        b.upload(sets)
    for step in range(4):
        eager.step()
        graph.step()
        for i in range(len(sets)):
            for f in ("x", "vx", "u", "e", "dedt", "P"):
                a, b = eager.instance_get(i, f), graph.instance_get(i, f)
                assert np.array_equal(a, b), (step, i, f)
    eager.close()
    graph.close()


def test_dedt_needs_the_formulation():
    from flecsph_amd import FsphError
    parts, prm = generators.random_box(500, seed=3)
    g = BodySystem(_abi.default_params(**dict(prm, thermokinetic=0)))
    g.upload(parts)
    g.update_iteration()
    with pytest.raises(FsphError):
        g.dedt()
    with pytest.raises(FsphError):
        g.integrate(_abi.KICK_E)
