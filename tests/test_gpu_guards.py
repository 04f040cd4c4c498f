"""State checks (check_nans, check_negativity, default_physics.h:918-979):
the driver's step returns FSPH_E_NAN / FSPH_E_NEG as the reference's
asserts would stop the run, and the oracle agrees on the same input."""
import numpy as np
import pytest

from flecsph_amd import BodySystem, FsphError, _abi, generators

pytestmark = pytest.mark.gpu


def _case():
    parts, prm = generators.random_box(3000, seed=41, h_spread=0.5)
    return dict(parts), dict(prm, precision=_abi.PREC_FP32)


def test_nan_in_u_stops_the_step(oracle_mod):
    parts, prm = _case()
    parts["u"] = parts["u"].copy()
    parts["u"][1234] = np.nan
    p = _abi.default_params(**prm)
    print("nan in u of particle 1234, initial iteration", flush=True)
    g = BodySystem(p)
    g.upload(parts)
    with pytest.raises(FsphError) as e:
        g.step()  # the initial iteration (no drift: positions stay finite)
    assert e.value.code == _abi.FSPH_E_NAN
    g.close()
    o = oracle_mod.OracleSystem(p, threads=4)
    o.upload(parts)
    o.step()
    assert oracle_mod.load().fo_check_state(o._h) == _abi.FSPH_E_NAN


def test_negative_internal_energy_is_reported(oracle_mod):
    parts, prm = _case()
    parts["u"] = parts["u"].copy()
    parts["u"][17] = -1.0
    p = _abi.default_params(**prm)
    g = BodySystem(p)
    g.upload(parts)
    with pytest.raises(FsphError) as e:
        g.check_state()
    assert e.value.code == _abi.FSPH_E_NEG
    assert "id 17" in str(e.value)
    g.close()
    o = oracle_mod.OracleSystem(p, threads=4)
    o.upload(parts)
    assert oracle_mod.load().fo_check_state(o._h) == _abi.FSPH_E_NEG


def test_clean_state_passes_and_checks_can_be_turned_off():
    parts, prm = _case()
    g = BodySystem(_abi.default_params(**prm))
    g.upload(parts)
    for _ in range(3):
        g.step()
    g.check_state()
    parts["u"] = parts["u"].copy()
    parts["u"][5] = -1.0
    g.upload(parts)
    g.set_checks(False)
    g.step()  # not checked
    with pytest.raises(FsphError):
        g.check_state()
    g.close()


def test_a_drift_needs_a_new_tree_before_the_neighbour_passes():
    """After fsph_integrate drifts the particles, the tree and tiles no
    longer describe them: the passes that use them (density, the Newton-h
    and learned-h updates) fail with FSPH_E_ARG until
    fsph_update_iteration, the order of hydro/main_driver.cc:187-193; a
    kick alone leaves them valid."""
    parts, prm = _case()
    g = BodySystem(_abi.default_params(**dict(prm, h_mode=_abi.H_NEWTON)))
    g.upload(parts)
    g.step()
    g.integrate(_abi.KICK_V)
    g.density_eos()
    g.integrate(_abi.DRIFT)
    for call in (g.density_eos, g.update_h):
        with pytest.raises(FsphError) as e:
            call()
        assert e.value.code == _abi.FSPH_E_ARG
    g.update_iteration()
    g.density_eos()
    g.update_h()
    g.close()
