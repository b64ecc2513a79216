"""GPU inflow emitter (sf_set_emitter) against the oracle's restatement of the
reference's CudaEmitter + OnUpdateSolver schedule + AddSphParticles.

Right after an emission the new particles sit, unsorted, after the others on
both sides (the GPU's slot order of the others equals the oracle's: the
neighbour structures are bit-exact), so their positions, velocities, colours
and masses compare bit for bit; the particle count after every call must
match exactly.  With params.dt binding the CFL sub-step, both sides take the
same sub-step count and therefore the same emission decisions."""
import numpy as np
import pytest

from sph_seepage_flows_amd import SFError, SFSystem, _abi, scenes

from _compare import norm_err

pytestmark = pytest.mark.gpu

RI = 1.0 / 30.0
EMIT = dict(pos=(0.3, 0.3, 0.4), vel=(0.0, 0.0, -10.0), col=(0.5, 0.8, 1.0), kind=_abi.SF_EMITTER_CIRCLE, radius=0.04)


def _pair(oracle_mod, cap_extra=4000, solver=None):
    sc = scenes.mixed(collider=False)
    p, b, o = sc.params()
    if solver is not None:
        p.solver_type = solver
    cap = sc.n + cap_extra
    g = SFSystem.from_scene(sc, max_num=cap)
    g.set_params(p)
    orc = oracle_mod.OracleSystem(sc, 64, params=(p, b, o))
    g.set_emitter(**EMIT)
    orc.set_emitter(cap, EMIT["pos"], EMIT["vel"], EMIT["col"], EMIT["kind"], EMIT["radius"])
    return sc, g, orc


def _check_appended(g, o, n0, tag):
    n = g.sizes()[0]
    assert n == o.size(), (tag, n, o.size())
    for f in ("pos", "vel", "col", "mass", "radius", "label"):
        np.testing.assert_array_equal(g.get(f)[n0:n], o.get(f)[n0:n].astype(g.get(f).dtype), err_msg=f"{tag} {f}")


def test_emitter_matches_oracle(oracle_mod):
    sc, g, o = _pair(oracle_mod)
    en, k, _ = g.emitter_state()
    assert en and k == o.emitter_state()[1] and k > 0
    print(f"[emitter] n={sc.n} samples={k}", flush=True)
    n0 = sc.n
    g.UpdateSystem(RI)
    o.update_system(RI)
    _check_appended(g, o, n0, "call 1")  # the first call emits
    assert g.emitter_state()[2] == k
    # the whole state one sub-step later, the emitted particles sorted in
    g.UpdateSystem(RI)
    o.update_system(RI)
    np.testing.assert_array_equal(g.get("sort_perm"), o.get("sort_perm"))
    for f, tol in (("pos", 1e-6), ("vel", 1e-5), ("density", 1e-5)):
        e = norm_err(g.get(f), o.get(f))
        print(f"[emitter] call 2 {f} {e:.2e}", flush=True)
        assert e < tol, (f, e)
    emissions = 1
    for call in range(3, 46):
        n0 = g.sizes()[0]
        g.UpdateSystem(RI)
        o.update_system(RI)
        assert g.GetNumOfSubTimeSteps() == o.num_sub_steps() == 333
        if g.sizes()[0] != n0 or o.size() != n0:
            _check_appended(g, o, n0, f"call {call}")
            emissions += 1
    assert emissions >= 2 and g.emitter_state()[2] == o.emitter_state()[2] == emissions * k


def test_emitter_sph_solver_schedule_and_capacity(oracle_mod):
    """Linear-EOS solver: emission every int(2 r / (|v| dt))-th call; with room
    for one emission only, the emitter switches itself off."""
    sc, g, o = _pair(oracle_mod, solver=_abi.SF_SPH_SOLVER)
    k = g.emitter_state()[1]
    for _ in range(30):
        g.UpdateSystem(RI)
        o.update_system(RI)
        assert g.sizes()[0] == o.size()
    assert g.emitter_state()[2] == o.emitter_state()[2] >= 2 * k
    sc2 = scenes.mixed(collider=False)
    g2 = SFSystem.from_scene(sc2, max_num=sc2.n + k + 1)
    g2.set_emitter(**EMIT)
    for _ in range(30):
        g2.UpdateSystem(RI)
    enabled, _, emitted = g2.emitter_state()
    assert emitted == k and not enabled and g2.sizes()[0] == sc2.n + k


def test_emitter_off_and_errors():
    sc = scenes.mixed(collider=False)
    g = SFSystem.from_scene(sc, max_num=sc.n + 1000)
    g.set_emitter(**{**EMIT, "run": False})  # armed but not running: no emission
    g.UpdateSystem(RI)
    assert g.sizes()[0] == sc.n and g.emitter_state()[2] == 0
    g.step(1e-4, 3)  # fixed-dt steps never emit
    assert g.sizes()[0] == sc.n
    with pytest.raises(SFError):
        g.set_emitter(**{**EMIT, "vel": (0.0, 0.0, 0.0)})
    with pytest.raises(SFError):
        g.set_emitter(**{**EMIT, "kind": 7})
    with pytest.raises(SFError):
        g.set_emitter(**{**EMIT, "radius": 1e6})
    with pytest.raises(SFError):
        g.set_emitter(**{**EMIT, "pos": (float("nan"), 0.0, 0.0)})


def test_emitter_sph_solver_near_zero_speed():
    """Linear-EOS solver with an emission speed of 1e-20 m/s: the period
    2 r / (|v| dt) ~ 2e22 lies beyond size_t, so it saturates (instead of
    converting an out-of-range quotient); the first call emits and no later
    one does."""
    sc = scenes.mixed(collider=False)
    p, b, o = sc.params()
    p.solver_type = _abi.SF_SPH_SOLVER
    g = SFSystem.from_scene(sc, max_num=sc.n + 4000)
    g.set_params(p)
    g.set_emitter(**{**EMIT, "vel": (0.0, 0.0, -1e-20)})
    k = g.emitter_state()[1]
    for _ in range(4):
        g.UpdateSystem(RI)
    assert g.emitter_state()[2] == k and g.sizes()[0] == sc.n + k
