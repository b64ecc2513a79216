"""The FP64 host tangent (the oracle instantiated with forward-mode dual
numbers) against central finite differences of the FP64 oracle itself.

The reference's model is only piecewise smooth: the sand damping switches
with sgn(a (v + dt a / 2)), liquid bridges switch on and off at contact and
rupture, and the water-sand adhesion (-4r^2/h + 6r - 2h)^(1/4) has unbounded
slope at r = h/2 and r = h.  Finite differences therefore converge to the
tangent only as the step shrinks below the distance to the nearest switch,
and in packed sand the reference's shear term (-k_s v_t with dt 2 k_s / m of
several hundred) amplifies perturbations by orders of magnitude within a few
steps.  The checks use two-step windows of the full model, where central
differences with FP64 steps of 1e-8 converge to the tangent (~1e-6)."""
import copy

import numpy as np
import pytest

from sph_seepage_flows_amd import scenes


def scene_of(name):
    if name == "column":
        return small_column()
    return scenes.mixed(collider=False)  # water and sand interpenetrate: drag, buoyancy, adhesion, seepage


def small_column(smooth=False):
    sc = scenes.column(n_sand=1500, n_water=500, width=0.06, height=0.1)
    if smooth:
        sc.physics.sf_a0 = 0.0
        sc.physics.sf_amc = 0.0
        sc.physics.sf_asat = 0.0
        sc.physics.damping = 0.0
    return sc


def test_tangent_values_equal_the_fp64_oracle(oracle_mod):
    sc = small_column(smooth=False)
    pos, sat, _, _ = oracle_mod.tangent_window(sc, 1e-4, 5)
    o = oracle_mod.OracleSystem(sc, 64)
    o.step(1e-4, 5)
    np.testing.assert_array_equal(pos, o.get_by_id("pos"))
    np.testing.assert_array_equal(sat, o.get_by_id("saturation"))


@pytest.mark.parametrize("name", ["column", "mixed"])
def test_tangent_matches_central_differences(name, oracle_mod):
    sc = scene_of(name)
    steps, dt = 2, 1e-4
    rng = np.random.default_rng(3)
    n = sc.n
    u = rng.standard_normal((n, 3))
    ux = rng.standard_normal((n, 3)) * 1e-3
    w = rng.standard_normal((n, 3))
    ws = rng.standard_normal(n) * 0.01
    vdot = np.zeros((2, n, 3)); vdot[0] = u
    xdot = np.zeros((2, n, 3)); xdot[1] = ux
    _, _, pd, sd = oracle_mod.tangent_window(sc, dt, steps, vdot=vdot, xdot=xdot)
    L = lambda o: float((w * o.get_by_id("pos")).sum() + (ws * o.get_by_id("saturation")).sum())

    def run(pos, vel):
        o = oracle_mod.OracleSystem(sc, 64)
        o.set_state(pos, vel)
        o.step(dt, steps)
        return L(o)

    p64, v64 = sc.pos.astype(np.float64), sc.vel.astype(np.float64)
    for q, (dp, dv) in enumerate(((0 * u, u), (ux, 0 * u))):
        tan = float((w * pd[q]).sum() + (ws * sd[q]).sum())
        h = 1e-8
        fd = (run(p64 + h * dp, v64 + h * dv) - run(p64 - h * dp, v64 - h * dv)) / (2 * h)
        print(f"direction {q}: tangent {tan:.8e} fd {fd:.8e}")
        assert abs(tan - fd) <= 1e-5 * abs(tan) + 1e-9


@pytest.mark.parametrize("name", ["column", "mixed"])
def test_parameter_tangents_match_central_differences(name, oracle_mod):
    sc = scene_of(name)
    steps, dt = 2, 1e-4
    rng = np.random.default_rng(4)
    w = rng.standard_normal((sc.n, 3))
    ws = rng.standard_normal(sc.n)
    pdot = np.eye(4)
    _, _, pd, sd = oracle_mod.tangent_window(sc, dt, steps, pdot=pdot)
    p, b, o = sc.params()
    names = ("sf_cd", "sf_c0", "sf_cmc", "dem_tan_friction_angle")
    for q, name in enumerate(names):
        tan = float((w * pd[q]).sum() + (ws * sd[q]).sum())
        v = getattr(p, name)
        # parameters are FP32 fields: the step actually taken is used.  The
        # friction step stays small: it moves contacts across the Coulomb cap.
        rel = 1e-6 if name == "dem_tan_friction_angle" else 1e-4
        h = np.float32(max(abs(v), 1e-2) * rel)

        def L(sign):
            pp = copy.copy(p)
            setattr(pp, name, float(np.float32(v) + sign * h))
            orc = oracle_mod.OracleSystem(sc, 64, params=(pp, b, o))
            orc.step(dt, steps)
            return float((w * orc.get_by_id("pos")).sum() + (ws * orc.get_by_id("saturation")).sum())

        step = (np.float32(np.float32(v) + h) - np.float32(np.float32(v) - h)) / 2
        lp, lm = L(1), L(-1)
        fd = (lp - lm) / (2 * float(step))
        # FP64 rounding floor of the difference quotient (sum of ~3n terms of size |L|)
        noise = 10 * np.finfo(np.float64).eps * (abs(lp) + float(np.abs(w).sum()) * 0.2) / (2 * float(step))
        print(f"{name}: tangent {tan:.8e} fd {fd:.8e} (rounding floor {noise:.1e})")
        assert abs(tan - fd) <= 1e-4 * max(abs(tan), abs(fd)) + noise, name
