"""Pins the oracle's kernel functions (cuda_sph_kernel.cuh:79-118,189-203,
cuda_sf_utils.cuh:36-82, cuda_dem_solver_common_gpu.cuh:20-199) against
closed forms, normalisations and independent numerical derivations."""
import math

import numpy as np
import pytest
from scipy import integrate, optimize

H = 0.04


def test_poly6_closed_form_and_normalisation(oracle_mod):
    w0 = 315.0 / (64.0 * math.pi * H ** 3)
    assert oracle_mod.poly6(H, 0.0) == pytest.approx(w0, rel=1e-12)
    assert oracle_mod.poly6(H, H) == 0.0
    assert oracle_mod.poly6(H, 1.5 * H) == 0.0
    # 4 pi int_0^h W(r) r^2 dr = 1
    val, _ = integrate.quad(lambda r: oracle_mod.poly6(H, r) * r * r, 0.0, H, epsabs=1e-14)
    assert 4 * math.pi * val == pytest.approx(1.0, rel=1e-9)


def spiky(r):
    """Normalised spiky kernel 15 / (pi h^6) (h - r)^3 (its gradient is SpikyKernelGrad)."""
    return 15.0 / (math.pi * H ** 6) * (H - r) ** 3 if r < H else 0.0


def test_spiky_normalisation_and_gradient_is_its_derivative(oracle_mod):
    val, _ = integrate.quad(lambda r: spiky(r) * r * r, 0.0, H, epsabs=1e-14)
    assert 4 * math.pi * val == pytest.approx(1.0, rel=1e-9)
    rng = np.random.default_rng(1)
    for _ in range(20):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        r = rng.uniform(0.05, 0.95) * H
        g = oracle_mod.spiky_grad(H, d * r)
        e = 1e-9
        dwdr = (spiky(r + e) - spiky(r - e)) / (2 * e)
        np.testing.assert_allclose(g, dwdr * d, rtol=1e-6)
    assert np.all(oracle_mod.spiky_grad(H, [H, 0, 0]) == 0)
    assert np.all(oracle_mod.spiky_grad(H, [1e-7, 0, 0]) == 0)  # below KIRI_EPSILON


def test_adhesion_special_values(oracle_mod):
    c = 0.007 / H ** 3.25
    assert oracle_mod.adhesion(H, 0.5 * H) == 0.0          # r <= h/2
    assert oracle_mod.adhesion(H, 0.3 * H) == 0.0
    assert oracle_mod.adhesion(H, 1.01 * H) == 0.0
    # -4h + 6h - 2h = 0 at r = h: the limit is 0 (exactly at r = h the
    # reference's arithmetic can cancel to a tiny negative and pow() gives NaN)
    assert oracle_mod.adhesion(H, H * (1 - 1e-9)) == pytest.approx(0.0, abs=c * H ** 0.25 * 1e-2)
    # maximum at r = 3h/4 where -4r^2/h + 6r - 2h = h/4
    assert oracle_mod.adhesion(H, 0.75 * H) == pytest.approx(c * (H / 4) ** 0.25, rel=1e-12)
    res = optimize.minimize_scalar(lambda r: -oracle_mod.adhesion(H, r), bounds=(0.51 * H, 0.99 * H),
                                   method="bounded", options={"xatol": 1e-12})
    assert res.x == pytest.approx(0.75 * H, rel=1e-5)


@pytest.mark.parametrize("ctl", [(0.7, 1.0, 0.01, 0.0), (0.0, 1.5, 0.5, 0.8), (0.2, 0.9, 0.25, 0.4)])
def test_bezier_against_parametric_curve(oracle_mod, ctl):
    py0, py1, px1, py2 = ctl
    G = lambda s: oracle_mod.bezier(py0, py1, px1, py2, s)
    assert G(-0.5) == py0 and G(0.0) == pytest.approx(py0)
    assert G(px1) == pytest.approx(py1, rel=1e-12)
    assert G(1.0) == py2 and G(1.7) == py2
    # second branch: quadratic Bezier from (px1, py1) through control (xc, py1) to (1, py2)
    xc = px1 + (1.0 + px1) / 4.0
    for s in np.linspace(px1 + 1e-6, 1 - 1e-6, 17):
        roots = np.roots([px1 - 2 * xc + 1.0, 2 * (xc - px1), px1 - s])
        t = [r.real for r in roots if abs(r.imag) < 1e-12 and -1e-9 <= r.real <= 1 + 1e-9][0]
        y = (1 - t) ** 2 * py1 + 2 * t * (1 - t) * py1 + t * t * py2
        assert G(s) == pytest.approx(y, rel=1e-9, abs=1e-12)
    # first branch, parameter t = s / px1 (the reference blends t^2 * py1 here, cuda_sf_utils.cuh:76)
    for s in np.linspace(0.0, px1, 7):
        t = s / px1
        assert G(s) == pytest.approx((1 - t) ** 2 * py0 + 2 * t * (1 - t) * py1 + t * t * py1, rel=1e-12)


def test_dem_head_on_and_friction_cap(oracle_mod):
    R, E, nu, mu = 0.01, 1e6, 0.3, 0.5
    kn, ks = E * R, E * R * nu
    d = 0.019  # overlap 0.001 along +x
    f = oracle_mod.dem_force([d, 0, 0], [0, 0, 0], 2 * R, kn, ks, mu)
    np.testing.assert_allclose(f, [-kn * (2 * R - d), 0, 0], rtol=1e-12)
    assert np.all(oracle_mod.dem_force([0.021, 0, 0], [0, 3, 0], 2 * R, kn, ks, mu) == 0)  # no overlap
    # small tangential speed: viscous shear -ks * v_t on j->i, so f = +ks * v_t... sign: f = -Fn - Fs, Fs = -ks v_t
    vt = 1e-4
    f = oracle_mod.dem_force([d, 0, 0], [0, vt, 0], 2 * R, kn, ks, mu)
    np.testing.assert_allclose(f, [-kn * 0.001, ks * vt, 0], rtol=1e-12)
    # large tangential speed: capped at mu * |Fn|
    f = oracle_mod.dem_force([d, 0, 0], [0, 10.0, 0], 2 * R, kn, ks, mu)
    np.testing.assert_allclose(f, [-kn * 0.001, mu * kn * 0.001, 0], rtol=1e-12)
    # the normal relative velocity does not enter (no dashpot)
    f2 = oracle_mod.dem_force([d, 0, 0], [5.0, 0, 0], 2 * R, kn, ks, mu)
    np.testing.assert_allclose(f2, [-kn * 0.001, 0, 0], rtol=1e-12)
    # pair antisymmetry
    rng = np.random.default_rng(2)
    for _ in range(10):
        dv = rng.normal(size=3) * 0.01
        vv = rng.normal(size=3)
        np.testing.assert_allclose(oracle_mod.dem_force(dv, vv, 2 * R, kn, ks, mu),
                                   -oracle_mod.dem_force(-dv, -vv, 2 * R, kn, ks, mu), rtol=1e-12, atol=1e-18)


def test_capillary_bridge_limits(oracle_mod):
    R = 0.01
    bez = [0.7, 1.0, 0.01, 0.0]
    theta = math.pi / 6
    vb = 4.0 / 3.0 * math.pi * R ** 3 * 1e-4
    srup = (1 + theta / 2) * (vb ** (1 / 3) + 0.1 * vb ** (2 / 3))  # rupture distance
    assert oracle_mod.capillary_rupture(R) == pytest.approx(srup, rel=1e-12)
    c = oracle_mod.bezier(*bez, 0.3)
    # H -> 0+: F -> 2 pi c R (cos(theta) + phi0 sin(theta)), phi0^2 = 2 sqrt(vb / (pi R)) / R
    phi0 = math.sqrt(2 * math.sqrt(vb / (math.pi * R)) / R)
    H0 = 1e-12
    f = oracle_mod.capillary_force([2 * R + H0, 0, 0], R, 0.3, bez)
    assert f[0] == pytest.approx(2 * math.pi * c * R * (math.cos(theta) + phi0 * math.sin(theta)), rel=1e-4)
    assert f[0] > 0 and f[1] == 0 and f[2] == 0  # attraction of i toward j (+x)
    # zero outside (0, s_rupture), monotone decay inside
    assert np.all(oracle_mod.capillary_force([2 * R, 0, 0], R, 0.3, bez) == 0)
    assert np.all(oracle_mod.capillary_force([2 * R + srup * 1.0001, 0, 0], R, 0.3, bez) == 0)
    fs = [oracle_mod.capillary_force([2 * R + h, 0, 0], R, 0.3, bez)[0] for h in np.linspace(1e-9, srup * 0.999, 40)]
    assert np.all(np.diff(fs) < 0)
    # linear in the coefficient curve
    f2 = oracle_mod.capillary_force([2 * R + srup / 3, 0, 0], R, 0.3, [2 * b for b in bez[:2]] + [bez[2], 2 * bez[3]])
    np.testing.assert_allclose(f2, 2 * oracle_mod.capillary_force([2 * R + srup / 3, 0, 0], R, 0.3, bez), rtol=1e-12)


def test_world_walls(oracle_mod):
    R, Rb, E, nu, mu = 0.01, 0.01, 1e6, 0.3, 0.5
    kn, rij = E * R, R + Rb
    lo, hi = [0, 0, 0], [1, 1, 1]
    free = oracle_mod.world_boundary_force([0.5, 0.5, 0.5], [1, 1, 1], R, Rb, E, nu, mu, lo, hi)
    assert np.all(free == 0)
    # 5 mm from the x-high wall: normal push kn * (rij - 0.005) toward -x
    f = oracle_mod.world_boundary_force([0.995, 0.5, 0.5], [0, 0, 0], R, Rb, E, nu, mu, lo, hi)
    np.testing.assert_allclose(f, [-kn * (rij - 0.005), 0, 0], rtol=1e-12)
    f = oracle_mod.world_boundary_force([0.5, 0.004, 0.5], [0, 0, 0], R, Rb, E, nu, mu, lo, hi)
    np.testing.assert_allclose(f, [0, kn * (rij - 0.004), 0], rtol=1e-12)
    # z-low wall triggers inside 2 rij but pushes only inside rij (cuda_dem_solver_common_gpu.cuh:191)
    assert np.all(oracle_mod.world_boundary_force([0.5, 0.5, 1.5 * rij], [0, 0, 0], R, Rb, E, nu, mu, lo, hi) == 0)
    f = oracle_mod.world_boundary_force([0.5, 0.5, 0.5 * rij], [0, 0, 0], R, Rb, E, nu, mu, lo, hi)
    np.testing.assert_allclose(f, [0, 0, kn * 0.5 * rij], rtol=1e-12)
    # sliding along the floor: friction opposes the motion, capped at mu * normal
    f = oracle_mod.world_boundary_force([0.5, 0.004, 0.5], [3.0, 0, 0], R, Rb, E, nu, mu, lo, hi)
    np.testing.assert_allclose(f, [-mu * kn * (rij - 0.004), kn * (rij - 0.004), 0], rtol=1e-12)
