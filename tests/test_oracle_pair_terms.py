"""Closed-form pins of the oracle's pair terms that conservation alone does not fix.

Each scene is a handful of particles whose densities, pressures, voidages and
forces follow in closed form from the reference's formulas, evaluated here in
float64 with this file's own kernels (the kernels themselves are pinned by
tests/test_oracle_kernels.py).  The SPH solver's linear EOS keeps negative
pressures, so pressure-driven terms are nonzero for isolated particles.  The
scenes are chosen so that a sign flip, a swapped density (rho instead of
rho / voidage, rho_i instead of rho0), or an i<->j-symmetric wrong factor
changes the expected value by far more than the tolerance.

Reference terms pinned (cuda_sph_sf_solver_gpu.cuh unless noted):
  water pressure term            :343-369
  water artificial viscosity     :371-406
  water seepage buoyancy + drag  :275-312
  sand buoyancy (p != 0)         :240-273
  boundary density               cuda_sph_solver_common_gpu.cuh:70-87
  boundary pressure (rho0 as the density argument, :944)  cuda_sph_solver_common_gpu.cuh:89-108
  boundary viscosity (rho0 as the density argument, :945) cuda_sph_solver_common_gpu.cuh:110-139
  sand DEM against label-1 boundary particles  :52-81 with ComputeDemForces
                                 cuda_dem_solver_common_gpu.cuh:20-51
  wet-sand colour / maxSaturation :952-977
"""
import math

import numpy as np
import pytest

from sph_seepage_flows_amd import params as P

from _scene_util import make_scene

H = float(np.float32(0.04))          # kernel radius as the FP32 parameter struct carries it
EPS = float(np.float32(1e-6))        # KIRI_EPSILON (kiri_pbs_pch.cuh:52), a float literal
W0 = 315.0 / (64.0 * math.pi * H ** 3)


def poly6(r):
    r = np.asarray(r, np.float64)
    return np.where(r * r < H * H, 315.0 / (64.0 * math.pi * H ** 9) * (H * H - r * r) ** 3, 0.0)


def spiky_grad(d):
    d = np.asarray(d, np.float64)
    r = np.linalg.norm(d, axis=-1)
    s = np.where((r < H) & (r >= EPS), -45.0 / (math.pi * H ** 6) * (H - r) ** 2 / np.maximum(r, 1e-30), 0.0)
    return s[..., None] * d


def f64(a):
    """FP32 upload value as the oracle sees it"""
    return np.asarray(np.asarray(a, np.float32), np.float64)


def prm_of(ph):
    p, b, o = P.make_params(ph)
    return p


def by_input_order(o, name):
    """field in upload order (the oracle sorts particles by cell)"""
    return o.get_by_id(name)


def g_vec(ph):
    return f64(ph.gravity)


# --------------------------------------------------------------------------
def test_water_triple_pressure_and_viscosity(oracle_mod):
    """Water A - B - C on a line at spacing d = 0.6 h (A and C out of range):
    rho_A = m (2 W0 + W(d)) differs from rho_B = m (2 W0 + 2 W(d)), so the
    pressure term's p_A / rho_A^2 + p_B / rho_B^2 is distinguishable from any
    symmetric mix such as (p_A + p_B) / (rho_A rho_B).  Linear EOS: pressures
    are negative (isolated particles are far below rho0), so the pressure
    force pulls A toward B.  A second run gives A a velocity toward B: the
    artificial viscosity then decelerates A by m Pi gradW with
    Pi = -nu / (rho_A + rho_B) (v.x) / (|x|^2 + eps)."""
    ph = P.Physics(solver_type=0, damping=0.0)
    prm = prm_of(ph)
    m, rho0, B, nu = float(prm.sph_mass), float(prm.sph_density), float(prm.sph_stiff), float(prm.sph_nu)
    d = 0.6 * H
    x = f64([[0.5, 0.5, 0.5], [0.5 + d, 0.5, 0.5], [0.5 + 2 * d, 0.5, 0.5]])
    r = np.linalg.norm(x[:, None] - x[None], axis=2)
    rho = m * (poly6(0.0) + poly6(r).sum(1))  # self counted twice
    assert rho[0] == pytest.approx(m * (2 * W0 + float(poly6(r[0, 1]))), rel=1e-12)
    assert abs(rho[1] - rho[0]) > 0.02 * rho[0]
    p = B * (rho - rho0)
    assert np.all(p < 0)

    def pressure_acc(i):
        a = np.zeros(3)
        for j in range(3):
            if j != i:
                a += -m * (p[i] / rho[i] ** 2 + p[j] / rho[j] ** 2) * spiky_grad(x[i] - x[j])
        return a

    def visc_acc(i, v):
        a = np.zeros(3)
        for j in range(3):
            if j == i:
                continue
            dp, dv = x[i] - x[j], v[i] - v[j]
            dd = float(dv @ dp)
            if dd < 0:
                pij = -nu / (rho[i] + rho[j]) * (dd / (dp @ dp + EPS))
                a += -m * pij * spiky_grad(dp)
        return a

    g = g_vec(ph)
    # run 1: at rest -> pressure only
    o = oracle_mod.OracleSystem(make_scene(x, [0, 0, 0], ph=ph), 64)
    o.step(1e-4)
    np.testing.assert_allclose(by_input_order(o, "density"), rho, rtol=1e-12)
    np.testing.assert_allclose(by_input_order(o, "pressure"), p, rtol=1e-12)
    acc = by_input_order(o, "acc")
    for i in range(3):
        np.testing.assert_allclose(acc[i], g + pressure_acc(i), rtol=1e-10, atol=1e-12 * np.abs(g).max())
    assert acc[0][0] > 0 and acc[2][0] < 0  # negative pressure: the ends are pulled toward the middle
    # a symmetric wrong factor would move A's x-acceleration by far more than the tolerance
    wrong = -m * (p[0] + p[1]) / (rho[0] * rho[1]) * spiky_grad(x[0] - x[1])[0]
    assert abs(wrong - pressure_acc(0)[0]) > 1e-3 * abs(pressure_acc(0)[0])

    # run 2: A moves toward B, C moves away from B -> viscosity on A and B only
    v = f64([[0.4, 0.1, 0.0], [0.0, 0.0, 0.0], [0.3, 0.0, 0.0]])
    o = oracle_mod.OracleSystem(make_scene(x, [0, 0, 0], vel=v, ph=ph), 64)
    o.step(1e-4)
    acc2 = by_input_order(o, "acc")
    for i in range(3):
        np.testing.assert_allclose(acc2[i], g + pressure_acc(i) + visc_acc(i, v), rtol=1e-10,
                                   atol=1e-12 * np.abs(g).max())
    assert acc2[0][0] < acc[0][0]          # approaching A is slowed
    assert np.all(visc_acc(2, v) == 0)     # separating C: no viscosity
    np.testing.assert_allclose(acc2[2], acc[2], rtol=1e-12)


# --------------------------------------------------------------------------
@pytest.mark.parametrize("offset", [(0.35, 0.0, 0.0), (0.2, -0.25, 0.1)])
def test_water_grain_pair_seepage_and_buoyancy(oracle_mod, offset):
    """One water particle w and one sand grain s at distance < h/2 (no
    adhesion: A(r) = 0 there).  Linear EOS: p_w = B (2 m_w W0 - rho0) < 0.
    Water voidage eps_w = 1 - (m_s / rho_s) W(d); the grain's voidage is eps_w
    (its one water neighbour), its average flow is v_w.  Then
      sand:  f_b = -m_s m_w p_w / (rho_s rho_w / eps_w) gradW(x_s - x_w)
             df  = g rho0 eps^2 / (cd eps^3 / (1-eps)^2) (v_w - v_s) m_s / rho_s
             acc_s = g + 2 (f_b + df) / m_s,  avgDrag_s = df / (m_w eps_w / rho_w W)
      water: acc_w = g - m_s p_w / ((rho_w / eps_w) rho_s) gradW(x_w - x_s)
                       - W / (rho_w / eps_w) avgDrag_s."""
    ph = P.Physics(solver_type=0, damping=0.0)
    prm = prm_of(ph)
    mw, ms = float(prm.sph_mass), float(prm.dem_mass)
    rho0, rhos, B, cd = float(prm.sph_density), float(prm.dem_density), float(prm.sph_stiff), float(prm.sf_cd)
    grav = abs(float(np.float32(ph.gravity[1])))
    xw = f64([0.5, 0.5, 0.5])
    xs = f64(np.float32(xw) + np.float32(H) * np.asarray(offset, np.float32))
    d = float(np.linalg.norm(xs - xw))
    assert 2 * d < H
    vw = f64([0.3, -0.2, 0.1])
    vs = f64([-0.05, 0.0, 0.02])
    Wd = float(poly6(d))
    rho_w = mw * 2 * W0
    p_w = B * (rho_w - rho0)
    assert p_w < 0
    eps_w = 1.0 - ms / rhos * Wd
    assert 0 < eps_w < 1
    eps_s = eps_w
    rho_t = rho_w / eps_w
    f_b = -ms * mw * p_w / (rhos * rho_t) * spiky_grad(xs - xw)
    C = cd * eps_s ** 3 / (1 - eps_s) ** 2
    df = grav * rho0 * eps_s ** 2 / C * (vw - vs) * ms / rhos
    vW = mw / rho_t * Wd
    avgdrag = df / vW
    g = g_vec(ph)
    acc_s = g + 2 * (f_b + df) / ms
    acc_w = g + (-ms * p_w / (rho_t * rhos)) * spiky_grad(xw - xs) - Wd / rho_t * avgdrag

    o = oracle_mod.OracleSystem(make_scene([xw, xs], [0, 1], vel=[vw, vs], ph=ph), 64)
    o.step(1e-4)
    dens, pres = by_input_order(o, "density"), by_input_order(o, "pressure")
    assert dens[0] == pytest.approx(rho_w, rel=1e-12) and dens[1] == rhos
    assert pres[0] == pytest.approx(p_w, rel=1e-12) and pres[1] == 0
    vo = by_input_order(o, "voidage")
    assert vo[0] == pytest.approx(eps_w, rel=1e-12) and vo[1] == pytest.approx(eps_s, rel=1e-12)
    assert by_input_order(o, "saturation")[1] == pytest.approx(mw / rho_w * Wd, rel=1e-12)
    np.testing.assert_allclose(by_input_order(o, "avg_flow_vel")[1], vw, rtol=1e-12)
    assert np.all(by_input_order(o, "adhesion_force") == 0)
    np.testing.assert_allclose(by_input_order(o, "avg_drag_force")[1], avgdrag, rtol=1e-10)
    acc = by_input_order(o, "acc")
    np.testing.assert_allclose(acc[1], acc_s, rtol=1e-10)
    np.testing.assert_allclose(acc[0], acc_w, rtol=1e-10)
    # the voidage enters both sides: with rho_w in place of rho_w / eps_w both
    # accelerations move by (1 - eps_w) of their pair parts, far above the tolerance
    assert abs(1 - eps_w) > 1e-3
    assert np.linalg.norm(acc_w - g) > 1e-3 * np.linalg.norm(g)


# --------------------------------------------------------------------------
def _sheet(y, n_side=13, spacing=0.02, cx=0.5, cz=0.5):
    k = (np.arange(n_side) - (n_side - 1) / 2) * spacing
    gx, gz = np.meshgrid(k + cx, k + cz, indexing="ij")
    return np.stack([gx.ravel(), np.full(gx.size, y), gz.ravel()], 1).astype(np.float32)


@pytest.mark.parametrize("vel,visc", [((0.0, 0.0, 0.0), False), ((0.15, -0.6, 0.05), True),
                                      ((0.05, 0.8, -0.1), False)])
def test_water_over_boundary_sheet(oracle_mod, vel, visc):
    """A water particle 0.45 h above a flat label-0 boundary sheet (spacing 2 r,
    reaching beyond h in every direction), off-centre.  With the volumes V_b
    the oracle computes (pinned by the sheet lattice sum in
    test_oracle_searcher.py):
      rho = 2 m W0 + sum_b rho0 V_b W(|x - x_b|)
      p   = B (rho - rho0)                         (linear EOS)
      a_p = sum_b -rho0 V_b p / rho0^2 gradW(x - x_b)        (rho0, not rho)
      a_v = sum_{b: v.(x-x_b) < 0} -V_b rho0 Pi gradW,  Pi = -bnu / (2 rho0) v.x / (|x|^2 + eps)
    Moving toward the sheet switches the viscosity on; moving away leaves a_p only."""
    ph = P.Physics(solver_type=0, damping=0.0)
    prm = prm_of(ph)
    m, rho0, B, bnu = float(prm.sph_mass), float(prm.sph_density), float(prm.sph_stiff), float(prm.sph_bnu)
    sheet = _sheet(0.5)
    x = f64([0.5 + 0.0037, 0.5 + 0.45 * H, 0.5 - 0.0051])
    v = f64(vel)
    o = oracle_mod.OracleSystem(make_scene([x], [0], vel=[v], ph=ph,
                                           boundary=(sheet, np.zeros(len(sheet), np.uint64))), 64)
    o.step(1e-4)
    bp, bv = o.get("b_pos"), o.get("b_volume")
    assert np.all(bv > 0)
    dp = x[None] - bp
    r = np.linalg.norm(dp, axis=1)
    assert (r < H).sum() >= 8
    rho = m * 2 * W0 + (rho0 * bv * poly6(r)).sum()
    p = B * (rho - rho0)
    assert rho < 0.9 * rho0  # far from rho0: rho in place of rho0 would change a_p by > 20 %
    gw = spiky_grad(dp)
    a = (-rho0 * bv * p / rho0 ** 2)[:, None] * gw
    dd = dp @ v
    on = dd < 0
    assert (on & (r < H)).any() == visc
    pij = -bnu / (2 * rho0) * (dd / ((dp * dp).sum(1) + EPS))
    a_v = np.where(on[:, None], (-bv * rho0 * pij)[:, None] * gw, 0.0)
    expect = g_vec(ph) + a.sum(0) + a_v.sum(0)
    assert o.get("density")[0] == pytest.approx(rho, rel=1e-12)
    assert o.get("pressure")[0] == pytest.approx(p, rel=1e-12)
    np.testing.assert_allclose(o.get("acc")[0], expect, rtol=1e-10, atol=1e-12)
    # negative pressure pulls the particle down onto the sheet
    assert a.sum(0)[1] < 0
    if visc:
        assert a_v.sum(0)[1] > 0  # moving down: the boundary viscosity pushes back up


# --------------------------------------------------------------------------
def _dem_force(dij, vij, rij, kn, ks, tanphi):
    """ComputeDemForces (cuda_dem_solver_common_gpu.cuh:20-51) for one pair"""
    dist = float(np.linalg.norm(dij))
    pen = rij - dist
    if pen <= 0:
        return np.zeros(3)
    n = dij / dist
    vt = vij - (vij @ n) * n
    fn = kn * pen * n
    fs = -ks * vt
    maxfs = (fn @ fn) * tanphi ** 2
    if fs @ fs > maxfs:
        fs = fs * math.sqrt(maxfs) / math.sqrt(fs @ fs)
    return -fn - fs


@pytest.mark.parametrize("vel", [(0.0, 0.0, 0.0), (0.003, -0.001, 0.002), (0.5, 0.0, -0.3)])
def test_grain_on_label1_boundary_particle(oracle_mod, vel):
    """A sand grain overlapping one label-1 boundary particle by 0.5 R
    (ComputeSFUniRadiusDemBoundaryForces, :52-81): dij = x_s - x_b and
    vij = -v_s, so the normal part is -kn (2R - |dij|) dij / |dij| -- directed
    from the grain TOWARD the boundary particle, the reference's sign as
    written (sand-sand pairs use dij = x_j - x_i and repel).  The shear part
    opposes the grain's tangential velocity and is capped at tan(phi) |Fn|
    (the last velocity exceeds the cap).  acc = g + 2 f / m (damping off).
    The same particle with label 0 exerts nothing."""
    ph = P.Physics(damping=0.0)
    prm = prm_of(ph)
    R, ms = float(prm.dem_particle_radius), float(prm.dem_mass)
    kn = float(prm.dem_young) * R
    ks = kn * float(prm.dem_poisson)
    tanphi = float(prm.dem_tan_friction_angle)
    xb = f64([0.5, 0.5, 0.5])
    u = np.array([0.6, 0.7, -0.39])
    u /= np.linalg.norm(u)
    xs = f64(np.float32(xb) + np.float32(1.5 * R) * u.astype(np.float32))
    vs = f64(vel)
    f = _dem_force(xs - xb, -vs, 2 * R, kn, ks, tanphi)
    pen = 2 * R - np.linalg.norm(xs - xb)
    assert pen > 0.4 * R
    n_hat = (xs - xb) / np.linalg.norm(xs - xb)
    assert f @ n_hat < 0  # toward the boundary particle
    g = g_vec(ph)
    for lab, expect in ((1, g + 2 * f / ms), (0, g)):
        o = oracle_mod.OracleSystem(make_scene([xs], [1], vel=[vs], ph=ph,
                                               boundary=(xb[None].astype(np.float32), np.array([lab], np.uint64))),
                                    64)
        o.step(1e-4)
        np.testing.assert_allclose(o.get("acc")[0], expect, rtol=1e-10, atol=1e-12)
    if np.linalg.norm(vs) > 0.1:  # capped shear: |Fs| = tan(phi) |Fn|
        fs = f + kn * pen * n_hat
        assert np.linalg.norm(fs) == pytest.approx(tanphi * kn * pen, rel=1e-9)


# --------------------------------------------------------------------------
def test_wet_sand_colour_both_branches(oracle_mod):
    """Two far-apart (water, grain) pairs at distance d: each grain's
    saturation is (m_w / rho_w) W(d) = W(d) / (2 W0).  Grain 1 has
    maxSaturation 0 < sat: colour = dry + (wet - dry) sat, maxSat = sat.
    Grain 2 has maxSaturation 0.9 > sat: colour and maxSat keep their values.
    Water colours never change (ComputeSFWetSandColor_CUDA, :952-977)."""
    ph = P.Physics()
    d = 0.3 * H
    pos = f64([[0.3, 0.5, 0.5], [0.3 + d, 0.5, 0.5], [0.7, 0.5, 0.5], [0.7 + d, 0.5, 0.5]])
    sc = make_scene(pos, [0, 1, 0, 1], ph=ph)
    col0 = np.array([[0.1, 0.2, 0.3], [0.11, 0.22, 0.33], [0.4, 0.5, 0.6], [0.44, 0.55, 0.66]], np.float32)
    sc.col = col0.copy()
    sc.max_sat = np.array([0.0, 0.0, 0.0, 0.9], np.float32)
    o = oracle_mod.OracleSystem(sc, 64)
    o.step(1e-4)
    sat, sat2 = (float(poly6(np.linalg.norm(pos[k + 1] - pos[k]))) / (2 * W0) for k in (0, 2))
    assert 0.2 < sat < 0.9 and 0.2 < sat2 < 0.9
    s = by_input_order(o, "saturation")
    assert s[1] == pytest.approx(sat, rel=1e-12) and s[3] == pytest.approx(sat2, rel=1e-12)
    dry, wet = f64(P.DRY_SAND), f64(P.WET_SAND)
    col, msat = by_input_order(o, "col"), by_input_order(o, "max_saturation")
    np.testing.assert_allclose(col[1], dry + (wet - dry) * sat, rtol=1e-12)
    assert msat[1] == pytest.approx(sat, rel=1e-12)
    np.testing.assert_array_equal(col[3], f64(col0[3]))
    assert msat[3] == pytest.approx(float(np.float32(0.9)), rel=1e-15)
    np.testing.assert_array_equal(col[[0, 2]], f64(col0[[0, 2]]))


def test_mixed_scene_runs_both_colour_branches(oracle_mod):
    """The mixed parity scene carries a saturation history: after one step
    some grains are wetter than their history (colour and maxSat updated) and
    some drier (kept), so the GPU-vs-oracle step tests cover both branches."""
    from sph_seepage_flows_amd import scenes

    sc = scenes.mixed()
    o = oracle_mod.OracleSystem(sc, 64)
    o.step(1e-4)
    sand = np.asarray(sc.label) == 1
    sat = o.get_by_id("saturation")[sand]
    ms0 = f64(sc.max_sat)[sand]
    ms1 = o.get_by_id("max_saturation")[sand]
    col0 = f64(sc.col)[sand]
    col1 = o.get_by_id("col")[sand]
    wetter = ms0 < sat
    assert 50 < wetter.sum() < sand.sum() - 50
    np.testing.assert_array_equal(ms1[~wetter], ms0[~wetter])
    np.testing.assert_array_equal(col1[~wetter], col0[~wetter])
    np.testing.assert_allclose(ms1[wetter], sat[wetter], rtol=1e-15)
    dry, wet = f64(P.DRY_SAND), f64(P.WET_SAND)
    np.testing.assert_allclose(col1[wetter], dry + (wet - dry) * sat[wetter][:, None], rtol=1e-12)
