"""GPU tests of the differentiable window (sf_diff_*): the hand-written
reverse passes against the FP64 host tangent, checkpoint invariance, the
torch.autograd.Function wrapper and the error paths.

The adjoint is checked through dot-product identities: for random seeds
(w, ws) on the final positions / last-step saturations and random
directions u, <adjoint(w, ws), u> from the GPU must equal
<w, J u> + <ws, J_sat u> from the FP64 tangent of the oracle (itself pinned
against central differences in test_oracle_tangent.py).  The GPU linearises
its FP32 trajectory and the oracle its FP64 one; packed sand amplifies their
difference quickly (the reference's shear term), so the windows with sand are
two steps long; the water-only window runs 40 steps."""
import copy

import numpy as np
import pytest

from sph_seepage_flows_amd import SFError, _abi, scenes
from sph_seepage_flows_amd.diff import DiffWindow, msd_loss, physical_grad, theta_of

pytestmark = pytest.mark.gpu

DT = 1e-4
# A direction is checked relative to its own FP64 tangent when that tangent
# is at least TAU times the largest of the window's directions; below that it
# is insensitive (the parameter has no effect the window can show, e.g. drag
# before any water touches the sand) and its error is measured against
# TAU times the largest.
TAU = 1e-4
TOL = 1e-3  # relative, per sensitive direction


def small_column():
    return scenes.column(n_sand=1500, n_water=500, width=0.06, height=0.1)


def water_only():
    sc = small_column()
    keep = sc.label == 0
    s = copy.copy(sc)
    for f in ("pos", "label", "mass", "radius", "col", "vel", "cda0asat", "amcamcp"):
        setattr(s, f, getattr(sc, f)[keep])
    s.name = "water_only"
    return s


SCENES = {"column": small_column, "mixed": lambda: scenes.mixed(collider=False), "water": water_only}


def dot_checks(sc, steps, K, oracle_mod, seed=0):
    n = sc.n
    print(f"[diff] {sc.name}: n={n} steps={steps} K={K}", flush=True)
    win = DiffWindow(sc, steps, DT, K)
    win.forward()
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((n, 3)).astype(np.float32)
    ws = rng.standard_normal(n).astype(np.float32)
    th, gx0, gv0 = win.backward(w, ws)
    u = rng.standard_normal((n, 3))
    ux = rng.standard_normal((n, 3)) * 1e-3
    vdot = np.zeros((4, n, 3)); vdot[0] = u
    xdot = np.zeros((4, n, 3)); xdot[1] = ux
    pdot = np.zeros((4, 4)); pdot[2, 0] = 1.0; pdot[3, 3] = 1.0
    _, _, pd, sd = oracle_mod.tangent_window(sc, DT, steps, vdot=vdot, pdot=pdot, xdot=xdot)
    _, _, pd2, sd2 = oracle_mod.tangent_window(sc, DT, steps, pdot=np.array([[0, 1, 0, 0], [0, 0, 1, 0]], float))
    J = lambda a, b: float((w * a).sum() + (ws * b).sum())
    return {
        "vel0": (float((gv0 * u).sum()), J(pd[0], sd[0])),
        "pos0": (float((gx0 * ux).sum()), J(pd[1], sd[1])),
        "sf_cd": (float(th[0]), J(pd[2], sd[2])),
        "tan_friction": (float(th[3]), J(pd[3], sd[3])),
        "sf_c0": (float(th[1]), J(pd2[0], sd2[0])),
        "sf_cmc": (float(th[2]), J(pd2[1], sd2[1])),
    }


def check_directions(name, res, tol=TOL):
    scale = max(abs(v[1]) for v in res.values())
    bad = {}
    for k, (gpu, ref) in res.items():
        sensitive = abs(ref) >= TAU * scale
        err = abs(gpu - ref) / (abs(ref) if sensitive else TAU * scale)
        print(f"[diff] {name} {k:13s} adjoint {gpu: .8e} tangent {ref: .8e} "
              f"{'rel' if sensitive else 'insensitive, err / (TAU max)'} {err:.2e}", flush=True)
        if not err <= tol:
            bad[k] = (gpu, ref, err)
    assert not bad, bad
    return scale


@pytest.mark.parametrize("name,steps", [("column", 2), ("mixed", 2), ("water", 40)])
def test_adjoint_matches_fp64_tangent(name, steps, oracle_mod):
    sc = SCENES[name]()
    res = dot_checks(sc, steps, 2, oracle_mod)
    check_directions(name, res)


def column_loss_window(sc, steps):
    """The configuration's loss on a window of the column: mean squared
    deviation of the final sand positions and saturations from the window
    run with the hidden parameters (drag D 4e3, capillary sigma 0.03,
    friction 0.55); parameters at (5e3, 0.02, 0.5).  Returns the window run
    at those parameters, the loss and its seeds."""
    sand = sc.label == 1
    hidden = DiffWindow(sc, steps, DT, 1)
    tp, ts = hidden.forward(theta_of(4e3, 0.03, 0.55))
    win = DiffWindow(sc, steps, DT, 1)
    pos, sat = win.forward(theta_of(5e3, 0.02, 0.5))
    loss, pb, sb = msd_loss(pos, sat, tp, ts, sand)
    return win, (tp, ts, sand), loss, pb, sb


def test_column_64k_adjoint_matches_fp64_tangent(oracle_mod):
    """The differentiable column at its full size (48k sand + 16k water, the
    configuration's scene), a two-step window of the configuration's loss:
    the adjoint of every direction (a random initial-velocity direction, the
    drag, the two capillary controls, friction) against the FP64 tangent of
    the same loss at the same parameters."""
    sc = scenes.column()
    steps = 2
    win, _, loss, pb, sb = column_loss_window(sc, steps)
    th, _, gv = win.backward(pb, sb)
    p, b, o = sc.params()
    for name, v in zip(("sf_cd", "sf_c0", "sf_cmc", "dem_tan_friction_angle"), theta_of(5e3, 0.02, 0.5)):
        setattr(p, name, float(v))
    rng = np.random.default_rng(0)
    u = rng.standard_normal((sc.n, 3))
    vdot = np.zeros((4, sc.n, 3)); vdot[0] = u
    pdot = np.zeros((4, 4)); pdot[1, 0] = 1.0; pdot[2, 1] = 1.0; pdot[3, 3] = 1.0
    _, _, pd, sd = oracle_mod.tangent_window(sc, DT, steps, vdot=vdot, pdot=pdot, params=(p, b, o))
    _, _, pd2, sd2 = oracle_mod.tangent_window(sc, DT, steps, pdot=np.array([[0, 0, 1.0, 0]]), params=(p, b, o))
    J = lambda a, c: float((pb * a).sum() + (sb * c).sum())
    res = {"vel0": (float((gv * u).sum()), J(pd[0], sd[0])), "sf_cd": (float(th[0]), J(pd[1], sd[1])),
           "sf_c0": (float(th[1]), J(pd[2], sd[2])), "sf_cmc": (float(th[2]), J(pd2[0], sd2[0])),
           "tan_friction": (float(th[3]), J(pd[3], sd[3]))}
    print(f"[diff] column_64k loss {loss:.6e}", flush=True)
    check_directions("column_64k", res)


def test_column_64k_gradient_vs_central_differences():
    """Central differences of the FP32 GPU forward at h = 1e-3 (relative) of
    the configuration's loss on the 64k column, one-step window, against the
    adjoint, for drag D, capillary sigma and friction mu.  Friction is the
    direction the window resolves (measured 1.1e-3 at h = 1e-3, 7.8e-4 at
    h = 1e-2: the FP32 rounding of the positions sets that floor).  Drag has
    no effect before water reaches the sand (differences exactly 0), and
    sigma's one-step effect on the positions is below their FP32 resolution
    (its differences are rounding noise near 1e-4 of friction's): both are
    insensitive here, their adjoint below TAU of the largest.  Over longer
    windows the FP32 loss is not differentiable at the scale of h (liquid
    bridges form and rupture, shear caps switch), so the differences do not
    converge; the FP64 tangent checks the windows above."""
    sc = scenes.column()
    win, (tp, ts, sand), loss, pb, sb = column_loss_window(sc, 1)
    th, _, _ = win.backward(pb, sb)
    D0, S0, M0 = 5e3, 0.02, 0.5
    g = physical_grad(th, D0)
    fd = {}
    for name, base in (("drag_D", D0), ("capillary_sigma", S0), ("friction_mu", M0)):
        h = 1e-3 * base
        L = []
        for sgn in (1, -1):
            a = {"drag_D": D0, "capillary_sigma": S0, "friction_mu": M0}
            a[name] = base + sgn * h
            p, s = win.forward(theta_of(a["drag_D"], a["capillary_sigma"], a["friction_mu"]))
            L.append(msd_loss(p, s, tp, ts, sand)[0])
        fd[name] = (L[0] - L[1]) / (2 * h)
    scale = max(abs(v) for v in fd.values())
    for name in fd:
        print(f"[fd] column_64k T=1 {name:16s} adjoint {g[name]: .6e} central difference {fd[name]: .6e}", flush=True)
    assert abs(fd["friction_mu"]) == scale
    assert abs(g["friction_mu"] - fd["friction_mu"]) <= 2e-3 * abs(fd["friction_mu"])
    for name in ("drag_D", "capillary_sigma"):
        assert abs(g[name]) <= TAU * scale  # insensitive in the adjoint
        assert abs(fd[name]) <= 10 * TAU * scale  # and no effect above the differences' noise
    assert fd["drag_D"] == 0.0


def test_checkpoint_spacing_does_not_change_gradients():
    sc = scenes.mixed(collider=False)
    rng = np.random.default_rng(1)
    w = rng.standard_normal((sc.n, 3)).astype(np.float32)
    ws = rng.standard_normal(sc.n).astype(np.float32)
    out = []
    for K in (1, 3, 7):
        win = DiffWindow(sc, 7, DT, K)
        pos, sat = win.forward()
        out.append((pos, sat) + win.backward(w, ws))
    for o in out[1:]:
        for a, b in zip(out[0], o):
            np.testing.assert_array_equal(a, b)


def test_window_forward_matches_plain_steps():
    sc = scenes.mixed(collider=False)
    win = DiffWindow(sc, 9, DT, 4)
    pos, sat = win.forward()
    from sph_seepage_flows_amd import SFSystem
    g = SFSystem.from_scene(sc)
    g.step(DT, 9)
    # the plain handle keeps cell order; compare as sets through the label/mass-free sort
    np.testing.assert_array_equal(np.sort(pos, axis=0), np.sort(g.get("pos"), axis=0))
    # repeated windows from the reset state are identical
    pos2, sat2 = win.forward()
    np.testing.assert_array_equal(pos, pos2)
    np.testing.assert_array_equal(sat, sat2)


def test_torch_autograd_function_matches_the_c_abi():
    torch = pytest.importorskip("torch")
    from sph_seepage_flows_amd.diff import seepage_window

    sc = small_column()
    steps = 6
    sand = sc.label == 1
    hidden = DiffWindow(sc, steps, DT, 3)
    tp, ts = hidden.forward(theta_of(4e3, 0.03, 0.55))
    theta0 = theta_of(5e3, 0.02, 0.5)
    # numpy path
    win = DiffWindow(sc, steps, DT, 3)
    pos, sat = win.forward(theta0, sc.vel)
    loss, pb, sb = msd_loss(pos, sat, tp, ts, sand)
    th, _, gv = win.backward(pb, sb)
    # torch path
    win2 = DiffWindow(sc, steps, DT, 3)
    dev = torch.device("cuda:0")
    theta = torch.tensor(theta0, requires_grad=True)
    vel0 = torch.tensor(sc.vel, device=dev, requires_grad=True)
    for _ in range(2):  # second call goes through sf_diff_reset with a device velocity
        theta.grad = None
        vel0.grad = None
        pt, st = seepage_window(win2, theta, vel0)
        m = torch.tensor(sand, device=dev)
        ns = float(sand.sum())
        dp = (pt - torch.tensor(tp, device=dev)) * m[:, None]
        ds = (st - torch.tensor(ts, device=dev)) * m
        lt = (dp.double() ** 2).sum() / ns + (ds.double() ** 2).sum() / ns
        lt.backward()
        assert abs(lt.item() - loss) <= 1e-6 * abs(loss)
        np.testing.assert_allclose(theta.grad.numpy(), th, rtol=1e-5, atol=1e-12)
        np.testing.assert_allclose(vel0.grad.cpu().numpy(), gv, rtol=1e-5, atol=1e-9)
    g = physical_grad(th, 5e3)
    assert set(g) == {"drag_D", "capillary_sigma", "friction_mu"} and all(np.isfinite(list(g.values())))


def test_diff_errors():
    from sph_seepage_flows_amd import BatchSystem, SFSystem
    import ctypes as C

    sc = scenes.mixed()  # has label-1 collider particles
    g = SFSystem.from_scene(sc)
    with pytest.raises(SFError) as e:
        _abi.check(_abi.lib().sf_diff_forward(g.handle, DT, 2, 1))
    assert e.value.code == _abi.SF_E_STATE
    m = SFSystem.from_scene(scenes.mixed(sf_type=1, collider=False))
    with pytest.raises(SFError):
        _abi.check(_abi.lib().sf_diff_forward(m.handle, DT, 2, 1))
    ok = SFSystem.from_scene(scenes.mixed(collider=False))
    with pytest.raises(SFError):  # backward before forward
        _abi.check(_abi.lib().sf_diff_backward(ok.handle, None, None, None, None, None, None, 0))
    with pytest.raises(SFError):
        _abi.check(_abi.lib().sf_diff_forward(ok.handle, DT, 0, 1))
    E = scenes.ensemble(n_scenes=2, n_min=2000, n_max=3000)
    b = BatchSystem.from_scenes(E)
    with pytest.raises(SFError):
        _abi.check(_abi.lib().sf_diff_forward(b.handle, DT, 2, 1))


def test_upload_releases_the_window():
    """A recorded window belongs to the particles it ran on: after a new
    upload the backward (and a reset) report that no window has run."""
    from sph_seepage_flows_amd import SFSystem

    sc = scenes.mixed(collider=False)
    g = SFSystem.from_scene(sc)
    _abi.check(_abi.lib().sf_diff_forward(g.handle, DT, 2, 1))
    g.upload_particles(scenes.mixed(collider=False, n_water=1000, n_sand=800))
    with pytest.raises(SFError) as e:
        _abi.check(_abi.lib().sf_diff_backward(g.handle, None, None, None, None, None, None, 0))
    assert e.value.code == _abi.SF_E_STATE
    with pytest.raises(SFError) as e:
        _abi.check(_abi.lib().sf_diff_reset(g.handle, None, 0))
    assert e.value.code == _abi.SF_E_STATE
