"""sf_update_frame against the reference driver's frame loop run from the host
(main.cpp:235-259, sf_update_system per sub-step): bit-identical particle
state, sub-step counts, nsub and emitter state, for the adaptive and the
fixed loop, the WCSPH and the SPH solver, with and without the inflow
emitter, and with a CFL dt bound by the accelerations (the case where the
host cannot know the sub-step count in advance)."""
import dataclasses

import numpy as np
import pytest

from sph_seepage_flows_amd import SF_SPH_SOLVER, SFSystem, scenes

pytestmark = pytest.mark.gpu
RI = 1.0 / 30.0
FIELDS = ("pos", "vel", "col", "label", "mass", "density", "saturation", "max_saturation", "acc")


def host_frame(g, ri, adaptive):
    """main.cpp:235-259 in FP32, sub-steps run."""
    if adaptive:
        rem, k = np.float32(ri), 0
        while rem > np.float32(1e-6):
            g.UpdateSystem(float(rem))
            rem = np.float32(rem - rem / np.float32(g.GetNumOfSubTimeSteps()))
            k += 1
        return k
    k = g.GetNumOfSubTimeSteps()
    for _ in range(k):
        g.UpdateSystem(ri)
    return k


def assert_same(a, b, tag):
    assert a.sizes() == b.sizes(), tag
    assert a.GetNumOfSubTimeSteps() == b.GetNumOfSubTimeSteps(), tag
    assert a.emitter_state() == b.emitter_state(), tag
    for f in FIELDS:
        np.testing.assert_array_equal(a.get(f), b.get(f), err_msg=f"{tag} {f}")


def run_pair(sc, frames, adaptive=True, emitter=None, max_num=None):
    a = SFSystem.from_scene(sc, max_num=max_num)
    b = SFSystem.from_scene(sc, max_num=max_num)
    if emitter:
        a.set_emitter(**emitter)
        b.set_emitter(**emitter)
    counts = []
    for fr in range(frames):
        ka = host_frame(a, RI, adaptive)
        kb, ms = b.update_frame(RI, adaptive)
        print(f"[frame] {sc.name} frame {fr + 1}: host {ka} sub-steps, device {kb} ({ms:.1f} ms), "
              f"n = {b.sizes()[0]}", flush=True)
        assert ka == kb
        assert_same(a, b, f"frame {fr + 1}")
        counts.append(kb)
    return counts


def test_adaptive_frames_small_block():
    counts = run_pair(scenes.small_block(), 2)
    assert counts == [333, 333]


def test_adaptive_frames_with_emitter():
    sc = scenes.small_block()
    em = scenes.main_emitter(sc.physics)
    run_pair(sc, 2, emitter=em, max_num=sc.n + 20000)


def test_adaptive_frames_acceleration_bound_cfl():
    """Slow sound (c0 = 2) and a small CFL scale: dt = 0.03 dtf follows the
    accelerations, below the fixed bound 0.03 * 0.4 h / c0, so the frame runs
    more sub-steps than the host's certain count and needs several reads."""
    sc = scenes.mixed(collider=False)
    sc.physics = dataclasses.replace(sc.physics, c0=2.0, time_step_limit_scale=0.03, dt=1.0)
    counts = run_pair(sc, 2)
    h = np.float32(sc.physics.h)
    dt_max = np.float32(0.03) * (np.float32(0.4) * h / np.float32(2.0))
    assert counts[0] > int(np.float32(RI) / dt_max)


def test_sph_solver_frames_with_emitter():
    sc = scenes.mixed(collider=False)
    sc.physics = dataclasses.replace(sc.physics, solver_type=SF_SPH_SOLVER, dt=2e-4)
    em = dict(scenes.main_emitter(sc.physics), vel=(0.0, 0.0, -0.5), radius=0.05)
    counts = run_pair(sc, 2, emitter=em, max_num=sc.n + 20000)
    assert counts == [166, 166]


def test_fixed_loop_frames():
    """main.cpp's non-adaptive branch: GetNumOfSubTimeSteps() steps of the
    whole render interval (1 in the first frame, the CFL count afterwards)."""
    sc = scenes.small_block()
    counts = run_pair(sc, 3, adaptive=False)
    assert counts[0] == 1 and counts[1] > 1
