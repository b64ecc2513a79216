"""One frame of the reference driver's adaptive loop (main.cpp:235-247) on the
GPU against the FP64 oracle's CudaBaseSystem::UpdateSystem loop
(oracle.update_system, cuda_base_system.cu:29-56).  The GPU runs the loop
from the host (sf_update_system per sub-step, recording every sub-step count)
and as one sf_update_frame; the two are identical bit for bit
(test_gpu_frame.py), so the recorded sequence is the device frame's too.

* small block, 1/30 s: the CFL dt is capped by params.dt, so the count
  sequence (333, 332, ..., 1: remaining -= remaining / nsub) is fixed by the
  FP32 remaining-time arithmetic and equal in full; the state after the
  frame is compared with the multi-step statistics' tolerances.
* mixed scene with a slow sound speed and a small CFL scale, where the
  sub-step dt follows max |acc|: the sequences agree while the FP32 and FP64
  accelerations do (the CPU oracle's own FP32 and FP64 runs first differ at
  sub-step 28, by 0.2 % in the count), the first 20 counts are compared
  exactly and the totals within 2 %."""
import dataclasses

import numpy as np
import pytest

from sph_seepage_flows_amd import SFSystem, scenes

from test_gpu_multistep import surface_profile_hist

pytestmark = pytest.mark.gpu
RI = np.float32(1.0 / 30.0)


def gpu_frame(sc):
    g = SFSystem.from_scene(sc)
    rem, seq = RI, []
    while rem > np.float32(1e-6):
        g.UpdateSystem(float(rem))
        k = g.GetNumOfSubTimeSteps()
        seq.append(k)
        rem = np.float32(rem - rem / np.float32(k))
    d = SFSystem.from_scene(sc)
    kd, _ = d.update_frame(float(RI))
    assert kd == len(seq)
    np.testing.assert_array_equal(d.get("pos"), g.get("pos"))
    return g, seq


def oracle_frame(oracle_mod, sc):
    o = oracle_mod.OracleSystem(sc, 64)
    rem, seq = RI, []
    while rem > np.float32(1e-6):
        o.update_system(float(rem))
        k = o.num_sub_steps()
        seq.append(k)
        rem = np.float32(rem - rem / np.float32(k))
    return o, seq


def test_adaptive_frame_small_block_vs_oracle(oracle_mod):
    sc = scenes.small_block()
    g, gs = gpu_frame(sc)
    o, os_ = oracle_frame(oracle_mod, sc)
    print(f"[frame vs oracle] small block: {len(gs)} / {len(os_)} sub-steps", flush=True)
    assert gs == os_ and len(gs) == 333
    tg, pg, hg = surface_profile_hist(g.get("pos"), g.get("label"), g.get("saturation"), 0.5, 0.82)
    to, po, ho = surface_profile_hist(o.get("pos"), o.get("label"), o.get("saturation"), 0.5, 0.82)
    print(f"water top {tg:.5f} vs {to:.5f}; sand profile max diff {np.abs(pg - po).max():.2e}", flush=True)
    assert abs(tg - to) < 1e-4
    assert np.abs(pg - po).max() < 1e-2
    np.testing.assert_array_equal(np.sort(g.get("mass")), np.sort(o.get("mass").astype(np.float32)))


def test_adaptive_frame_cfl_bound_vs_oracle(oracle_mod):
    sc = scenes.mixed(collider=False)
    sc.physics = dataclasses.replace(sc.physics, c0=2.0, time_step_limit_scale=0.03, dt=1.0)
    g, gs = gpu_frame(sc)
    o, os_ = oracle_frame(oracle_mod, sc)
    print(f"[frame vs oracle] CFL-bound mixed: {len(gs)} / {len(os_)} sub-steps; first counts {gs[:6]} / "
          f"{os_[:6]}", flush=True)
    assert gs[:20] == os_[:20]
    assert abs(len(gs) - len(os_)) <= 0.02 * len(os_)
    lab = g.get("label")
    sg = g.get("saturation")[lab == 1].mean()
    so = o.get("saturation")[o.get("label") == 1].mean()
    print(f"mean sand saturation {sg:.5f} / {so:.5f}", flush=True)
    assert abs(sg - so) < 2e-2 * abs(so)
