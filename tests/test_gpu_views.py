"""The reference-layout device arrays (sf_particle_view, the getters of
CudaSFParticles, cuda_sf_particles.cuh:170-190, and GetCellStart /
GetGridIdxArray, cuda_neighbor_searcher.cuh:44-50), sf_update_solver
(CudaWCSphSFSolver::UpdateSolver on a BuildGNSearcher's structures) and the
device-array bgeo export (ExportBgeoFileCUDA, kiri_utils.h:364-371)."""
import numpy as np
import pytest

from sph_seepage_flows_amd import SFSystem, scenes
from sph_seepage_flows_amd.system import export_bgeo_device

from _hip import d2h, h2d

pytestmark = pytest.mark.gpu

F3 = ["pos", "vel", "acc", "col", "avg_flow_vel", "avg_drag_force", "adhesion_force", "avg_adhesion_force",
      "cda0asat"]
F1 = ["mass", "radius", "density", "pressure", "voidage", "saturation", "max_saturation"]


def read_view(g, name):
    n, nb, C = g.sizes()
    a = g.view(name)
    if name in F3:
        return d2h(a, (n, 3), np.float32)
    if name in F1:
        return d2h(a, n, np.float32)
    if name == "amcamcp":
        return d2h(a, (n, 2), np.float32)
    if name in ("label", "sorted_keys"):
        return d2h(a, n, np.uint64)
    if name in ("cell_start", "b_cell_start"):
        return d2h(a, C + 1, np.uint64)
    raise KeyError(name)


@pytest.mark.parametrize("sf_type", [0, 1])
def test_views_follow_every_step(sf_type):
    sc = scenes.mixed(sf_type=sf_type)
    g = SFSystem.from_scene(sc)
    names = F3 + F1 + ["label", "amcamcp", "sorted_keys", "cell_start", "b_cell_start"]
    g.BuildGNSearcher()
    addr = {k: g.view(k) for k in names}
    for step in range(3):
        g.step(1e-4)
        for k in names:
            assert g.view(k) == addr[k], f"{k}: the view moved"
            v = read_view(g, k)
            ref = g.get(k)
            if k in ("sorted_keys", "cell_start", "b_cell_start"):
                ref = ref.astype(np.uint64)
            np.testing.assert_array_equal(v, ref, err_msg=f"step {step}: {k}")
        print(f"step {step}: {len(names)} views equal the state", flush=True)


def test_view_writes_enter_the_next_step():
    """Velocities, colours, maxSaturation written into the views between steps
    give the same step as sf_set_field with the same values; positions and
    masses written into the views give the same step as a fresh handle
    uploaded with the changed state."""
    sc = scenes.mixed()
    a = SFSystem.from_scene(sc)
    b = SFSystem.from_scene(sc)
    a.step(1e-4, 2)
    b.step(1e-4, 2)
    rng = np.random.default_rng(5)
    n = a.sizes()[0]
    vel = (a.get("vel") + rng.normal(0, 0.05, (n, 3))).astype(np.float32)
    col = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    ms = rng.uniform(0, 0.4, n).astype(np.float32)
    h2d(a.view("vel"), vel)
    h2d(a.view("col"), col)
    h2d(a.view("max_saturation"), ms)
    b.set_field("vel", vel)
    b.set_field("col", col)
    b.set_field("max_saturation", ms)
    a.step(1e-4, 2)
    b.step(1e-4, 2)
    for k in ("pos", "vel", "col", "max_saturation", "acc", "density"):
        np.testing.assert_array_equal(a.get(k), b.get(k), err_msg=k)

    # positions and labels through the views
    pos = a.get("pos")
    pos[:, 1] += np.float32(0.003)
    lab = a.get("label")
    h2d(a.view("pos"), pos)
    st = {k: a.get(k) for k in ("vel", "col", "mass", "radius", "max_saturation", "cda0asat", "amcamcp")}
    h2d(a.view("label"), lab)
    fresh = scenes.Scene("fresh", sc.physics, pos, lab, st["mass"], st["radius"], st["col"], st["vel"],
                         st["cda0asat"], st["amcamcp"], sc.bpos, sc.blabel, {}, st["max_saturation"])
    c = SFSystem.from_scene(fresh)
    a.step(1e-4)
    c.step(1e-4)
    for k in ("pos", "vel", "col", "label", "density", "acc"):
        np.testing.assert_array_equal(a.get(k), c.get(k), err_msg=k)


def test_update_solver_after_build_equals_update_system():
    sc = scenes.small_block()
    a = SFSystem.from_scene(sc)
    b = SFSystem.from_scene(sc)
    with pytest.raises(Exception):
        b.UpdateSolver(1.0 / 30.0)  # no build yet
    for q in range(4):
        a.UpdateSystem(1.0 / 30.0)
        b.BuildGNSearcher()
        b.UpdateSolver(1.0 / 30.0)
        assert a.GetNumOfSubTimeSteps() == b.GetNumOfSubTimeSteps()
    for k in ("pos", "vel", "acc", "density", "saturation", "cell_start"):
        np.testing.assert_array_equal(a.get(k), b.get(k), err_msg=k)
    with pytest.raises(Exception):
        b.UpdateSolver(1.0 / 30.0)  # the build was used


def test_export_bgeo_device_equals_state_export(tmp_path):
    sc = scenes.mixed()
    g = SFSystem.from_scene(sc)
    g.step(1e-4, 3)
    n = g.sizes()[0]
    f1 = g.export_bgeo(str(tmp_path / "a"), "0001")
    f2 = export_bgeo_device(str(tmp_path / "b"), "0001", g.view("pos"), g.view("col"), g.view("radius"),
                            g.view("label"), n)
    data = open(f1, "rb").read()
    assert len(data) > 36 * n and data == open(f2, "rb").read()


@pytest.mark.parametrize("how", ["view", "device_buffer"])
def test_position_handout_after_a_step_changes_nothing(how):
    """A step's advect hashes the next build's keys and takes the newcomers'
    ranks; handing the positions out afterwards makes the next build hash and
    rank again, which must start from zero counts.  The run with the handout
    equals the run without it bit for bit (regression: the ranks were counted
    twice, which shifted the permutation)."""
    sc = scenes.small_block()
    a = SFSystem.from_scene(sc)
    b = SFSystem.from_scene(sc)
    a.step(1e-4, 5)
    b.step(1e-4, 5)
    if how == "view":
        b.view("pos")
    else:
        b.device_buffer("pos")
    for _ in range(4):
        a.step(1e-4, 25)
        b.step(1e-4, 25)
        for k in ("sort_perm", "cell_start", "pos", "vel"):
            np.testing.assert_array_equal(a.get(k), b.get(k), err_msg=k)
