"""GPU tests of the batched-scene handle, graph replay and the launch-shape
autotuner.  A batched scene runs the same per-particle arithmetic in the same
neighbour order as a single-scene handle of that scene, so every comparison
here is bit-exact."""
import numpy as np
import pytest

from sph_seepage_flows_amd import BatchSystem, SFError, SFSystem, _abi, scenes

pytestmark = pytest.mark.gpu

DT = 1e-4
FIELDS = ("pos", "vel", "density", "pressure", "voidage", "saturation", "avg_drag_force", "col", "acc")


def small_ensemble(n_scenes=5, seed=3):
    return scenes.ensemble(seed=seed, n_scenes=n_scenes, n_min=2_000, n_max=6_000)


def test_batch_matches_single_scene_handles():
    E = small_ensemble()
    print("[batch] counts", [e.n for e in E], flush=True)
    b = BatchSystem.from_scenes(E)
    off = b.offsets()
    assert list(np.diff(off)) == [e.n for e in E]
    b.step(DT, 12)
    for q, sc in enumerate(E):
        s = SFSystem.from_scene(sc)
        s.step(DT, 12)
        for name in FIELDS:
            np.testing.assert_array_equal(b.get_scene(name, q), s.get(name), err_msg=f"scene {q} {name}")
        s.close()


def test_batch_neighbour_structures_per_scene():
    E = small_ensemble(3, seed=5)
    b = BatchSystem.from_scenes(E)
    b.step(DT, 3)
    b.BuildGNSearcher()
    n, nb, cells = b.sizes()
    g = E[0].params()[1].grid_size
    c1 = g.x * g.y * g.z
    assert cells == 3 * c1
    cs = b.get("cell_start")
    keys = b.get("sorted_keys")
    off = b.offsets()
    for q, sc in enumerate(E):
        s = SFSystem.from_scene(sc)
        s.step(DT, 3)
        s.BuildGNSearcher()
        k1 = s.get("sorted_keys")
        np.testing.assert_array_equal(keys[off[q]:off[q + 1]] - np.uint32(q * c1), k1)
        np.testing.assert_array_equal(cs[q * c1:(q + 1) * c1 + 1] - np.uint32(off[q]), s.get("cell_start"))
        np.testing.assert_array_equal(b.get("nbr_count")[off[q]:off[q + 1]], s.get("nbr_count"))
        s.close()


def test_graph_replay_matches_eager_steps():
    sc = scenes.small_block()
    a = SFSystem.from_scene(sc)
    g = SFSystem.from_scene(sc)
    a.step(DT, 7)
    g.step_graph(DT, 7)  # 3 graph replays of two steps + 1 eager step
    for name in ("pos", "vel", "density", "col"):
        np.testing.assert_array_equal(a.get(name), g.get(name))
    a.step(DT, 6)
    g.step_graph(DT, 6)  # parity changed: the graph is captured again
    np.testing.assert_array_equal(a.get("pos"), g.get("pos"))
    # eager work between replays, then replays again
    a.step(DT, 1)
    g.step(DT, 1)
    a.step(DT, 4)
    g.step_graph(DT, 4)
    np.testing.assert_array_equal(a.get("pos"), g.get("pos"))


def test_batch_graph_and_autotune_do_not_change_results():
    E = small_ensemble(4, seed=9)
    a = BatchSystem.from_scenes(E)
    t = BatchSystem.from_scenes(E)
    res = t.autotune(reps=3)
    print("[batch] tuned", res, flush=True)
    assert set(res) == set(SFSystem.NEIGHBOUR_PASSES)
    for nb, cap, best, dflt in res.values():
        assert nb in (64, 128, 256) and cap in (8, 16) and 0 < best <= dflt * 1.0001
    a.step(DT, 9)
    t.step_graph(DT, 9)
    for name in FIELDS:
        np.testing.assert_array_equal(a.get(name), t.get(name), err_msg=name)


@pytest.mark.parametrize("nb,cap", [(64, 8), (128, 16), (256, 8)])
def test_launch_shapes_give_identical_results(nb, cap):
    sc = scenes.mixed()
    a = SFSystem.from_scene(sc)
    b = SFSystem.from_scene(sc)
    for name in SFSystem.NEIGHBOUR_PASSES:
        b.set_launch_config(name, nb, cap)
        assert b.launch_config(name) == (nb, cap)
    a.step(DT, 4)
    b.step(DT, 4)
    for name in FIELDS:
        np.testing.assert_array_equal(a.get(name), b.get(name), err_msg=name)


def test_interface_lanes_change_only_rounding():
    """1, 2 and 4 lanes per particle in the interface passes regroup their
    sums: after several steps of the mixed scene the states agree to FP32
    rounding (relative to each field's magnitude), not bit for bit."""
    sc = scenes.mixed()
    ref = SFSystem.from_scene(sc)
    ref.step(DT, 6)
    for lanes in (2, 4):
        s = SFSystem.from_scene(sc)
        s.set_interface_lanes(lanes)
        s.step(DT, 6)
        for name in ("pos", "vel", "voidage", "saturation", "avg_drag_force"):
            a, b = ref.get(name), s.get(name)
            scale = max(float(np.abs(a).max()), 1e-30)
            err = float(np.abs(a - b).max()) / scale
            print(f"[lanes {lanes}] {name} {err:.2e}", flush=True)
            assert err < 2e-5, (lanes, name, err)
    with pytest.raises(SFError):
        ref.set_interface_lanes(3)


def test_batch_matches_single_scene_handles_with_interface_lanes():
    E = small_ensemble(3, seed=17)
    b = BatchSystem.from_scenes(E)
    b.set_interface_lanes(4)
    b.step(DT, 8)
    for q, sc in enumerate(E):
        s = SFSystem.from_scene(sc)
        s.set_interface_lanes(4)
        s.step(DT, 8)
        for name in FIELDS:
            np.testing.assert_array_equal(b.get_scene(name, q), s.get(name), err_msg=f"scene {q} {name}")
        s.close()


def test_scene_params_change_only_their_scene():
    E = small_ensemble(3, seed=11)
    a = BatchSystem.from_scenes(E)
    b = BatchSystem.from_scenes(E)
    sp = [BatchSystem.scene_params_of(e.physics) for e in E]
    sp[1].dem_tan_friction_angle = 0.05
    sp[1].sf_cd = 50.0
    b.set_scene_params(sp)
    a.step(DT, 30)
    b.step(DT, 30)
    np.testing.assert_array_equal(a.get_scene("pos", 0), b.get_scene("pos", 0))
    np.testing.assert_array_equal(a.get_scene("pos", 2), b.get_scene("pos", 2))
    assert not np.array_equal(a.get_scene("pos", 1), b.get_scene("pos", 1))


def test_batch_errors():
    E = small_ensemble(2, seed=13)
    p, bp, o = E[0].params()
    b = BatchSystem(p, bp, o, [E[0].n, E[1].n], [BatchSystem.scene_params_of(e.physics) for e in E])
    with pytest.raises(SFError) as e:
        b.upload_particles(E[0])  # not the sum of the scene counts
    assert e.value.code == _abi.SF_E_INVALID
    b2 = BatchSystem.from_scenes(E)
    with pytest.raises(SFError) as e:
        b2.UpdateSystem(1 / 30)
    assert e.value.code == _abi.SF_E_STATE
    bad = [BatchSystem.scene_params_of(x.physics) for x in E]
    bad[0].sf_cd = float("nan")
    with pytest.raises(SFError):
        b2.set_scene_params(bad)
    b2.step(DT, 2)  # the handle still works with its previous table
    with pytest.raises(SFError):
        b2.set_launch_config("density", 96, 16)
