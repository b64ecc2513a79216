"""The incremental stable sort (default) against the LSD radix sort: the same
permutation, sorted keys and cellStart at every step, and therefore the same
trajectory bit for bit.  (The 1M neighbour-structure test in
test_gpu_multistep.py checks the incremental path against the oracle's
std::stable_sort as well.)"""
import numpy as np
import pytest

from sph_seepage_flows_amd import BatchSystem, SFSystem, scenes

pytestmark = pytest.mark.gpu

DT = 1e-4


def pair(make):
    a = make()
    b = make()
    a.set_sort_strategy(False)  # radix every step
    b.set_sort_strategy(True)
    return a, b


@pytest.mark.parametrize("name", ["mixed", "small_block"])
def test_incremental_sort_equals_radix_sort(name):
    sc = getattr(scenes, name)()
    a, b = pair(lambda: SFSystem.from_scene(sc))
    for step in range(40):
        a.step(DT)
        b.step(DT)
        if step % 8 == 0 or step == 39:
            for f in ("sort_perm", "sorted_keys", "cell_start"):
                np.testing.assert_array_equal(a.get(f), b.get(f), err_msg=f"step {step} {f}")
    for f in ("pos", "vel", "density", "saturation", "col"):
        np.testing.assert_array_equal(a.get(f), b.get(f), err_msg=f)


def test_incremental_sort_fast_motion():
    # large random velocities: many particles change cell every step
    sc = scenes.mixed(vel_sigma=3.0)
    a, b = pair(lambda: SFSystem.from_scene(sc))
    for step in range(12):
        a.step(DT)
        b.step(DT)
        np.testing.assert_array_equal(a.get("sort_perm"), b.get("sort_perm"), err_msg=f"step {step}")
    np.testing.assert_array_equal(a.get("pos"), b.get("pos"))


def test_incremental_sort_batched_and_graph():
    E = scenes.ensemble(seed=21, n_scenes=4, n_min=2000, n_max=5000)
    a, b = pair(lambda: BatchSystem.from_scenes(E))
    a.step(DT, 10)
    b.step_graph(DT, 10)
    for f in ("sort_perm", "sorted_keys", "cell_start", "pos", "vel", "density"):
        np.testing.assert_array_equal(a.get(f), b.get(f), err_msg=f)


def test_incremental_sort_after_reupload():
    sc = scenes.small_block()
    a, b = pair(lambda: SFSystem.from_scene(sc))
    a.step(DT, 5)
    b.step(DT, 5)
    b.upload_particles(sc)  # back to the unsorted upload order: the next build must use the radix sort
    a.upload_particles(sc)
    a.step(DT, 5)
    b.step(DT, 5)
    np.testing.assert_array_equal(a.get("pos"), b.get("pos"))


def test_prefix_scan_branches_are_identical():
    """The slot prefixes of the incremental sort and of the water prefix
    have two branches: each prefix block sums the earlier tiles' counts
    itself (up to 4096 tiles of 2048 slots, every scene here) or a one-block
    scan launch runs in between (above; threshold 0 forces it).  Both must
    give the same structures and trajectories, bit for bit."""
    from sph_seepage_flows_amd import _abi
    lib = _abi.lib()
    makers = {
        "mixed": lambda: SFSystem.from_scene(scenes.mixed(vel_sigma=1.0)),
        "batched": lambda: BatchSystem.from_scenes(scenes.ensemble(seed=21, n_scenes=3, n_min=2000, n_max=5000)),
    }
    for name, make in makers.items():
        a = make()
        a.step(DT, 6)
        prev = lib.sf_set_prefix_scan_threshold(0)
        try:
            assert prev == 4096
            b = make()
            b.step(DT, 6)
        finally:
            lib.sf_set_prefix_scan_threshold(-1)
        for f in ("sort_perm", "sorted_keys", "cell_start", "pos", "vel", "density", "saturation"):
            np.testing.assert_array_equal(a.get(f), b.get(f), err_msg=f"{name}: {f}")
