"""GPU searcher parity: grid keys, sort permutation, cellStart/End, reordered
state and neighbour counts must equal the oracle's bit for bit
(hash helper_thrust.cuh:56-130, sort + cellStart cuda_neighbor_searcher.cu:36-116)."""
import numpy as np
import pytest

from sph_seepage_flows_amd import SFSystem, scenes
from sph_seepage_flows_amd.scenes import Scene

pytestmark = pytest.mark.gpu

EXACT = ["particle2cell", "sort_perm", "sorted_keys", "cell_start", "nbr_count", "nbr_candidates", "label"]


def _check(scene, oracle_mod, builds=1):
    g = SFSystem.from_scene(scene)
    o = oracle_mod.OracleSystem(scene, 64)
    pos0 = scene.pos.copy()
    for b in range(builds):
        print(f"[{scene.name}] build {b} n={scene.n} nb={scene.nb}", flush=True)
        g.BuildGNSearcher()
        o.build_neighbors()
        for f in EXACT:
            np.testing.assert_array_equal(g.get(f), o.get(f), err_msg=f"{scene.name} build {b}: {f}")
        perm = o.get("sort_perm")
        pos0 = pos0[perm]
        np.testing.assert_array_equal(g.get("pos"), pos0, err_msg="reordered positions")
        np.testing.assert_array_equal(g.get("mass"), o.get("mass").astype(np.float32))
    return g, o


def test_small_block_neighbors(oracle_mod):
    g, o = _check(scenes.small_block(), oracle_mod, builds=2)
    # second build on sorted data keeps the order
    np.testing.assert_array_equal(g.get("sort_perm"), np.arange(g.sizes()[0], dtype=np.uint32))
    cs = g.get("cell_start")
    assert cs[0] == 0 and cs[-1] == g.sizes()[0] and np.all(np.diff(cs.astype(np.int64)) >= 0)


def test_mixed_neighbors_multi(oracle_mod):
    _check(scenes.mixed(sf_type=1), oracle_mod)


def test_boundary_structures(oracle_mod):
    sc = scenes.mixed()
    g = SFSystem.from_scene(sc)
    o = oracle_mod.OracleSystem(sc, 64)
    np.testing.assert_array_equal(g.get("b_cell_start"), o.get("b_cell_start"))
    np.testing.assert_array_equal(g.get("b_pos"), o.get("b_pos").astype(np.float32))
    np.testing.assert_array_equal(g.get("b_label"), o.get("b_label"))
    vg, vo = g.get("b_volume"), o.get("b_volume")
    rel = np.max(np.abs(vg - vo) / np.abs(vo))
    print(f"boundary volume max rel err {rel:.2e}")
    assert rel < 1e-5


def _scene_from(pos, label, ph, name, bpos=None, blabel=None):
    pos = np.ascontiguousarray(pos, np.float32)
    n = len(pos)
    lab = np.ascontiguousarray(label, np.uint64)
    mass = np.where(lab == 0, np.float32(0.0064), np.float32(0.0111)).astype(np.float32)
    rad = np.full(n, 0.01, np.float32)
    col = np.zeros((n, 3), np.float32)
    return Scene(name, ph, pos, lab, mass, rad, col, bpos=bpos, blabel=blabel)


def test_edge_dense_cell_outside_nan_ragged(oracle_mod):
    """One cell holding 3000 particles, particles outside the box and exactly on
    its faces (clamped to the edge cells), NaN positions (cell 0), and a count
    that is not a multiple of the sort tile."""
    from sph_seepage_flows_amd import params as P
    ph = P.Physics(hi=(0.5, 0.5, 0.5))
    rng = np.random.default_rng(3)
    dense = np.float32(0.2) + rng.random((3000, 3), dtype=np.float32) * np.float32(0.0399)
    spread = rng.random((1201, 3), dtype=np.float32) * np.float32(0.5)
    outside = np.array([[-0.1, 0.2, 0.2], [0.6, 0.2, 0.2], [0.2, -1e-7, 0.2], [0.5, 0.5, 0.5], [0.0, 0.0, 0.0],
                        [0.2, 0.2, 7.0], [np.nan, 0.2, 0.2], [0.2, np.inf, 0.2], [0.2, 0.2, -np.inf]], np.float32)
    pos = np.concatenate([spread, dense, outside])
    label = (rng.random(len(pos)) < 0.4).astype(np.uint64)
    _check(_scene_from(pos, label, ph, "edge"), oracle_mod, builds=2)


def test_single_and_empty(oracle_mod):
    from sph_seepage_flows_amd import params as P
    ph = P.Physics(hi=(0.5, 0.5, 0.5))
    _check(_scene_from(np.array([[0.11, 0.22, 0.33]]), np.array([1]), ph, "single"), oracle_mod)
    sc = _scene_from(np.zeros((0, 3)), np.zeros(0), ph, "empty")
    g = SFSystem.from_scene(sc, max_num=16)
    g.BuildGNSearcher()
    assert np.all(g.get("cell_start") == 0)
    g.step(1e-4)
    assert g.sizes()[0] == 0


def test_three_pass_sort_large_grid(oracle_mod):
    """1 m scene grid (108 x 54 x 54 = 314,928 cells -> 19-bit keys, 3 radix
    passes) on a 60k-particle subset, shuffled so the sort has real work."""
    big = scenes.drop_in_1m()
    rng = np.random.default_rng(5)
    idx = rng.permutation(big.n)[:60001]
    sc = Scene("drop_in_subset", big.physics, big.pos[idx].copy(), big.label[idx].copy(), big.mass[idx].copy(),
               big.radius[idx].copy(), big.col[idx].copy(), bpos=big.bpos, blabel=big.blabel)
    _check(sc, oracle_mod, builds=2)
