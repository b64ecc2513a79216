"""Seepage embankment (2.5M-grain trapezoidal dam + 1.5M reservoir water,
4M particles, 1.17M cells) at full size: one step against the FP64 oracle
with the single-step tolerances, and neighbour structures rebuilt by the
oracle from the GPU's own positions after 20 steps, bit for bit."""
import numpy as np
import pytest

from sph_seepage_flows_amd import SFSystem, scenes
from sph_seepage_flows_amd.scenes import Scene

from _compare import norm_err, report
from test_gpu_step import TOL

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dam():
    return scenes.dam()


def test_dam_single_step_vs_fp64_oracle(dam, oracle_mod):
    """Tolerance per quantity: the single-step tolerance, or twice the error
    of the FP32 oracle (the same formulas evaluated in plain FP32 on the CPU)
    where that is larger.  In this geometry the water-sand adhesion
    A(r) ~ (-4r^2/h + 6r - 2h)^(1/4) (cuda_sph_kernel.cuh:189-203) is evaluated
    at many pair distances just above h/2, where its argument is a cancellation
    of O(h) terms and the quarter power amplifies the FP32 rounding (FP32 oracle:
    adhesion 3.9e-5, average adhesion 2.0e-3, sand acceleration 1.08e-5)."""
    print(f"[dam] n={dam.n} nb={dam.nb}", flush=True)
    g = SFSystem.from_scene(dam)
    o = oracle_mod.OracleSystem(dam, 64)
    o32 = oracle_mod.OracleSystem(dam, 32)
    g.step(1e-4)
    o.step(1e-4)
    o32.step(1e-4)
    np.testing.assert_array_equal(g.get("sort_perm"), o.get("sort_perm"))
    np.testing.assert_array_equal(g.get("cell_start"), o.get("cell_start"))
    label = o.get("label")
    errs, e32 = {}, {}
    for q, (ph, tol) in TOL.items():
        name = "acc" if q.startswith("acc") else q
        m = None if ph is None else (label == ph)
        ref = o.get(name)
        errs[q] = norm_err(g.get(name), ref, m)
        e32[q] = norm_err(o32.get(name), ref, m)
    print(report("dam step 1 (GPU FP32 vs FP64)", errs), flush=True)
    print(report("dam step 1 (CPU FP32 oracle vs FP64)", e32), flush=True)
    bad = {q: (e, e32[q]) for q, e in errs.items() if not e <= max(TOL[q][1], 2 * e32[q])}
    assert not bad, f"dam: above tolerance {bad}"
    for q in ("density", "pressure", "saturation", "voidage", "acc_water"):
        assert errs[q] <= TOL[q][1]
    # the initial saturation history survives where the step's saturation is lower
    ms = g.get("max_saturation")[label == 1]
    assert ms.max() > 0.09 and ms.min() >= 0


def test_dam_neighbour_structures_after_20_steps(dam, oracle_mod):
    g = SFSystem.from_scene(dam)
    g.step(1e-4, 20)
    pos = g.get("pos")
    assert np.all(np.isfinite(pos))
    lab = g.get("label")
    assert int((lab == 1).sum()) == 2_500_000 and int((lab == 0).sum()) == 1_500_000
    np.testing.assert_array_equal(np.sort(g.get("mass")), np.sort(dam.mass))  # mass conserved exactly
    sc = Scene(dam.name, dam.physics, pos, lab, g.get("mass"), g.get("radius"), g.get("col"), g.get("vel"),
               None, None, None, None)
    o = oracle_mod.OracleSystem(sc, 32)
    o.build_neighbors()
    g.BuildGNSearcher()
    for f in ("particle2cell", "sort_perm", "sorted_keys", "cell_start", "nbr_count", "nbr_candidates"):
        np.testing.assert_array_equal(g.get(f), o.get(f), err_msg=f)
