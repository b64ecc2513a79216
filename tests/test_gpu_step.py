"""GPU single-step parity against the FP64 oracle.

Error metric per quantity: max_i |gpu_i - ref_i| / max_i |ref_i| over the
particles the quantity is defined on (vectors in the Euclidean norm).  The GPU
computes in FP32; the oracle restates the reference's kernels in FP64 on the
same FP32 inputs and the same (bit-exact) neighbour structures.
"""
import numpy as np
import pytest

from sph_seepage_flows_amd import SFSystem, scenes

from _compare import norm_err, point_err, report

pytestmark = pytest.mark.gpu

# quantity -> (phase mask: None all, 0 water, 1 sand; tolerance)
TOL = {
    "density": (0, 1e-5),
    "pressure": (0, 1e-5),
    "voidage": (None, 1e-5),
    "saturation": (1, 1e-5),
    "avg_flow_vel": (None, 1e-5),
    "adhesion_force": (0, 1e-5),
    # F / vS and df / vW: at a phase fringe the denominator holds only
    # kernel-edge terms c (h^2 - r^2)^3 with r ~ h, whose FP32 relative error
    # is ~ 6e-8 * 3 h^2 / (h^2 - r^2); the ratio inherits it.
    "avg_adhesion_force": (0, 5e-5),
    "avg_drag_force": (1, 5e-5),
    "acc_water": (0, 1e-5),
    "acc_sand": (1, 1e-5),
    "max_saturation": (1, 1e-5),
    "col": (None, 1e-5),
    "vel": (None, 1e-5),
    "pos": (None, 1e-6),
}


def _compare(g, o, tag, skip=(), o32=None):
    """Normwise errors against the FP64 oracle `o`, each within TOL; then,
    per particle (each particle's error relative to its own value, values
    below 1e-3 of the largest measured against that floor), the GPU's worst
    particle against the worst particle of the same formulas evaluated in
    plain FP32 (`o32`, the oracle at 32 bits, same inputs): within 1e-5 or
    4x the FP32 evaluation's, so no particle is worse than FP32 arithmetic
    makes it (near-cancelling sums such as p ~ rho - rho0 or a grain's net
    force at rest are ill-conditioned in any FP32 code)."""
    np.testing.assert_array_equal(g.get("sort_perm"), o.get("sort_perm"))
    np.testing.assert_array_equal(g.get("cell_start"), o.get("cell_start"))
    label = o.get("label")
    errs = {}
    for q, (ph, tol) in TOL.items():
        if q in skip:
            continue
        name = "acc" if q.startswith("acc") else q
        mask = None if ph is None else (label == ph)
        errs[q] = norm_err(g.get(name), o.get(name), mask)
    print(report(tag, errs), flush=True)
    pts = {}
    for q, (ph, tol) in TOL.items():
        if q in skip:
            continue
        name = "acc" if q.startswith("acc") else q
        pts[q] = point_err(g.get(name), o.get(name), None if ph is None else (label == ph))
    print(report(tag + " per particle", pts), flush=True)
    if o32 is not None:
        p32 = {}
        for q, (ph, tol) in TOL.items():
            if q in skip:
                continue
            name = "acc" if q.startswith("acc") else q
            p32[q] = point_err(o32.get(name), o.get(name), None if ph is None else (label == ph))
        print(report(tag + " per particle, FP32 oracle", p32), flush=True)
        bad = {q: (pts[q], p32[q]) for q in pts if not pts[q] <= max(1e-5, 4.0 * p32[q])}
        assert not bad, f"{tag}: per-particle error above 4x the FP32 evaluation's {bad}"
    bad = {q: e for q, e in errs.items() if not e <= TOL[q][1]}
    assert not bad, f"{tag}: above tolerance {bad}"
    return errs


@pytest.mark.parametrize("sf_type", [0, 1])
def test_single_step_mixed(sf_type, oracle_mod):
    sc = scenes.mixed(sf_type=sf_type)
    print(f"[mixed sf_type={sf_type}] n={sc.n} nb={sc.nb}", flush=True)
    g = SFSystem.from_scene(sc)
    o = oracle_mod.OracleSystem(sc, 64)
    o32 = oracle_mod.OracleSystem(sc, 32)
    g.step(1e-4)
    o.step(1e-4)
    o32.step(1e-4)
    errs = _compare(g, o, f"mixed sf_type={sf_type}", o32=o32)
    # the scene must actually exercise the coupling terms
    for f in ("saturation", "avg_drag_force", "avg_adhesion_force", "adhesion_force"):
        assert np.abs(o.get(f)).max() > 0, f


@pytest.mark.parametrize("scene", ["mixed", "small_block"])
def test_single_step_interface_lanes(scene, oracle_mod):
    """The avg-flow and voidage/adhesion passes with 4 lanes per particle
    (sf_set_interface_lanes) meet the same single-step tolerances."""
    sc = scenes.mixed() if scene == "mixed" else scenes.small_block()
    g = SFSystem.from_scene(sc)
    g.set_interface_lanes(4)
    o = oracle_mod.OracleSystem(sc, 64)
    g.step(1e-4)
    o.step(1e-4)
    _compare(g, o, f"{scene} lanes=4")


def test_single_step_small_block(oracle_mod):
    sc = scenes.small_block()
    g = SFSystem.from_scene(sc)
    o = oracle_mod.OracleSystem(sc, 64)
    o32 = oracle_mod.OracleSystem(sc, 32)
    g.step(1e-4)
    o.step(1e-4)
    o32.step(1e-4)
    _compare(g, o, "small_block", o32=o32)


def test_single_step_clamped_edge_cells(oracle_mod):
    """A 0.06 m wide seepage column with h = 0.0124 m: the grid holds 4 whole
    cells across, so the particles beyond 4 h are hashed into the last
    column of cells (the hash clamps), and their neighbours in it lie beyond
    the cell's nominal extent.  avgDrag is left out: in this state every
    grain's water lies at r > 0.97 h (the water block starts one spacing
    above the sand), where the kernel's h^2 - r^2 keeps few FP32 digits and
    the ratio df / vW inherits them; the grains' accelerations, which carry
    df, are checked."""
    sc = scenes.column(n_sand=1500, n_water=500, width=0.06, height=0.1)
    g = SFSystem.from_scene(sc)
    o = oracle_mod.OracleSystem(sc, 64)
    o32 = oracle_mod.OracleSystem(sc, 32)
    g.step(1e-4)
    o.step(1e-4)
    o32.step(1e-4)
    pos = sc.pos
    assert (pos[:, 0] > 4 * sc.physics.h).any() and (pos[:, 2] > 4 * sc.physics.h).any()
    _compare(g, o, "column_clamped", skip=("avg_drag_force",), o32=o32)


def _lattice_scene(name, blocks, box=0.5):
    """Scene of lattice blocks [(origin, spacing, counts, label)] with the
    mixed scene's physics (r = 0.01, h = 0.04) in a cubic box."""
    from sph_seepage_flows_amd import params as P
    ph = P.Physics(hi=(box, box, box))
    parts = []
    for q, (origin, spacing, counts, label) in enumerate(blocks):
        n = counts[0] * counts[1] * counts[2]
        pos = scenes.lattice(origin, spacing, counts, n, jitter=0.1 * spacing, seed=21, stream=q)
        parts.append((pos, label))
    return scenes._assemble(name, ph, parts)


def test_single_step_dense_clump(oracle_mod):
    """Water and sand packed at h/4 (64 particles per cell): a wave's union of
    three cells in a neighbour column holds ~190 candidates, more than the
    density pass's 128-slot LDS stage, so the staged walk runs in chunks."""
    sc = _lattice_scene("dense_clump", [((0.2, 0.2, 0.2), 0.01, (10, 10, 10), 0),
                                        ((0.3, 0.2, 0.2), 0.01, (6, 10, 10), 1)])
    g = SFSystem.from_scene(sc)
    o = oracle_mod.OracleSystem(sc, 64)
    o32 = oracle_mod.OracleSystem(sc, 32)
    g.step(1e-4)
    o.step(1e-4)
    o32.step(1e-4)
    counts = np.diff(g.get("cell_start").astype(np.int64))
    assert counts.max() >= 43, counts.max()  # 3 cells of a column exceed 128
    _compare(g, o, "dense_clump", o32=o32)


@pytest.mark.parametrize("nw,ns", [(1, 0), (0, 1), (63, 2), (64, 0), (0, 65), (255, 3), (257, 256), (2, 511)])
def test_single_step_phase_counts(nw, ns, oracle_mod):
    """Phase-balanced launches at edge counts: one phase empty, counts around
    the 64-lane wave and 256-thread block, one phase of a single particle."""
    blocks = []
    if nw:
        blocks.append(((0.1, 0.1, 0.1), 0.02, (nw, 1, 1) if nw < 8 else (8, 8, (nw + 63) // 64), 0))
    if ns:
        blocks.append(((0.12, 0.11, 0.1), 0.02, (ns, 1, 1) if ns < 8 else (8, 8, (ns + 63) // 64), 1))
    sc = _lattice_scene(f"phases_{nw}_{ns}", blocks)
    n_water_built = int((sc.label == 0).sum())  # the blocks hold whole lattice layers; keep nw and ns
    keep = np.concatenate([np.arange(nw), n_water_built + np.arange(ns)])
    sc = scenes.Scene(sc.name, sc.physics, sc.pos[keep].copy(), sc.label[keep].copy(), sc.mass[keep].copy(),
                      sc.radius[keep].copy(), sc.col[keep].copy(), bpos=sc.bpos, blabel=sc.blabel)
    assert (int((sc.label == 0).sum()), int((sc.label == 1).sum())) == (nw, ns)
    g = SFSystem.from_scene(sc)
    o = oracle_mod.OracleSystem(sc, 64)
    o32 = oracle_mod.OracleSystem(sc, 32)
    g.step(1e-4)
    o.step(1e-4)
    o32.step(1e-4)
    np.testing.assert_array_equal(g.get("sort_perm"), o.get("sort_perm"))
    # The lattices sit at rest spacing, so p ~ (rho - rho0) amplifies the
    # FP32 rounding of rho; each quantity is held to 1e-5 or twice the error
    # of the same formulas in plain FP32 (the FP32 oracle), whichever is larger.
    for q in ("density", "pressure", "vel", "pos", "acc"):
        ref = o.get(q)
        err = norm_err(g.get(q), ref, None)
        floor = 2.0 * norm_err(o32.get(q), ref, None)
        print(f"[phases {nw}+{ns}] {q} {err:.2e} (FP32 oracle {floor / 2:.2e})", flush=True)
        assert err <= max(1e-5, floor), (nw, ns, q, err, floor)


def test_update_system_cfl(oracle_mod):
    """UpdateSystem(renderInterval) with the WCSPH CFL sub-step count
    (ComputeSubTimeStepsByCFL, cuda_wcsph_sf_solver.cu:35-58)."""
    sc = scenes.small_block()
    g = SFSystem.from_scene(sc)
    o = oracle_mod.OracleSystem(sc, 64)
    for k in range(3):
        ms = g.UpdateSystem(1.0 / 30.0)
        o.update_system(1.0 / 30.0)
        assert g.GetNumOfSubTimeSteps() == o.num_sub_steps(), k
        assert ms > 0.0
    print("CFL sub-steps", g.GetNumOfSubTimeSteps(), flush=True)
    assert norm_err(g.get("pos"), o.get("pos")) < 1e-6


def test_sph_solver_linear_eos(oracle_mod):
    """SPH_SOLVER: linear EOS with negatives kept, advect with params.dt
    (CudaSphSFSolver::UpdateSolver, cuda_sph_sf_solver.cpp:14-152)."""
    sc = scenes.mixed()
    sc.physics.solver_type = 0
    g = SFSystem.from_scene(sc)
    o = oracle_mod.OracleSystem(sc, 64)
    o32 = oracle_mod.OracleSystem(sc, 32)
    g.UpdateSystem(1e-3)
    o.update_system(1e-3)
    o32.update_system(1e-3)
    assert g.GetNumOfSubTimeSteps() == o.num_sub_steps() == 10
    _compare(g, o, "sph linear eos", o32=o32)
