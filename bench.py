#!/usr/bin/env python3
"""Headline benchmark: coupled SPH-DEM seepage particle-steps/s on one MI355X.

Workload (default): the 1M-particle drop-in scene -- 600k WCSPH water + 400k
DEM sand in a 2 x 1 x 1 m box, jitter seed 1, dt = 1e-4, capillary and
seepage drag on.  One "step" = hash + stable sort + reorder + cellStart +
the five solver passes + advect over all particles, inputs resident in HBM.

Prints ONE JSON line.  `roofline` is for the dominant pass: its algorithmic
HBM bytes per launch (byte model below) divided by its mean duration, timed
with HIP events on the stream the passes run on, inside the timed region
(events around that pass on every EVENT_EVERY-th step only, since each event
between two kernels costs the stream a few microseconds; an untimed profiled
run before it gives the per-pass breakdown `pass_ms_profiled` and picks the
pass).
`cpu_baseline` times the FP32 C++ oracle (a restatement of the reference's
kernels; the reference itself has no CPU path) on the same scene on the host
cores, for a bounded number of steps.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "coupled particle-steps/s @1M particles (%HBM roofline); adjoint/forward time"
PROFILE_STEPS = 10  # untimed steps with an event between every two passes
EVENT_EVERY = 10  # timed steps: events around the dominant pass on every 10th step
HBM_PEAK_GBS = 8000.0


def pass_bytes(n_w: int, n_s: int, nb: int, cells: int, multi: bool, listed=(0, 0)) -> dict:
    """Algorithmic HBM bytes per launch of each pass: every array element the
    pass must touch, read or written once (neighbour re-reads are served
    on-chip and not counted).  Layout: pm/vr/cs/flow/adh/xf/acc float4 (16 B),
    rp/es float2 (8 B), keys/perm/cellStart/lists u32 (4 B), flags u8,
    boundary float4; the water momentum pass's two 16-byte neighbour records
    per particle (every water particle has one mass, as in the benchmark
    scenes).  listed = (water, sand) particles with the other phase in their
    27 cells: the interface passes run over those only, and the density pass
    writes the empty sums' results of the rest that a later pass reads
    (water records, sand voidage / saturation) plus one flag byte each."""
    n = n_w + n_s
    aux = 24 if multi else 0
    lw, ls = listed
    return {
        # stepping loop: the advect computes the next keys (below), the stable
        # sort is the incremental one (change flags and both prefixes, the
        # newcomer lists, the placement: keys, prefixes, permutation, sorted
        # keys; per cell the newcomer counts and starts)
        "sort": n * 44 + 12 * (cells + 1),
        # perm, persistent state read by perm and written, water flag; the
        # water prefix then reads the flags and writes prefix + phase list
        "gather": n * (4 + 2 * (48 + aux) + 1) + n * (1 + 4 + 4),
        # per cell: old start, stayer / newcomer prefixes, water prefix read,
        # cellStart, (start, prefix) table and newcomer counter written; per
        # particle: phase list + position read, position in phase order written
        "cell_start": (cells + 1) * (16 + 16) + n * (4 + 16 + 16),
        # plist, own position (water: pm, sand: pph), water candidates' pph, rho/p (water);
        # sand: position record, velocity read and written in phase order;
        # unlisted water: velocity read, both records written; unlisted
        # sand: es + flag; listed water: velocity read and written, (rho, p);
        # listed: list entry
        "density": n_w * (4 + 16 + 16 + 8 + 1) + n_s * (4 + 16 + 1 + 16 + 32) + (n_w - lw) * (16 + 32)
                   + (n_s - ls) * 8 + lw * (32 + 16) + (lw + ls) * 4 + 8 * (cells + 1) + 4 * (cells + 1) + cells
                   + nb * 16,
        # listed only: list + slot, own position, phase-index position / velocity
        # (/ rho for water), flow, es, water velocity, records and voidage, sand
        # saturation (twice: for P3, and beside the phase-order velocity for P4)
        "avg_flow": lw * (4 + 4 + 16 + 16 + 16 + 16 + 8 + 16 + 8 + 16 + 32 + 4) + ls * (4 + 4 + 16 + 16 + 16 + 16 + 8 + 4 + 4)
                    + 8 * (cells + 1),
        # listed only: list + slot, own position, phase-index position / (rho, voidage, sat);
        # sand voidage + adhesion 0, water adhesion + avgAdhesion (slot and phase index)
        # and the buoyancy record (own (rho, p, voidage) read, position + factor written)
        "voidage_adhesion": lw * (4 + 4 + 16 + 16 + 16 + 48 + 16 + 16) + ls * (4 + 4 + 16 + 16 + 16 + 4 + 16)
                            + 8 * (cells + 1),
        # per grain: plist, position, velocity, es, colour read and written,
        # interface flag, avgDrag (and its record) and acc written; the sand
        # candidates' phase-order positions and velocities touched once; the cell table
        "sand_momentum": n_s * (4 + 16 + 16 + 8 + 16 + 16 + 1 + 16 + 16 + 16) + n_s * 32 + 8 * (cells + 1),
        # both records of every particle once (own and candidates); per water
        # particle plist, rho/p and acc written; the cell and boundary
        # tables, boundary particles
        "water_momentum": n * 32 + n_w * (4 + 8 + 16) + 8 * (cells + 1)
                          + 4 * (cells + 1) + cells + nb * 16,
        # position, velocity, acc read; position, velocity written; sand colour;
        # the next key written and the current sorted key read (change flag)
        "advect": n * (16 * 3 + 16 * 2) + n_s * 16 + n * (4 + 4),
    }


def interface_counts(pos, lab, b) -> tuple[int, int]:
    """(water, sand) particles with a particle of the other phase in their 27
    grid cells: the interface passes' particles (the density pass's test),
    counted on the host from FP32 cell coordinates as the grid hash computes
    them (b: the sf_boundary_params)."""
    import numpy as np

    g = np.array([b.grid_size.x, b.grid_size.y, b.grid_size.z], np.int64)
    lo = np.array([b.lowest_point.x, b.lowest_point.y, b.lowest_point.z], np.float32)
    q = (np.asarray(pos, np.float32) - lo) / np.float32(b.kernel_radius)
    c = np.where(q >= 0, np.where(q >= g.astype(np.float32), g - 1, np.floor(np.nan_to_num(q)).astype(np.int64)), 0)
    key = (c[:, 0] * g[1] + c[:, 1]) * g[2] + c[:, 2]
    lab = np.asarray(lab)
    C = int(g.prod())
    cnt = [np.bincount(key[lab == ph], minlength=C) for ph in (0, 1)]
    other = np.zeros(len(key), np.int64)
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                x, y, z = c[:, 0] + dx, c[:, 1] + dy, c[:, 2] + dz
                ok = (x >= 0) & (x < g[0]) & (y >= 0) & (y < g[1]) & (z >= 0) & (z < g[2])
                k = np.where(ok, (x * g[1] + y) * g[2] + z, 0)
                other += np.where(ok, np.where(lab == 0, cnt[1][k], cnt[0][k]), 0)
    listed = other > 0
    return int((listed & (lab == 0)).sum()), int((listed & (lab == 1)).sum())


def cpu_baseline(scene, steps: int, dt: float, small_steps: int, one_thread_steps: int) -> dict:
    """The FP32 C++ oracle (OpenMP over particles) on the host cores this
    process may use (the GPU's CPU share, OMP_NUM_THREADS when set), on the
    headline scene; plus a 1-thread sample of the same scene and the small
    block scene (24,576 particles) for `small_steps` steps."""
    import oracle

    from sph_seepage_flows_amd import scenes

    oracle.build()
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    want = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or avail
    cores = oracle.threads(min(want, avail))
    o = oracle.OracleSystem(scene, 32)
    o.step(dt, 1)  # warm-up (first-touch allocation)
    t = time.perf_counter()
    o.step(dt, steps)
    sec = time.perf_counter() - t
    out = {"value": scene.n * steps / sec, "unit": "particle-steps/s", "cores": cores, "kind": "port",
           "sample": f"FP32 C++ oracle (OpenMP, {cores} threads), {steps} steps of the same {scene.n}-particle "
                     f"scene after 1 warm-up step; {sec:.1f} s",
           "host_cpus_reported": os.cpu_count(), "cpus_in_affinity": avail}
    if one_thread_steps:
        oracle.threads(1)
        t = time.perf_counter()
        o.step(dt, one_thread_steps)
        sec1 = time.perf_counter() - t
        out["one_thread"] = {"value": scene.n * one_thread_steps / sec1, "cores": 1,
                             "sample": f"{one_thread_steps} further steps of the same scene; {sec1:.1f} s"}
        oracle.threads(cores)
    o.close()
    if small_steps:
        sb = scenes.small_block()
        o = oracle.OracleSystem(sb, 32)
        t = time.perf_counter()
        o.step(dt, small_steps)
        secs = time.perf_counter() - t
        o.close()
        out["small_block"] = {"value": sb.n * small_steps / secs, "cores": cores,
                              "sample": f"{sb.name}: {sb.n} particles, {small_steps} steps from the initial "
                                        f"state; {secs:.1f} s"}
    return out


def ensemble_bench(dt: float, steps_batch: int, steps_plain: int) -> dict:
    """Batched ensemble: 256 water-block + sand-pile scenes (5k-60k particles
    each, per-scene friction / capillary / drag, seed 3) stepped by one
    batched handle (HIP-graph replay, tuned launch shapes) versus the plain
    loop of 256 single-scene handles, each stepped in turn."""
    from sph_seepage_flows_amd import BatchSystem, SFSystem, scenes

    E = scenes.ensemble()
    n = sum(e.n for e in E)
    b = BatchSystem.from_scenes(E, device=0)
    b.step(dt, 2)
    tuned = b.autotune(reps=5)
    b.step_graph(dt, 2)  # capture outside the timed region
    b.synchronize()
    t = time.perf_counter()
    b.step_graph(dt, steps_batch)
    tb = (time.perf_counter() - t) / steps_batch
    b.close()
    hs = [SFSystem.from_scene(sc, device=0) for sc in E]
    for h in hs:
        h.step(dt, 1)
    t = time.perf_counter()
    for h in hs:
        h.step(dt, steps_plain)
    tp = (time.perf_counter() - t) / steps_plain
    for h in hs:
        h.close()
    return {"workload": f"256 scenes, {n} particles (seed 3)", "ms_per_step": tb * 1e3,
            "particle_steps_per_s": n / tb, "plain_loop_ms_per_step": tp * 1e3,
            "plain_loop_particle_steps_per_s": n / tp, "speedup_vs_plain_loop": tp / tb,
            "steps": [steps_batch, steps_plain],
            "launch_shapes": {k: [v[0], v[1]] for k, v in tuned.items()}}


def diff_bench(dt: float, check: bool) -> dict:
    """Differentiable seepage column: 48k sand + 16k water poured from above,
    200-step window, checkpoints every 20 steps.  Loss = mean squared deviation
    of the final sand positions and saturations from a target produced with
    hidden parameters (drag D 4e3, capillary sigma 0.03, friction 0.55);
    gradient at (5e3, 0.02, 0.5) and w.r.t. the initial velocities.  Times the
    window's forward, its backward (segment recompute + reverse passes) and a
    plain forward of the same steps.  With `check`, the gradient of a 2-step
    window of the same scene is compared with the oracle's FP64 tangent (a
    dot-product identity per direction).  The column is small and well mixed,
    so both the window's handle and the plain-forward handle run the
    interface passes with 4 lanes per particle."""
    import numpy as np

    from sph_seepage_flows_amd import SFSystem, scenes
    from sph_seepage_flows_amd.diff import DiffWindow, msd_loss, physical_grad, theta_of

    sc = scenes.column()
    T, K = 200, 20
    sand = sc.label == 1
    th_hidden, th0 = theta_of(4e3, 0.03, 0.55), theta_of(5e3, 0.02, 0.5)
    lanes = 4
    win = DiffWindow(sc, T, dt, K, device=0)
    win.sys.set_interface_lanes(lanes)
    tp, ts = win.forward(th_hidden)
    pos, sat = win.forward(th0)
    loss, pb, sb = msd_loss(pos, sat, tp, ts, sand)
    win.backward(pb, sb)  # warm-up (tape allocation)
    reps = 3
    t = time.perf_counter()
    for _ in range(reps):
        win.forward(th0)
    tf = (time.perf_counter() - t) / reps
    t = time.perf_counter()
    for _ in range(reps):
        th, gx, gv = win.backward(pb, sb)
    tb = (time.perf_counter() - t) / reps
    g = SFSystem.from_scene(sc, device=0)
    g.set_params(win.sys._p)
    g.set_interface_lanes(lanes)
    g.step(dt, 2)
    g.synchronize()
    t = time.perf_counter()
    g.step(dt, T)
    tplain = time.perf_counter() - t
    out = {"workload": f"column: {int(sand.sum())} sand + {int((~sand).sum())} water, {T} steps, dt={dt}, "
                       f"checkpoint every {K}, interface passes {lanes} lanes per particle",
           "forward_ms": tf * 1e3, "adjoint_ms": tb * 1e3,
           "adjoint_over_forward": tb / tf, "plain_forward_ms": tplain * 1e3, "adjoint_over_plain_forward": tb / tplain,
           "loss": loss, "grad": physical_grad(th, 5e3), "grad_vel0_norm": float(np.linalg.norm(gv)),
           "grad_finite": bool(np.isfinite(th).all() and np.isfinite(gv).all())}
    if check:
        import oracle

        oracle.build()
        oracle.threads(min(16, len(os.sched_getaffinity(0))))
        steps = 2
        w2 = DiffWindow(sc, steps, dt, 1, device=0)
        h2 = DiffWindow(sc, steps, dt, 1, device=0)
        tp2, ts2 = h2.forward(th_hidden)
        p2, s2 = w2.forward(th0)
        _, pb2, sb2 = msd_loss(p2, s2, tp2, ts2, sand)
        th2, _, gv2 = w2.backward(pb2, sb2)
        p, b, o = sc.params()
        for name, v in zip(("sf_cd", "sf_c0", "sf_cmc", "dem_tan_friction_angle"), th0):
            setattr(p, name, float(v))
        rng = np.random.default_rng(0)
        u = rng.standard_normal((sc.n, 3))
        vdot = np.zeros((4, sc.n, 3)); vdot[0] = u
        pdot = np.zeros((4, 4)); pdot[1, 0] = 1.0; pdot[2, 1] = 1.0; pdot[3, 3] = 1.0
        _, _, pd, sd = oracle.tangent_window(sc, dt, steps, vdot=vdot, pdot=pdot, params=(p, b, o))
        _, _, pd2, sd2 = oracle.tangent_window(sc, dt, steps, pdot=np.array([[0, 0, 1.0, 0]]), params=(p, b, o))
        J = lambda a, c: float((pb2 * a).sum() + (sb2 * c).sum())
        pairs = {"vel0_direction": (float((gv2 * u).sum()), J(pd[0], sd[0])), "sf_cd": (float(th2[0]), J(pd[1], sd[1])),
                 "sf_c0": (float(th2[1]), J(pd[2], sd[2])), "sf_cmc": (float(th2[2]), J(pd2[0], sd2[0])),
                 "tan_friction": (float(th2[3]), J(pd[3], sd[3]))}
        out["check_2_steps_vs_fp64_tangent"] = {
            k: {"adjoint": a, "tangent": r, "rel": abs(a - r) / max(abs(r), 1e-30)} for k, (a, r) in pairs.items()}
    return out


def frame_bench(scene, frames: int) -> dict:
    """The reference driver's adaptive frame loop (main.cpp:235-247, 1/30 s
    frames): sf_update_system called per sub-step from a Python host loop
    (the loop main.cpp runs, one host read of the sub-step count per call)
    against sf_update_frame (the same loop inside the library; identical
    results).  Wall-clock ms per frame after one untimed frame each."""
    import numpy as np

    import sph_seepage_flows_amd as sf

    ri = 1.0 / 30.0

    def host_frame(g):
        rem, k = np.float32(ri), 0
        while rem > np.float32(1e-6):
            g.UpdateSystem(float(rem))
            rem = np.float32(rem - rem / np.float32(g.GetNumOfSubTimeSteps()))
            k += 1
        return k

    res = {"workload": f"{scene.name}: {scene.n} particles, 1/30 s frames", "frames": frames}
    for mode in ("host_loop", "device_frame"):
        g = sf.SFSystem.from_scene(scene, device=0)
        run = host_frame if mode == "host_loop" else (lambda g: g.update_frame(ri)[0])
        run(g)
        g.synchronize()
        t0 = time.perf_counter()
        subs = sum(run(g) for _ in range(frames))
        g.synchronize()
        wall = time.perf_counter() - t0
        g.close()
        res[mode] = {"ms_per_frame": wall * 1e3 / frames, "sub_steps_per_frame": subs / frames,
                     "us_per_sub_step": wall * 1e6 / subs, "particle_steps_per_s": scene.n * subs / wall}
    res["speedup"] = res["host_loop"]["ms_per_frame"] / res["device_frame"]["ms_per_frame"]
    return res


def mixed_bench(scene, dt: float, steps: int) -> dict:
    """The headline scene after 0.2 s of simulated time (six 1/30 s frames of
    the reference driver's loop from the initial state): the water column has
    run into the sand pile, so every neighbour pass carries interface work.
    Then the per-pass breakdown of 10 profiled steps and `steps` steps timed
    with events around the water momentum pass (every EVENT_EVERY-th step) only."""
    import sph_seepage_flows_amd as sf

    g = sf.SFSystem.from_scene(scene, device=0)
    for _ in range(6):
        g.update_frame(1.0 / 30.0)
    _, passes = g.step_profiled(dt, PROFILE_STEPS, per_pass=True)
    total_ms, _ = g.step_timed(dt, steps, "water_momentum", every=min(EVENT_EVERY, steps))
    n = g.sizes()[0]
    g.close()
    return {"workload": f"{scene.name} after 0.2 s (six 1/30 s frames)", "steps": steps,
            "ms_per_step": total_ms / steps, "particle_steps_per_s": n * steps / (total_ms / 1e3),
            "pass_ms_profiled": passes}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--cpu-small-steps", type=int, default=2000,
                    help="steps of the small-block scene in the CPU baseline (0: skip)")
    ap.add_argument("--cpu-one-thread-steps", type=int, default=2,
                    help="1-thread steps of the headline scene in the CPU baseline (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", default="drop_in_1m", choices=["drop_in_1m", "small_block", "dam_4m"])
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble measurement")
    ap.add_argument("--no-diff", action="store_true", help="skip the differentiable-window measurement")
    ap.add_argument("--parts", default="step,mixed,ensemble,diff,frame",
                    help="comma list of step (the headline workload), mixed (the same scene after 0.2 s), ensemble, "
                         "diff, frame; a profiler can run one")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on exactly one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU slab decomposition is design-only)", file=sys.stderr)
        return 2
    rank = int(os.environ.get("RANK", "0"))

    import numpy as np

    import sph_seepage_flows_amd as sf
    from sph_seepage_flows_amd import scenes

    parts = set(args.parts.split(","))
    scene = {"drop_in_1m": scenes.drop_in_1m, "small_block": scenes.small_block, "dam_4m": scenes.dam}[args.workload]()
    dt = 1e-4
    out = {
        "metric": METRIC,
        "value": None,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": scene.name, "parallelism": "single-gpu"},
        "roofline": None,
        "cpu_baseline": None,
        "batched": None,
        "differentiation": None,
        "frame_loop": None,
        "mixed_state": None,
    }
    if "step" in parts:
        print(f"[bench] scene {scene.name}: n={scene.n} nb={scene.nb}", file=sys.stderr, flush=True)
        g = sf.SFSystem.from_scene(scene, device=0)
        n, nb, cells = g.sizes()
        if args.warmup:
            g.step(dt, args.warmup)
        # untimed: per-pass breakdown (an event between every two passes) and
        # the dominant pass; the timed region then records events around that
        # pass only, on every EVENT_EVERY-th step, since each event between
        # kernels costs the stream a few microseconds
        _, pass_ms = g.step_profiled(dt, PROFILE_STEPS, per_pass=True)
        dom = max(pass_ms, key=pass_ms.get)
        g.synchronize()
        t0 = time.perf_counter()
        total_ms, dom_ms = g.step_timed(dt, args.steps, dom, every=min(EVENT_EVERY, args.steps))
        wall = time.perf_counter() - t0
        # neighbour-pair throughput of the dominant pass: its particles' pairs
        # within h (self included: the passes evaluate it, with a zero term)
        # and their 27-cell candidates, counted on the final state
        cnt = g.get("nbr_count").astype(np.int64)
        cand = g.get("nbr_candidates").astype(np.int64)
        lab = g.get("label")
        mine = {"water_momentum": lab == 0, "density": lab == 0, "sand_momentum": lab == 1}.get(dom, lab >= 0)
        pairs, cands = int(cnt[mine].sum()), int(cand[mine].sum())
        nbr_pairs = {"particles": int(mine.sum()), "pairs_within_h_per_launch": pairs,
                     "candidates_per_launch": cands, "pairs_per_s": pairs / (dom_ms * 1e-3),
                     "candidates_per_s": cands / (dom_ms * 1e-3)}
        listed = interface_counts(g.get("pos"), lab, g._b)
        stiff = g._p.sph_stiff
        g.close()
        # the configuration's own run: 500 steps from the initial state, one
        # event pair around all of them (skipped by short profiling runs)
        run500 = None
        if args.steps >= 100:
            g5 = sf.SFSystem.from_scene(scene, device=0)
            t500, _ = g5.step_profiled(dt, 500, per_pass=False)
            g5.close()
            run500 = {"steps": 500, "ms_per_step_mean": t500 / 500, "particle_steps_per_s": n * 500 / (t500 / 1e3)}
            print(f"[bench] 500 steps from the initial state: {t500 / 500:.4f} ms/step", file=sys.stderr, flush=True)
        ms_per_step = total_ms / args.steps
        print(f"[bench] {args.steps} steps: {ms_per_step:.3f} ms/step (events), wall {wall * 1e3 / args.steps:.3f} "
              f"ms/step; {dom} {dom_ms:.4f} ms; passes of {PROFILE_STEPS} profiled steps "
              f"{json.dumps({k: round(v, 4) for k, v in pass_ms.items()})}", file=sys.stderr, flush=True)
        n_s = int(np.count_nonzero(scene.label == 1))
        n_w = n - n_s
        pb = pass_bytes(n_w, n_s, nb, cells, scene.physics.sf_type == 1, listed)
        ach = pb[dom] / (dom_ms * 1e-3) / 1e9
        step_bytes = sum(pb.values())
        traffic = None
        prof = ROOT / "profiles" / f"pmc_{dom}_{scene.name}.json"
        if prof.exists():
            try:
                traffic = json.loads(prof.read_text()).get("hbm_bytes_per_launch")
            except Exception:
                traffic = None
        # The neighbour passes are bound by instruction issue, not HBM: VALU
        # busy = (VALU wave-instructions per launch, from the committed PMC
        # profile) x 4 cycles / (1024 SIMDs x 2.4 GHz x this run's mean launch time).
        valu = None
        vprof = ROOT / "profiles" / f"pmc_valu_{dom}_{scene.name}.json"
        if vprof.exists():
            try:
                insts = float(json.loads(vprof.read_text())["valu_insts_per_launch"])
                valu = {"valu_insts_per_launch": insts, "busy_frac": insts * 4 / (1024 * 2.4e9 * dom_ms * 1e-3),
                        "source": str(vprof.relative_to(ROOT))}
            except Exception:
                valu = None
        out.update({
            "value": n * args.steps / (total_ms / 1e3),
            "ms_per_step": ms_per_step,
            "config": {"workload": f"{scene.name}: {n_w} water + {n_s} sand, {nb} boundary particles, "
                                   f"{cells} grid cells, dt=1e-4, FP32, Tait stiffness B = {stiff:.1f} Pa (the "
                                   f"reference driver's 0.001 c^2 rho0 / 7 with c = 100, main.cpp:48,53; "
                                   f"c0 = 40 m/s would give c0^2 rho0 / 7 = 228571 Pa)", "n_particles": n,
                       "tait_stiffness_pa": stiff, "parallelism": "single-gpu"},
            "roofline": {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": ach / HBM_PEAK_GBS, "traffic": traffic, "alg_bytes_per_launch": pb[dom],
                         "mean_launch_ms": dom_ms, "valu": valu, "neighbour_pairs": nbr_pairs},
            "step_roofline": {"alg_bytes_per_particle_step": step_bytes / n,
                              "achieved_GBs": step_bytes / (ms_per_step * 1e-3) / 1e9,
                              "frac": step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS},
            "pass_ms_profiled": pass_ms,
            "interface_particles": {"water": listed[0], "sand": listed[1]},
            "run_500_steps": run500,
        })
    if "mixed" in parts:
        print("[bench] mixed state (6 frames first)...", file=sys.stderr, flush=True)
        out["mixed_state"] = mixed_bench(scene, dt, 100)
        print(f"[bench] mixed state: {json.dumps(out['mixed_state'])}", file=sys.stderr, flush=True)
    if "ensemble" in parts and not args.no_ensemble:
        print("[bench] batched ensemble...", file=sys.stderr, flush=True)
        out["batched"] = ensemble_bench(dt, 50, 5)
        print(f"[bench] batched: {json.dumps(out['batched'])}", file=sys.stderr, flush=True)
    if "diff" in parts and not args.no_diff:
        print("[bench] differentiable window...", file=sys.stderr, flush=True)
        out["differentiation"] = diff_bench(dt, check=not args.no_cpu_baseline)
        print(f"[bench] differentiation: {json.dumps(out['differentiation'])}", file=sys.stderr, flush=True)
    if "frame" in parts:
        print("[bench] frame loop...", file=sys.stderr, flush=True)
        out["frame_loop"] = [frame_bench(scenes.small_block(), 5), frame_bench(scene, 3)]
        print(f"[bench] frame loop: {json.dumps(out['frame_loop'])}", file=sys.stderr, flush=True)
    if rank == 0 and "step" in parts and not args.no_cpu_baseline:
        print(f"[bench] cpu baseline ({args.cpu_steps} oracle steps)...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(scene, args.cpu_steps, dt, args.cpu_small_steps, args.cpu_one_thread_steps)
    if rank == 0:
        print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
