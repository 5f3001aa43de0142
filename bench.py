"""Benchmark: VBD vertex-updates/s on the 512x512 cloth (one MI355X).

Workload: the 512x512 cloth (262,144 vertices, 522,242 StVK triangles, 782,341 bending hinges,
512 anchored vertices), Y drawn log-uniform in [1e3, 1e5] from --seed (default 3: Y = 4.95e4, a draw
for which the reference VBD converges), dt = 1/60, 20 iterations per step, FP32 on the GPU.
Inputs are resident in HBM before the timed region. A step = one VBD::step (inertial predict, 20
iterations x 4 colour sweeps, velocity update), captured once in a hipGraph and replayed.

value      = vertices x iterations x steps / wall time of the timed steps (synchronised both sides)
roofline   = the vertex-solve kernel: compulsory bytes per launch / mean launch duration, launch
             durations from HIP events on the solver's stream over an event-timed replay of the same
             steps right after the throughput region
cpu_baseline = the FP64 OpenMP oracle (restatement of spg's VBD::step), closed-form derivatives,
             repeated first steps of the same scene on the host cores (about 10 s of CPU work), plus a
             one-thread sample and a dual-number (AD-cost) sample
traffic    = HBM bytes per vertex-solve launch from the committed rocprofv3 counter passes of this
             workload and launch shape (profiles/), null when the shape differs

learned    = time to a 1e-4 relative FP64 force residual, learned warm start vs plain VBD, on the same
             cloth (network trained in this run on the GPU from 256 config-4-style cloths)
batched    = 256-cloth batch (config 4) throughput, one launch per colour for all instances
volumetric = 64^3-cell Stable Neo-Hookean cube (config 3) throughput
xpbd       = the XPBD parallel Gauss-Seidel step (spg's XPBD(true)) on the same cloth: stencil projections/s,
             per-type projection-kernel rooflines, and the FP64 oracle's XPBD step on the host cores
further_energies = VBD on the same cloth built from the further energies (Baraff-Witkin membrane and bending,
             continuum edge springs), and the oracle's step (one iteration) on the same cloth on the host cores
rigid_bodies = VBD on rigid bodies (FP64 6x6 blocks) on a 512x512 lattice of boxes joined by springs and
             orientation alignments, with the colour-solve roofline and the oracle's step on the same lattice
implicit   = the Baraff-Witkin implicit Euler step (FP64: element Hessians, block-sparse assembly, Jacobi CG with
             the reference's 1e-5 / 300-iteration settings) on the same cloth, the roofline of the symmetric
             block-sparse SpMV sampled inside the CG, and the oracle's step on a 128x128 cloth on the host cores

Usage: python bench.py [--gpus 1] [--steps 200] [--warmup 10] [--seed 3] [--no-cpu] [--no-learned]
                       [--no-batched] [--no-volumetric] [--no-xpbd] [--xpbd-substeps 20] [--no-implicit]
                       [--no-rigid] [--rigid-n 512]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--lds", type=int, default=-1, help="lds_staging of the vertex solve (with --lanes/--block)")
    ap.add_argument("--timing-steps", type=int, default=20)
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-warm-start comparison")
    ap.add_argument("--learned-steps", type=int, default=20, help="trajectory steps of the learned comparison")
    ap.add_argument("--no-batched", action="store_true", help="skip the 256-cloth batched measurement")
    ap.add_argument("--no-volumetric", action="store_true", help="skip the 64^3 cube measurement")
    ap.add_argument("--no-further", action="store_true", help="skip the further-energies VBD measurement")
    ap.add_argument("--no-xpbd", action="store_true", help="skip the XPBD measurement")
    ap.add_argument("--xpbd-substeps", type=int, default=20, help="XPBD substeps per step")
    ap.add_argument("--xpbd-steps", type=int, default=30, help="timed XPBD steps")
    ap.add_argument("--no-implicit", action="store_true", help="skip the implicit Euler measurement")
    ap.add_argument("--no-rigid", action="store_true", help="skip the rigid-body VBD measurement")
    ap.add_argument("--rigid-n", type=int, default=512, help="side of the rigid-body lattice")
    ap.add_argument("--implicit-steps", type=int, default=4, help="timed Baraff-Witkin steps")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run the plain copy kernel of BYTES (counter calibration) and exit")
    ap.add_argument("--barrier-calibration", action="store_true",
                    help="only measure a barrier inside one cooperative launch against back-to-back launches, and exit")
    args = ap.parse_args()

    if args.barrier_calibration:
        from spg_amd import native
        rows = []
        for grid in (256, 512, 1024, 2048):
            for group in (grid, 8):
                b, l = native.bench_barrier(grid, block=256, group=group, rounds=500)
                rows.append({"grid": grid, "block": 256, "group": group, "us_per_barrier": b, "us_per_launch": l})
                print(json.dumps(rows[-1]), file=sys.stderr, flush=True)
        print(json.dumps({"barrier_calibration": rows}))
        return

    if args.copy_calibration:
        from spg_amd import native
        ms = native.bench_copy(args.copy_calibration, repeats=5)
        print(json.dumps({"copy_bytes": args.copy_calibration, "ms": ms,
                          "GB/s": 2 * args.copy_calibration / (ms * 1e-3) / 1e9}))
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: the VBD step runs on one GPU here; runs on several GPUs are not supported "
              "(instances would shard by object, a single mesh by vertex ownership with per-colour halo "
              "exchange)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from spg_amd import native, scenes

    s = scenes.cloth_config2(seed=args.seed, n=args.n)
    obj = s.vbd_object()
    t0 = time.perf_counter()
    ctx = native.VbdContext([obj], device=0)
    create_s = time.perf_counter() - t0  # colouring + incidence streams + upload (vbd.cpp:14-18 on the first step)
    tuned = None
    if args.lanes or args.block:
        kw = {}
        if args.lanes:
            kw["lanes_per_vertex"] = args.lanes
        if args.block:
            kw["block_size"] = args.block
        if args.lds >= 0:
            kw["lds_staging"] = args.lds
        ctx.set_launch(**kw)
    elif not args.no_autotune:
        tuned = ctx.autotune(trials=4)
    ctx.set_launch(use_graph=1)
    V = s.n
    iters = s.iterations
    print(f"# scene {s.meta} colours {ctx.num_colors()} launch {ctx.launch()}", file=sys.stderr, flush=True)

    # Every step starts from the workload's initial state (device checkpoint restored on the GPU inside
    # the timed region): the reference VBD on this cloth converges on its first step but its trajectory
    # leaves the basin within ~10 steps, so repeated first steps are the meaningful unit of work.
    ctx.save_state()
    for _ in range(args.warmup):
        ctx.restore_state()
        ctx.step(s.dt, 1, iters, s.gravity)
    ctx.synchronize()

    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.restore_state()
        ctx.step(s.dt, 1, iters, s.gravity)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    value = V * iters * args.steps / wall
    print(f"# timed {args.steps} steps in {wall * 1e3:.2f} ms", file=sys.stderr, flush=True)

    # event-timed replay of the same steps (non-graph launches, one event pair per vertex-solve launch)
    ctx.reset_stats()
    ctx.set_timing(True)
    for _ in range(args.timing_steps):
        ctx.restore_state()
        ctx.step(s.dt, 1, iters, s.gravity)
    ctx.synchronize()
    st = ctx.stats()
    ctx.set_timing(False)
    mean_ms = st["solve_ms"] / max(1, st["solve_launches"])
    bytes_per_launch = st["algorithmic_bytes"] / max(1, st["solve_launches"])
    achieved = bytes_per_launch / (mean_ms * 1e-3) / 1e9
    x, v = ctx.state()
    finite = bool(np.all(np.isfinite(x)))
    traffic, traffic_src = measured_traffic(ctx.launch(), args)
    # the force residual of the headline step, FP64 on the GPU: one substep of `iters` sweeps from the
    # workload's initial state, |r(x)| / |r(x~)| (the oracle's value for the same x, x~ is in the CPU leg)
    ctx.restore_state()
    ctx.begin_substep(s.dt, s.gravity)
    x_tilde, _ = ctx.state()
    ctx.iterate(iters)
    r_x, r_xt = float(ctx.residual()[0]), float(ctx.residual_at_inertial()[0])
    x_iter, _ = ctx.state()
    res_us, res_bytes = ctx.bench_residual(20)
    ctx.end_substep()
    step_residual = {"gpu_fp64_rho": r_x / r_xt, "gpu_fp64_residual": r_x, "gpu_fp64_residual_at_inertial": r_xt,
                     "iterations": iters, "state": "one substep from the workload's initial state",
                     "residual_us": res_us, "residual_in_solve_launches": res_us / (mean_ms * 1e3),
                     "roofline": {"bound": "hbm", "kernel": "hingeKcKernel + residualNbKernel + reduceSegmentsKernel + "
                                                           "reduceObjectsKernel (one FP64 residual evaluation)",
                                  "achieved": res_bytes / (res_us * 1e-6) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": res_bytes / (res_us * 1e-6) / 1e9 / HBM_PEAK_GBS,
                                  "algorithmic_bytes": res_bytes, "timing": "HIP events over 20 back-to-back "
                                                                             "evaluations on the context's stream"}}
    kernel = {0: "solveColorKernel", 1: "solveColorKernel (element rows staged in LDS)",
              2: "solveColorNbKernel"}[ctx.launch()["lds_staging"]]

    batched = None if args.no_batched else batched_throughput(args)
    volumetric = None if args.no_volumetric else volumetric_throughput(args)
    further = None if args.no_further else further_throughput(args)
    xpbd = None if args.no_xpbd else xpbd_throughput(s, args)
    implicit = None if args.no_implicit else implicit_throughput(s, args)
    rigid = None if args.no_rigid else rigid_throughput(args)
    learned_cmp = None if args.no_learned else learned_comparison(s, args, ctx.launch())

    cpu = None
    setup = {"gpu_create_s": create_s,
             "what": "spg_vbd_create: greedy colouring, incidence streams and neighbourhood rows built on the host, "
                     "uploaded (the reference does this inside its first step, vbd.cpp:14-18); not in the timed "
                     "region"}
    if not args.no_cpu:
        cpu = cpu_baseline(s)
        setup["oracle_precompute_s"] = cpu.pop("precompute_s")
        step_residual.update(oracle_residual(s, x_iter, x_tilde))

    out = {
        "metric": "VBD vertex-updates/s on 512² cloth; time to 1e-4 residual, learned vs plain",
        "value": value,
        "unit": "vertex-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"cloth {args.n}x{args.n} (V={V}, {s.meta['triangles']} StVK triangles, "
                               f"{s.meta['hinges']} bending hinges, top row anchored), Y={s.meta['young']:.5g} "
                               f"(seed {args.seed}), dt=1/60, {iters} iterations/step, hipGraph replay",
                   "vertices": V, "iterations": iters, "colors": ctx.num_colors(),
                   "launch": ctx.launch(), "autotuned": tuned},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "traffic_note": "2.0 FETCH_SIZE + 1.0 WRITE_SIZE per launch (rocprofv3 PMC, scaled by the same "
                                     "counters on a copy kernel): bytes between the L2 caches and the memory fabric. "
                                     "The 512x512 working set (~170 MB of records and state) stays in the 256 MB "
                                     "Infinity Cache between sweeps, so this is an upper bound on DRAM bytes",
                     "kernel": kernel, "mean_launch_us": mean_ms * 1e3,
                     "algorithmic_bytes_per_launch": bytes_per_launch,
                     "launches_timed": st["solve_launches"]},
        "solver_stats": {"fp64_fallbacks": st["fp64_fallbacks"], "skipped_blocks": st["skipped_blocks"],
                         "state_finite": finite},
        "step_residual": step_residual,
        "setup": setup,
        "cpu_baseline": cpu,
        "learned": learned_cmp,
        "batched": batched,
        "volumetric": volumetric,
        "further_energies": further,
        "xpbd": xpbd,
        "implicit": implicit,
        "rigid_bodies": rigid,
    }
    print(json.dumps(out))


def _throughput(scene_list, steps):
    """Autotuned, graph-captured steps of the scenes advanced together (each step from the initial
    state); returns (vertex-updates/s, ms per step, launch, event-timed solve launch us, algorithmic
    GB/s of the solve, the time and roofline fraction of one FP64 residual evaluation after one substep's
    sweeps, all positions finite, context setup s)."""
    import numpy as np
    from spg_amd import native

    t0 = time.perf_counter()
    ctx = native.VbdContext([b.vbd_object() for b in scene_list], device=0)
    setup = time.perf_counter() - t0
    tuned = ctx.autotune(trials=4)
    ctx.set_launch(use_graph=1)
    V = sum(b.n for b in scene_list)
    iters = scene_list[0].iterations
    dt, g = scene_list[0].dt, scene_list[0].gravity
    ctx.save_state()
    for _ in range(3):
        ctx.restore_state()
        ctx.step(dt, 1, iters, g)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.restore_state()
        ctx.step(dt, 1, iters, g)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ctx.reset_stats()
    ctx.set_timing(True)
    ctx.restore_state()
    ctx.step(dt, 1, iters, g)
    ctx.synchronize()
    st = ctx.stats()
    ctx.set_timing(False)
    finite = all(bool(np.isfinite(ctx.state(o)[0]).all()) for o in range(len(scene_list)))
    mean_ms = st["solve_ms"] / max(1, st["solve_launches"])
    # one FP64 residual evaluation (the convergence check) at the state after one substep's sweeps
    ctx.restore_state()
    ctx.begin_substep(dt, g)
    ctx.iterate(iters)
    res_us, res_bytes = ctx.bench_residual(20)
    ctx.end_substep()
    return {"vertices": V, "colors": ctx.num_colors(), "launch": tuned, "value": V * iters * steps / wall,
            "unit": "vertex-updates/s", "ms_per_step": wall * 1e3 / steps, "steps": steps,
            "solve_launch_us": mean_ms * 1e3,
            "solve_achieved_GBs": st["algorithmic_bytes"] / max(1, st["solve_launches"]) / (mean_ms * 1e-3) / 1e9,
            "fp64_fallbacks": st["fp64_fallbacks"], "skipped_blocks": st["skipped_blocks"],
            "residual_us": res_us, "residual_frac": res_bytes / (res_us * 1e-6) / 1e9 / HBM_PEAK_GBS,
            "setup_s": setup, "state_finite": finite}


def batched_throughput(args, count=256, steps=30):
    """256 independent cloths of side U{16..128} advanced together: one launch per colour covers every
    instance, launch shape from the autotuner, each step captured in a hipGraph."""
    from spg_amd import scenes

    batch = scenes.cloth_batch(count=count, seed=7)
    out = {"workload": f"{count} cloths, side U{{16..128}} (seed 7), Y log-U[1e3,1e5], bending log-U[1e-5,1e-3], "
                       f"top row anchored, {batch[0].iterations} iterations/step, one launch per colour for all "
                       f"instances, hipGraph replay"}
    out.update(_throughput(batch, steps))
    return out


def volumetric_throughput(args, steps=20):
    """The 64^3-cell Stable Neo-Hookean cube (274,625 vertices, 1,572,864 tets, 8 colours)."""
    from spg_amd import scenes

    cube = scenes.cube_config3()
    out = {"workload": "cube 64^3 cells, 6 tets per cell (V=274625, 1572864 SNH tets, Y 1e5, nu 0.45), bottom face "
                       "anchored, rest scaled 1.05 in y, 20 iterations/step, hipGraph replay"}
    out.update(_throughput([cube], steps))
    return out


def further_throughput(args, steps=20):
    """VBD on the bench's cloth built from the further energies: Baraff-Witkin membrane and bending, and
    continuum springs on the mesh edges (one extra incidence class whose vertex blocks come from forward-mode
    passes over the closed-form element gradients). CPU baseline: the oracle's VBD step (dual numbers, as the
    reference's TinyAD) on the same 512x512 cloth, one iteration per step, about 5 s on the host cores."""
    import numpy as np
    from spg_amd import scenes

    def build(n):
        base = scenes.cloth_config2(seed=3, n=n)
        yt = base.meta["young"] * 1e-3
        return scenes.cloth(n, young=base.meta["young"], poisson=0.3, thickness=1e-3, bending=1e-4,
                            mass=1e-3 * (32.0 / n) ** 2, pin="top_row", seed=3, iterations=20,
                            membrane="membrane_bw", bending_kind="bending_bw", edge_springs="spring_continuum",
                            edge_k=0.1 * yt)

    cloth = build(512)
    out = {"workload": "cloth 512x512 as the main line (Y from seed 3), Baraff-Witkin membrane (kWeft = kWarp = Y t, "
                       "kBias = Y t / 2.6) and bending (k 1e-4), continuum springs on the 784,385 edges "
                       "(k 0.1 Y t), top row anchored, 20 iterations/step, hipGraph replay"}
    out.update(_throughput([cloth], steps))
    out["roofline"] = {"bound": "hbm", "kernel": "solveColorKernel (further-energy classes)",
                       "achieved": out["solve_achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": out["solve_achieved_GBs"] / HBM_PEAK_GBS, "mean_launch_us": out["solve_launch_us"]}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as ol
        cores = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
        ol.set_threads(cores)
        o = ol.from_scene(cloth)
        o.colors()  # first-step precomputation outside the sample
        n, t = 0, 0.0
        while t < 5.0:
            o.set_state(cloth.x, np.zeros_like(cloth.x))
            t0 = time.perf_counter()
            o.step(cloth.dt, 1, 1, cloth.gravity, mode=ol.AD)
            t += time.perf_counter() - t0
            n += 1
        out["cpu_baseline"] = {"value": cloth.n * n / t, "unit": "vertex-updates/s", "cores": cores,
                               "kind": "port", "sample": f"{n} first steps of 1 iteration of the same 512x512 cloth, "
                                                         f"FP64 dual numbers, {t:.1f} s"}
    return out


def rigid_throughput(args, steps=20):
    """VBD on rigid bodies (FP64, 6x6 blocks): an n x n lattice of unit boxes (n = --rigid-n, default 512:
    262,144 bodies, 523,264 springs and as many orientation alignments, one anchored row, random tilts up to
    0.05 rad), dt 1/60, 10 iterations per step over 2 colours, hipGraph replay. Roofline: the colour-solve kernel,
    algorithmic bytes of one sweep (the body's state read and written, and per incident element its record and
    the other body's centre and rotation) over the summed HIP-event durations of the sweep's launches.
    CPU baseline: the oracle's step (dual numbers, as the reference's TinyAD) on the same lattice, one iteration
    per step."""
    import numpy as np
    from spg_amd import native, scenes

    sc = scenes.rb_lattice(args.rigid_n)
    ctx = native.RigidContext([sc])
    for _ in range(3):
        ctx.step(sc.dt, 1, sc.iterations, sc.gravity)
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.step(sc.dt, 1, sc.iterations, sc.gravity)
    wall = time.perf_counter() - t0
    st = ctx.stats()
    samples = [ctx.profile_substep(sc.dt, sc.iterations, sc.gravity) for _ in range(3)]
    ms = min(m for m, _ in samples)
    launches = samples[0][1]
    sweep_ms = ms / sc.iterations
    achieved = st["solve_bytes"] / (sweep_ms * 1e-3) / 1e9
    x = ctx.state()[0]
    out = {"workload": f"rigid-body lattice {args.rigid_n}x{args.rigid_n} ({sc.nb} boxes, {sc.meta['springs']} "
                       f"springs, {sc.meta['alignments']} orientation alignments, {sc.meta['anchors']} anchors), "
                       f"dt=1/60, {sc.iterations} iterations/step, FP64, hipGraph replay",
           "value": sc.nb * sc.iterations * steps / wall, "unit": "body-updates/s", "steps": steps,
           "ms_per_step": wall * 1e3 / steps, "colors": st["n_colors"],
           "skipped_blocks": st["skipped_blocks"], "state_finite": bool(np.all(np.isfinite(x))),
           "roofline": {"bound": "hbm", "kernel": "rbSolveKernel", "achieved": achieved, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                        "algorithmic_bytes_per_sweep": st["solve_bytes"], "launches_per_sweep": launches // sc.iterations,
                        "sweep_us": sweep_ms * 1e3, "mean_launch_us": ms * 1e3 / launches}}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as ol
        cores = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
        ol.set_threads(cores)
        n, t = 0, 0.0
        o = ol.rb_from_scene(sc)
        o.colors()  # first-step precomputation outside the sample
        while t < 5.0:
            o.set_state(x=sc.x, theta=sc.theta, v=np.zeros_like(sc.x), omega=np.zeros_like(sc.x))
            t0 = time.perf_counter()
            o.step(sc.dt, 1, 1, sc.gravity)
            t += time.perf_counter() - t0
            n += 1
        out["cpu_baseline"] = {"value": sc.nb * n / t, "unit": "body-updates/s", "cores": cores, "kind": "port",
                               "sample": f"{n} first steps of 1 iteration of the same {args.rigid_n}x{args.rigid_n} "
                                         f"lattice, FP64 dual numbers, {t:.1f} s"}
    return out


def xpbd_throughput(scene, args):
    """XPBD(true) on the bench's cloth: each step is args.xpbd_substeps substeps of one projection of every
    stencil, in stencil-colour launches (14 bending + 7 membrane + 1 anchor per substep), captured in a
    hipGraph; every step starts from the initial state. Roofline per projection kernel type from an
    event-timed step right after the throughput region."""
    import numpy as np
    from spg_amd import native

    nsub, steps = args.xpbd_substeps, args.xpbd_steps
    t0 = time.perf_counter()
    ctx = native.XpbdContext([scene.xpbd_object()], parallel_gauss_seidel=True)
    setup = time.perf_counter() - t0
    ctx.set_launch(use_graph=1, block_size=256)
    n_st = sum(len(st) for _, st, _ in scene.energies)
    ctx.save_state()
    for _ in range(3):
        ctx.restore_state()
        ctx.step(scene.dt, nsub, scene.gravity)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.restore_state()
        ctx.step(scene.dt, nsub, scene.gravity)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ctx.reset_stats()
    ctx.set_timing(True)
    ctx.restore_state()
    ctx.step(scene.dt, nsub, scene.gravity)
    ctx.synchronize()
    st = ctx.stats()
    ctx.set_timing(False)
    finite = bool(np.isfinite(ctx.state()[0]).all())
    per_type = {}
    for k, d in st["per_type"].items():
        ms = d["ms"] / d["launches"]
        gbs = d["bytes"] / d["launches"] / (ms * 1e-3) / 1e9
        per_type[k] = {"launches": d["launches"], "mean_launch_us": ms * 1e3, "achieved_GBs": gbs,
                       "frac": gbs / HBM_PEAK_GBS, "bytes_per_launch": d["bytes"] / d["launches"]}
    dom = max(per_type, key=lambda k: st["per_type"][k]["ms"])
    out = {"workload": f"same cloth, XPBD parallel Gauss-Seidel, dt=1/60, {nsub} substeps/step, "
                       f"{ctx.num_launches()} projection launches per substep, hipGraph replay",
           "value": n_st * nsub * steps / wall, "unit": "stencil-projections/s",
           "ms_per_step": wall * 1e3 / steps, "steps": steps, "substeps": nsub, "stencils": n_st,
           "projection_ms_per_substep": st["projection_ms"] / nsub,
           "roofline": {"bound": "hbm", "kernel": f"xpbdProjectKernel<{dom}>", "achieved": per_type[dom]["achieved_GBs"],
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": per_type[dom]["frac"]},
           "per_type": per_type, "fp64_solves": st["fp64_solves"], "setup_s": setup, "state_finite": finite}
    if not args.no_cpu:
        out["cpu_baseline"] = xpbd_cpu_baseline(scene)
    return out


def implicit_throughput(scene, args):
    """ImplicitEulerBaraffWitkin::step on the bench's cloth: element Hessians and forces (FP64 forward mode),
    upper-block assembly, A = M - dt^2 K, Jacobi CG (tolerance 1e-5, 300 iterations). Every step starts from
    the initial state. Roofline: the symmetric block-sparse product of A inside the CG, the first product of
    every 8-iteration chunk timed with HIP events on the context's stream during two timed steps."""
    import numpy as np
    from spg_amd import native

    steps = args.implicit_steps
    t0 = time.perf_counter()
    ctx = native.ImplicitContext([scene.vbd_object()], "baraff_witkin")
    setup = time.perf_counter() - t0
    x0, v0 = ctx.state()
    ctx.step(scene.dt, 1, scene.gravity)  # warm-up
    ctx.set_state(0, x0, v0)
    ctx.reset_stats()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.set_state(0, x0, v0)
        ctx.step(scene.dt, 1, scene.gravity)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    st = ctx.stats()
    finite = bool(np.isfinite(ctx.state()[0]).all())
    ctx.reset_stats()
    ctx.set_timing(True)
    timing_steps = 2
    for _ in range(timing_steps):
        ctx.set_state(0, x0, v0)
        ctx.step(scene.dt, 1, scene.gravity)
    tst = ctx.stats()
    ctx.set_timing(False)
    ctx.set_state(0, x0, v0)
    spmv_ms = tst["spmv_sampled_ms"] / max(1, tst["spmv_sampled"])
    b2b_ms = ctx.bench_spmv(1.0, -scene.dt ** 2, repeats=50)
    bytes_spmv = ctx.spmv_bytes()
    achieved = bytes_spmv / (spmv_ms * 1e-3) / 1e9
    dofs = ctx.num_dofs()
    iters = st["cg_iterations"] / steps
    converged = st["last_cg_error"] <= 1e-5
    status = ("CG converged" if converged else
              f"CG not converged: capped at 300 iterations as the reference (implicitEulerBase.h:26-27), relative "
              f"error {st['last_cg_error']:.2e}")
    out = {"workload": f"same cloth, ImplicitEulerBaraffWitkin (FP64), dt=1/60, 1 substep, CG tolerance 1e-5 / "
                       f"300 iterations, {ctx.num_blocks()} 3x3 blocks, every step from the initial state; {status}",
           "cg_converged": converged,
           "value": dofs * iters * steps / wall, "unit": "DOF x CG-iterations/s",
           "ms_per_step": wall * 1e3 / steps, "steps": steps, "dofs": dofs, "cg_iterations_per_step": iters,
           "last_cg_error": st["last_cg_error"], "assembly_ms_per_step": tst["assembly_ms"] / timing_steps,
           "cg_ms_per_step": tst["cg_ms"] / timing_steps,
           "cg_us_per_iteration": tst["cg_ms"] * 1e3 / max(1, tst["cg_iterations"]),
           "roofline": {"bound": "hbm", "kernel": "symSpmvKernel", "achieved": achieved, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "mean_launch_us": spmv_ms * 1e3,
                        "sampled_launches": tst["spmv_sampled"], "algorithmic_bytes_per_launch": bytes_spmv,
                        "back_to_back_launch_us": b2b_ms * 1e3},
           "setup_s": setup, "state_finite": finite}
    if not args.no_cpu:
        out["cpu_baseline"] = implicit_cpu_baseline()
    return out


def implicit_cpu_baseline(budget_s=10.0):
    """The oracle's Baraff-Witkin step (dual-number Hessians as TinyAD, triplet assembly, the Eigen CG
    restatement with an OpenMP SpMV) on a 128x128 cloth of the same construction, repeated from its initial
    state for about budget_s; reported per DOF x CG iteration like the GPU value. One oracle step of the 512x512
    cloth itself takes about 80 s on 8 host cores (dual-number assembly plus 300 CG iterations), above what a
    bench run can spend, so the sample is the smaller mesh and says so."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    import numpy as np
    from spg_amd import scenes

    cores = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    ol.set_threads(cores)
    sc = scenes.cloth_config2(seed=3, n=128)
    o = ol.from_scene(sc)
    v0 = np.zeros_like(sc.x)
    n, t, its = 0, 0.0, 0
    while t < budget_s:
        o.set_state(sc.x, v0)
        t0 = time.perf_counter()
        info = o.implicit_step(ol.BARAFF_WITKIN, sc.dt, 1, sc.gravity)
        t += time.perf_counter() - t0
        its += info["cg_iterations"]
        n += 1
    dofs = 3 * sc.n
    return {"value": dofs * its / t, "unit": "DOF x CG-iterations/s", "cores": cores, "kind": "port",
            "mesh": "128x128 (the GPU line runs 512x512)",
            "sample": f"{n} Baraff-Witkin steps of a 128x128 cloth of the same construction ({dofs} DOFs, "
                      f"{its / n:.0f} CG iterations per step), FP64, {t:.1f} s ({t / n * 1e3:.0f} ms per step)"}


def xpbd_cpu_baseline(scene, budget_s=8.0):
    """FP64 oracle XPBD step (OpenMP over the stencils of a colour, as xpbd.cpp:91 parallelises) on the host
    cores: repeated one-substep steps from the initial state for about budget_s."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    import numpy as np

    cores = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    ol.set_threads(cores)
    o = ol.from_scene(scene)
    v0 = np.zeros_like(scene.x)
    n_st = sum(len(st) for _, st, _ in scene.energies)
    o.xpbd_step(scene.dt, 1, scene.gravity, parallel=True)  # first step: stencil colouring (xpbd.cpp:18-21)
    n, t = 0, 0.0
    while t < budget_s:
        o.set_state(scene.x, v0)
        t0 = time.perf_counter()
        o.xpbd_step(scene.dt, 1, scene.gravity, parallel=True)
        t += time.perf_counter() - t0
        n += 1
    return {"value": n_st * n / t, "unit": "stencil-projections/s", "cores": cores, "kind": "port",
            "sample": f"{n} one-substep XPBD steps of the same cloth after a first step that coloured the "
                      f"stencils, FP64, dual-number constraint gradients (as TinyAD in the reference), {t:.1f} s"}


def learned_comparison(scene, args, launch):
    """Learned warm start against plain VBD on the bench's own cloth. The network is trained here, on
    the GPU, from the solver's own converged trajectories of 256 config-4-style cloths (2M vertex
    samples). Then, for each state of a reference trajectory and from the same device checkpoint: plain VBD
    from x~ with the FP64 convergence check every 1, 2, 4 or 8 sweeps; the history extrapolation x~ + (x* -
    x~) of the previous substep (the non-learned start); and the network's prediction, the last two behind
    the deferred gate; each iterates until the FP64 relative force residual is at most 1e-4, timed with the
    host clock around the whole substep. Finally a deployment run, in which each of the three paths continues
    from its own 1e-4 solves."""
    import numpy as np
    from spg_amd import learned, scenes

    t0 = time.perf_counter()
    train_set = scenes.cloth_batch(count=256, seed=7)
    f, t = learned.collect(train_set, steps=10, samples=2_000_000, rel_tol=1e-5, max_iterations=1500, check_every=10)
    collect_s = time.perf_counter() - t0
    # the learned start predicts x* - x~; a second network, trained on the same data, predicts the change of the
    # previous substep's offset on top of the history start (x~ + c + s MLP), for comparison
    t0 = time.perf_counter()
    net, losses = learned.train(f, t, epochs=12, batch=4096, lr=2e-3, seed=2, history_base=False)
    train_s = time.perf_counter() - t0
    net_hist, losses_hist = learned.train(f, t, epochs=12, batch=4096, lr=2e-3, seed=2, history_base=True)
    nets = {"learned": net, "learned_on_history": net_hist}
    lk = {k: launch[k] for k in ("lanes_per_vertex", "block_size", "lds_staging")}
    rows = learned.compare(scene, nets, steps=args.learned_steps, rel_tol=1e-4, max_iterations=3000, check_every=1,
                           smoothing=2, reference_tol=1e-5, reference_iterations=2000, cadences=(1, 2, 4, 8),
                           launch=lk)
    sm = learned.summarize(rows)
    drows = learned.deploy(scene, nets, steps=args.learned_steps, rel_tol=1e-4, check_every=1, smoothing=2, launch=lk)
    dsm = learned.summarize_deploy(drows)

    def clean(r):
        out = {}
        for k, v in r.items():
            if k == "plain_by_cadence":
                out[k] = {str(ce): {"iterations": c["iterations"], "rho": c["rho"], "ms": round(c["ms"], 4)}
                          for ce, c in v.items()}
            elif isinstance(v, float) and k.endswith("_ms"):
                out[k] = round(v, 4)
            elif isinstance(v, float) and not np.isfinite(v):
                out[k] = None
            else:
                out[k] = v  # residual ratios unrounded
        return out

    return {"metric": "wall time to a 1e-4 relative FP64 force residual per substep, learned warm start vs plain VBD",
            "workload": f"same cloth as above, {args.learned_steps} consecutive steps of a reference trajectory "
                        f"(plain VBD, 2000 iterations per step); from each state: plain VBD from x~ (checks every "
                        f"1/2/4/8 sweeps), the history start x~ + c (c = (x* - x~) of the previous substep), the "
                        f"learned start x~ + s MLP (the network predicts x* - x~) and, for comparison, x~ + c + s MLP' "
                        f"(a second network predicts the change of c), the starts behind the deferred gate, 2 "
                        f"one-ring smoothing passes of the network output, checks after every sweep; warm start "
                        f"(features, inference, gate residual) included",
            "plain_ms": sm["plain_ms"], "learned_ms": sm["learned_ms"], "time_ratio": sm["time_ratio"],
            "plain_iterations": sm["plain_iterations"], "learned_iterations": sm["learned_iterations"],
            "iteration_ratio": sm["iteration_ratio"],
            "plain_ms_by_check_cadence": {str(k): v for k, v in sm["plain_ms_by_cadence"].items()},
            "plain_iterations_by_check_cadence": {str(k): v for k, v in sm["plain_iterations_by_cadence"].items()},
            "best_plain_check_cadence": sm["best_plain_cadence"], "best_plain_ms": sm["best_plain_ms"],
            "time_ratio_vs_best_plain": sm["time_ratio_vs_best_plain"],
            "history_ms": sm["history_ms"], "history_iterations": sm["history_iterations"],
            "history_all_converged": sm["history_all_converged"],
            "time_ratio_vs_history": sm["time_ratio_vs_history"],
            "iteration_ratio_vs_history": sm["iteration_ratio_vs_history"],
            "learned_on_history": sm["learned_on_history"],
            "accepted": sm["accepted"], "fell_back": sm["fell_back"], "all_converged": sm["all_converged"],
            "deployment": {"what": "one context per path (plain, history, learned, learned on history) from the "
                                   "initial state, each continuing from its own 1e-4 solves "
                                   f"for {args.learned_steps} steps (checks after every sweep)", **dsm,
                           "per_step": [clean(r) for r in drows]},
            "per_step": [clean(r) for r in rows],
            "training": {"samples": int(f.shape[0]), "scenes": len(train_set), "collect_s": collect_s,
                         "train_s": train_s, "epochs": len(losses), "final_loss": float(losses[-1]),
                         "final_loss_change_target": float(losses_hist[-1]),
                         "target_variance": float(np.var(t, axis=0).sum())}}


def oracle_residual(scene, x, x_tilde):
    """The FP64 oracle's relative force residual |r(x)| / |r(x~)| for the GPU's state (the checker, in the CPU
    leg of the bench)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol

    o = ol.from_scene(scene)
    r = o.residual(x, x_tilde, scene.dt)
    r0 = o.residual(x_tilde, x_tilde, scene.dt)
    return {"oracle_fp64_rho": r / r0, "oracle_fp64_residual": r, "oracle_fp64_residual_at_inertial": r0}


def measured_traffic(launch, args):
    """HBM bytes per vertex-solve launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes of
    this workload (scaled by the same counters on a copy kernel of known size), for the launch shape
    the bench uses; None when that shape was not profiled."""
    path = os.path.join(ROOT, "profiles", "solveColorKernel_cloth512_traffic.json")
    if args.n != 512 or args.seed != 3 or not os.path.exists(path):
        return None, None
    with open(path) as f:
        t = json.load(f)
    for sh in t.get("shapes", []):
        if (sh["lanes"], sh["block"], sh["lds"]) == (launch["lanes_per_vertex"], launch["block_size"],
                                                     launch["lds_staging"]):
            return sh["traffic_bytes"], ("profiles/solveColorKernel_cloth512_traffic.json (rocprofv3 --pmc, same "
                                         "kernel and launch shape)")
    return None, None


def cpu_baseline(scene):
    """FP64 oracle (OpenMP restatement of VBD::step) on the host cores: repeated first steps of the same
    scene for a bounded sample (~10 s closed form on all threads, ~5 s on one thread, ~8 s with
    forward-mode dual numbers), colouring/incidence excluded (done before timing)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    import numpy as np

    cores = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    o = ol.from_scene(scene)
    ol.set_threads(cores)
    t0 = time.perf_counter()
    o.colors()  # first-call precomputation outside the timed region (vbd.cpp:14-18), timed on its own
    precompute_s = time.perf_counter() - t0
    v0 = np.zeros_like(scene.x)

    def timed(threads, iterations, mode, budget_s):
        """Repeated first steps from the initial state until about budget_s of CPU work."""
        ol.set_threads(threads)
        n, t = 0, 0.0
        while t < budget_s:
            o.set_state(scene.x, v0)
            t0 = time.perf_counter()
            o.step(scene.dt, 1, iterations, scene.gravity, mode=mode)
            t += time.perf_counter() - t0
            n += 1
        return scene.n * iterations * n / t, n, t

    cf, n_cf, t_cf = timed(cores, scene.iterations, ol.CLOSED_FORM, 10.0)
    cf1, n_cf1, t_cf1 = timed(1, 2, ol.CLOSED_FORM, 5.0)
    # AD mode (TinyAD-like cost: two full-element derivative passes per incidence)
    ad, n_ad, t_ad = timed(cores, 1, ol.AD, 8.0)
    return {"value": cf, "unit": "vertex-updates/s", "cores": cores, "kind": "port", "precompute_s": precompute_s,
            "sample": f"{n_cf} first steps ({scene.iterations} iterations each) of the same {scene.n}-vertex cloth, "
                      f"FP64 closed-form derivatives, {t_cf:.1f} s",
            "one_core_value": cf1,
            "one_core_sample": f"{n_cf1} first steps of 2 iterations on 1 thread, closed form, {t_cf1:.1f} s",
            "ad_mode_value": ad,
            "ad_mode_sample": f"{n_ad} first steps of 1 iteration on {cores} threads, FP64 forward-mode dual numbers, "
                              f"{t_ad:.1f} s"}


if __name__ == "__main__":
    main()
