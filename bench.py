"""Benchmark: AVBD step throughput on the 21^3-vertex soft block.

Workload (one step = the whole AVBD frame: hash, pairs, narrow phase,
persistence map, inertial predict, contact adjacency, 20 iterations of the
static-colored primal sweep + dual, BDF1 finalize): a 20x20x20-cell tet block
(9261 vertices, 48000 Kuhn tets, 1 m), Stable Neo-Hookean E = 1e5 Pa,
nu = 0.4, 1000 kg/m^3, dropped from 0.5 m with seeded N(0, 0.1) m/s velocity
onto the ground, dt = 1/60.  Inputs are resident on the GPU; the timed region
is K steps bracketed by device synchronisation.

Prints one JSON line: value = vertex-iters/s (dynamic vertices x iterations
x steps / wall time), the step rate when every frame iterates to residual
1e-4 (cap 200 sweeps), the primal color-sweep roofline measured live with
HIP events, the oracle CPU baseline on a bounded sample, and the learned
warm start against the plain one (frames iterated to residual 1e-3 -- the
FP32 iterate of this block floors near 5e-4 -- and the residual after the
fixed 20 sweeps), with the MLP forward's matrix-core roofline.

--workload cloth / batched run the 256^2 cloth drape or the 4096 batched
rigid scenes instead (reported the same way; the default is the soft block).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "avbd-metal_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0
METRIC = "vertex-iters/s + ms/step to residual 1e-4; batched+learned vs plain AVBD"


def traffic_from_profiles(name):
    """HBM bytes per primal launch from a committed rocprofv3 PMC summary, when present."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return float(json.load(f)["hbm_bytes_per_launch"])
    except Exception:
        return None


def cpu_baseline(scene, unit, target_s=10.0):
    """Oracle (C++ FP32 restatement of the reference CPU path) on the host:
    one core (the serial Gauss-Seidel, the reported value), then every color
    split over the host's cores (identical results), each on a bounded
    sample of about target_s seconds."""
    import oracle

    def rate(threads, budget):
        o = oracle.OracleSolver(scene)
        o.set_threads(threads)
        o.step()  # warm-up step (first contact detection, caches)
        t0 = time.perf_counter()
        o.step()
        one = time.perf_counter() - t0
        n = max(1, min(500, int(budget / max(one, 1e-3))))
        t0 = time.perf_counter()
        o.step(n)
        dt = time.perf_counter() - t0
        return scene.numDynamic * scene.settings.iterations * n / dt, n, dt

    v1, n1, dt1 = rate(1, target_s)
    cores = max(1, min(16, os.cpu_count() or 1))
    out = {"value": v1, "unit": unit, "cores": 1, "kind": "port",
           "sample": f"{n1} steps of the same workload after 2 warm-up steps (about {target_s:.0f} s of work), "
                     f"serial oracle FP32, {dt1:.2f} s"}
    if cores > 1:
        vn, nn, dtn = rate(cores, target_s / 2)
        out["all_cores"] = {"value": vn, "cores": cores,
                            "sample": f"{nn} steps, each color's bodies split over {cores} OpenMP threads, "
                                      f"{dtn:.2f} s"}
    return out


def primal_roofline(solver, pmc):
    """Algorithmic bytes of one primal color launch over its graph-replayed
    duration (HIP events on the solver stream) against the HBM peak."""
    tp = solver.timePrimal(reps=50)
    st = solver.stats()
    bytes_per_launch = tp["sweep_bytes"] / tp["launches_per_sweep"]
    achieved = bytes_per_launch / (tp["avg_launch_us"] * 1e-6) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(pmc) if pmc else None,
            "kernel": f"k_primal<{st.lanes_per_body}>", "avg_launch_us": round(tp["avg_launch_us"], 3),
            "algorithmic_bytes_per_launch": round(bytes_per_launch)}


def learned_comparison(args):
    """Learned vs plain warm start: training pairs from the plain solver on
    other seeds of the --train-cells block (every frame iterated to residual
    2e-4 or 200 sweeps), GPU training, then frames of the --eval-cells block
    (default: the metric's scene) iterated to `tol` with and without the
    learned start (plain fallback active), the residual after the fixed
    iteration count, and the MLP forward roofline on the evaluated rows."""
    import numpy as np
    from avbd_amd import GPUSolver, WarmStartMLP, demos
    from avbd_amd.learned import collect_samples

    t0 = time.perf_counter()
    train = [demos.soft_block_scene(cells=args.train_cells, seed=10 + k) for k in range(args.train_seeds)]
    X, T, label_its = collect_samples(train, args.train_frames)
    t_collect = time.perf_counter() - t0
    m = WarmStartMLP(seed=1)
    t0 = time.perf_counter()
    losses = m.train(X, T, epochs=args.epochs, batch=4096, lr=1e-3, seed=2)
    t_train = time.perf_counter() - t0

    eval_cells = args.eval_cells or args.cells
    eval_scene = demos.soft_block_scene(cells=eval_cells)
    nv = eval_scene.numDynamic

    def run(learned):
        s = eval_scene
        g = GPUSolver(s)
        if learned:
            g.setWarmStart(m, smooth=2)
        its, acc, reached = [], [], 0
        t0 = time.perf_counter()
        for _ in range(args.learned_frames):
            n, r = g.stepToResidual(args.learned_tol, 200, 5)
            its.append(n)
            reached += r <= args.learned_tol
            acc.append(g.stats().ws_accepted)
        wall = time.perf_counter() - t0
        g2 = GPUSolver(s)
        if learned:
            g2.setWarmStart(m, smooth=2)
        res = []
        for _ in range(args.learned_frames):
            g2.step()
            res.append(g2.residual())
        return {"iterations_mean": round(float(np.mean(its)), 2), "ms_per_step": round(wall * 1000 / len(its), 4),
                "reached_fraction": reached / len(its),
                "learned_start_kept": round(float(np.mean(acc)), 3) if learned else None,
                "residual_after_fixed_iterations_mean": float(np.mean(res)),
                "residual_after_fixed_iterations_median": float(np.median(res))}

    plain, learned = run(False), run(True)
    us = m.timeForward(nv, reps=50)
    flops = 2.0 * nv * (24 * 128 + 128 * 128 + 128 * 3)
    tflops = flops / (us * 1e-6) / 1e12
    out = {"evaluated_on": f"soft block {eval_cells}^3 cells ({nv} vertices, {len(eval_scene.tets)} tets)",
           "trained_on": f"soft block {args.train_cells}^3 cells, seeds 10..{9 + args.train_seeds}",
           "tolerance": args.learned_tol, "frames": args.learned_frames, "max_iterations": 200,
           "training": {"samples": int(len(X)), "scenes": f"{args.train_seeds} seeds x {args.train_frames} frames",
                        "label_sweeps_mean": round(label_its, 1), "collect_s": round(t_collect, 2),
                        "epochs": args.epochs, "train_s": round(t_train, 2),
                        "loss_first_last": [float(losses[0]), float(losses[-1])], "target_variance": float(T.var())},
           "plain": plain, "learned": learned,
           "step_rate_ratio": round(plain["ms_per_step"] / learned["ms_per_step"], 3),
           "iteration_ratio": round(plain["iterations_mean"] / learned["iterations_mean"], 3)}
    roof = {"bound": "mfma", "achieved": round(tflops, 2), "peak": 2500.0, "unit": "TFLOP/s",
            "frac": round(tflops / 2500.0, 5), "traffic": None, "kernel": "k_mlp_forward",
            "avg_launch_us": round(us, 3), "rows": nv, "flops_per_row": 2 * (24 * 128 + 128 * 128 + 128 * 3)}
    return out, roof


def batched_learned_comparison(args):
    """Batched + learned vs plain AVBD on the same uneven soft blocks: the
    plain path steps each block with its own solver (plain warm start), the
    batched paths step all blocks in one solver, without and with the learned
    start (trained on other draws of the same generator).  Frames iterate to
    residual `tol` (per solver: each block for the plain path, all blocks
    together for the batched paths), cap 200 sweeps."""
    import numpy as np
    from avbd_amd import GPUSolver, WarmStartMLP, demos
    from avbd_amd.learned import collect_samples

    n, frames, tol = args.batch_scenes, args.batch_frames, args.learned_tol
    t0 = time.perf_counter()
    train = [demos.batched_soft_scene(n, seed_base=(6 << 32) + 1000 * k) for k in range(2)]
    X, T, _ = collect_samples(train, args.train_frames)
    m = WarmStartMLP(seed=1)
    m.train(X, T, epochs=args.epochs, batch=4096, lr=1e-3, seed=2)
    t_train = time.perf_counter() - t0

    def frames_to_tol(solvers):
        its = []
        t0 = time.perf_counter()
        for _ in range(frames):
            for g in solvers:
                k, _ = g.stepToResidual(tol, 200, 5)
                its.append(k)
        return (time.perf_counter() - t0) * 1000.0 / frames, float(np.mean(its))

    sep = [GPUSolver(sc) for sc in demos.batched_soft_scene(n, separate=True)]
    ms_plain, it_plain = frames_to_tol(sep)
    for g in sep:
        g.close()
    scene = demos.batched_soft_scene(n)
    gb = GPUSolver(scene)
    ms_batched, it_batched = frames_to_tol([gb])
    gl = GPUSolver(scene)
    gl.setWarmStart(m, smooth=2)
    ms_learned, it_learned = frames_to_tol([gl])
    return {"scenes": n, "vertices": scene.numDynamic, "tets": len(scene.tets), "frames": frames,
            "tolerance": tol, "train_s": round(t_train, 2), "train_samples": int(len(X)),
            "plain_per_scene_solvers": {"ms_per_step": round(ms_plain, 3), "iterations_mean": round(it_plain, 2)},
            "batched_plain": {"ms_per_step": round(ms_batched, 3), "iterations_mean": round(it_batched, 2)},
            "batched_learned": {"ms_per_step": round(ms_learned, 3), "iterations_mean": round(it_learned, 2)},
            "step_rate_ratio_batched_learned_vs_plain": round(ms_plain / ms_learned, 2),
            "step_rate_ratio_batching_only": round(ms_plain / ms_batched, 2),
            "step_rate_ratio_learned_only": round(ms_batched / ms_learned, 3)}


def rigid_learned_comparison(args, scene):
    """Batched rigid scenes with the learned start of the rigid bodies (a
    6-output network trained on other draws of the generator, labels from the
    plain solver at 100 sweeps) against the plain guess: the residual after
    the fixed iteration count over `frames` frames."""
    import numpy as np
    from avbd_amd import GPUSolver, WarmStartMLP, demos
    from avbd_amd.learned import collect_samples

    t0 = time.perf_counter()
    train = [demos.batched_scene(64, seed_base=(7 << 32) + 1000 * k)[0] for k in range(3)]
    X, T, _ = collect_samples(train, 60, tol=1e-4, max_iterations=100, check_every=10, rigid=True)
    m = WarmStartMLP(seed=1, outputs=6)
    losses = m.train(X, T, epochs=args.epochs, batch=4096, lr=1e-3, seed=2)
    t_train = time.perf_counter() - t0

    def run(learned):
        g = GPUSolver(scene)
        if learned:
            g.setWarmStart(rigid=m)
        res, kept = [], []
        t0 = time.perf_counter()
        for _ in range(args.learned_frames):
            g.step()
            res.append(g.residual())
            kept.append(g.stats().ws_accepted)
        wall = time.perf_counter() - t0
        return {"residual_after_fixed_iterations_median": float(np.median(res)),
                "residual_after_fixed_iterations_mean": float(np.mean(res)),
                "ms_per_step_with_residual_readback": round(wall * 1000 / len(res), 4),
                "learned_start_kept": round(float(np.mean(kept)), 3) if learned else None}

    plain, learned = run(False), run(True)
    return {"trained_on": "3 x 64 batched rigid scenes (other seeds), 60 frames, labels at 100 sweeps",
            "train_samples": int(len(X)), "train_s": round(t_train, 2),
            "loss_first_last": [float(losses[0]), float(losses[-1])], "frames": args.learned_frames,
            "plain": plain, "learned": learned,
            "residual_ratio_median": round(plain["residual_after_fixed_iterations_median"] /
                                           max(learned["residual_after_fixed_iterations_median"], 1e-30), 3)}


def build(workload, cells):
    from avbd_amd import demos

    if workload == "cloth":
        s = demos.cloth_scene()
        desc = (f"cloth drape 256^2 ({s.numDynamic} vertices, {len(s.tris)} triangles, {len(s.springs)} "
                f"structural springs) onto a fixed r = 0.5 m sphere, {s.settings.iterations} iterations, dt 1/60")
        return s, desc, "vertex-iters/s", None
    if workload == "batched":
        s, ranges = demos.batched_scene()
        desc = (f"4096 batched uneven rigid scenes ({s.numDynamic} bodies, {len(s.joints)} hinges) in one solver, "
                f"{s.settings.iterations} iterations, dt 1/60, contact-aware dynamic coloring")
        return s, desc, "body-iters/s", None
    s = demos.soft_block_scene(cells=cells)
    desc = (f"soft block {cells}^3 cells, {s.numDynamic} vertices, {len(s.tets)} SNH tets, "
            f"{s.settings.iterations} iterations, dt 1/60, static coloring")
    return s, desc, "vertex-iters/s", f"primal_soft_block_{cells + 1}_pmc.json"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--cells", type=int, default=20)
    ap.add_argument("--workload", choices=["soft_block", "cloth", "batched"], default="soft_block")
    ap.add_argument("--residual-steps", type=int, default=30, help="frames iterated to residual 1e-4")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-vs-plain warm start comparison")
    ap.add_argument("--learned-tol", type=float, default=1e-3)
    ap.add_argument("--learned-frames", type=int, default=60)
    ap.add_argument("--train-seeds", type=int, default=3)
    ap.add_argument("--train-frames", type=int, default=60)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--train-cells", type=int, default=20, help="block of the learned start's training data")
    ap.add_argument("--eval-cells", type=int, default=0, help="block it is evaluated on (0 = --cells)")
    ap.add_argument("--batch-scenes", type=int, default=64, help="uneven soft blocks of the batched comparison")
    ap.add_argument("--batch-frames", type=int, default=30)
    ap.add_argument("--scale-cells", type=int, default=55, help="block of the at-scale primal roofline (0 = skip)")
    ap.add_argument("--lanes", type=int, default=0, help="lanes per body (0 = autotune)")
    ap.add_argument("--sweep", type=int, default=0, help="1 per-color launches, 2 persistent (0 = autotune)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this backend runs on one GPU; --gpus {args.gpus} is not supported "
              f"(independent scenes would shard per GPU with no per-step exchange)", file=sys.stderr)
        sys.exit(2)

    from avbd_amd import GPUSolver

    scene, desc, unit, pmc = build(args.workload, args.cells)
    solver = GPUSolver(scene, lanesPerBody=args.lanes, sweepMode=args.sweep)
    st = solver.stats()
    for _ in range(args.warmup):
        solver.submitStep()
    solver.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        solver.submitStep()
    solver.synchronize()
    wall = time.perf_counter() - t0
    iters = scene.settings.iterations
    nv = scene.numDynamic
    value = nv * iters * args.steps / wall
    ms_step = wall * 1000.0 / args.steps
    residual = solver.residual()
    st = solver.stats()

    # the second half of the metric: frames iterated until residual <= 1e-4 (cap 200)
    res = None
    if args.residual_steps > 0:
        its, reached = [], 0
        t0 = time.perf_counter()
        for _ in range(args.residual_steps):
            n, r = solver.stepToResidual(1e-4, 200, 5)
            its.append(n)
            reached += r <= 1e-4
        rwall = time.perf_counter() - t0
        res = {"tolerance": 1e-4, "max_iterations": 200, "check_every": 5, "steps": args.residual_steps,
               "ms_per_step": round(rwall * 1000.0 / args.residual_steps, 4),
               "iterations_mean": round(sum(its) / len(its), 2), "reached_fraction": reached / len(its)}

    out = {"metric": METRIC, "value": round(value, 1), "unit": unit, "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded SplitMix64 scene)",
           "config": {"workload": desc, "dynamic_bodies": nv, "iterations": iters,
                      "colors": st.num_colors, "lanes_per_body": st.lanes_per_body,
                      "sweep_mode": "persistent" if st.sweep_mode == 2 else "per-color launches",
                      "end_residual": residual, "residual_1e-4": res}}
    if args.workload == "cloth":
        out["config"].update({"soft_contacts_last_frame": st.num_soft, "dat_pairs_last_frame": st.dat_pairs})

    # roofline of the dominant kernel (per-color primal sweep), HIP events on the solver stream
    if args.workload != "batched":
        out["roofline"] = primal_roofline(solver, pmc)
    # the same kernel where a color fills the chip: the 1M-tet block (56^3 vertices)
    if args.workload == "soft_block" and args.scale_cells > 0 and args.scale_cells != args.cells:
        from avbd_amd import demos as _demos
        big = _demos.soft_block_scene(cells=args.scale_cells)
        t_create = time.perf_counter()
        gb = GPUSolver(big, lanesPerBody=args.lanes, sweepMode=args.sweep)
        create_ms = (time.perf_counter() - t_create) * 1e3
        gb.step(3)
        gb.synchronize()
        roof = primal_roofline(gb, f"primal_soft_block_{args.scale_cells + 1}_pmc.json")
        roof["workload"] = (f"soft block {args.scale_cells}^3 cells ({big.numDynamic} vertices, {len(big.tets)} tets), "
                            f"the same kernel at the 1M-tet size")
        roof["create_ms"] = round(create_ms, 1)   # scene packing + upload precompute + coloring + autotune
        out["roofline_at_scale"] = roof
        gb.close()
    if args.workload == "soft_block" and not args.no_learned:
        cmp, roof = learned_comparison(args)
        out["config"]["learned_vs_plain"] = cmp
        out["roofline_mlp"] = roof
        out["config"]["batched_learned_vs_plain"] = batched_learned_comparison(args)
    if args.workload == "batched" and not args.no_learned:
        out["config"]["rigid_learned_vs_plain"] = rigid_learned_comparison(args, scene)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scene, unit)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
