#!/usr/bin/env python
"""Throughput of the implicit elastodynamics step on one MI355X.

Workload (default): the embedded higher-order scene - a 24^3 Hex20 background grid over
[-1, 1]^3 cut by a hollow cylinder (r in [0.6, 0.8], |z| <= 0.8), 8-point reduced quadrature
per cell, Stable Neo-Hookean E = 5e5, nu = 0.45, rho = 1000, bottom cap fixed, gravity,
dt = 1/200, 20 Backward-Euler steps, Newton with backtracking line search to a 1e-6 relative
residual (|F| <= dt * 1e-6 * |M a_rep|), Schwarz-preconditioned CG (rel. 1e-5, <= 500
iterations), all in FP64. Synthetic, seeded inputs.

A "step" is one time step. The timed region runs K steps from the rest state after W
untimed warm-up steps (also from rest), bracketed by device synchronizations.

Printed JSON line: the metric, the roofline of the dominant kernel (the fused CG
system-matrix kernel: direction update + SpMV + p.Ap) measured with kernel-timestamp HIP events
inside the library on the solver stream during a second pass over the same steps (the timed
pass replays captured CG batch graphs, which per-kernel events would break up), and the CPU
baseline: the FP64 oracle (the C++ restatement of
the reference's CPU path, OpenMP over element colors) on a bounded sample of the same
workload, with the displacement error of the GPU run against it.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--config", type=int, default=2, choices=[1, 2, 3, 4, 5],
                   help="1: Tet10 cube, 2: embedded Hex20 cylinder, 3: embedded Hex27 cylinder (48^3), "
                        "4: config 2 with the learned coarse correction, 5: batched independent scenes")
    p.add_argument("--harvest-steps", type=int, default=10, help="config 4: plain steps that harvest snapshots")
    p.add_argument("--train-steps", type=int, default=600, help="config 4: Adam steps")
    p.add_argument("--scenes", type=int, default=64, help="config 5: number of scenes")
    p.add_argument("--cpu-steps", type=int, default=10,
                   help="oracle steps timed for the CPU baseline (0: skip); 10 steps of the default workload "
                        "are 10-15 s of CPU work on 16 threads")
    p.add_argument("--no-profile", action="store_true", help="do not bracket the CG kernel with events")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="run a plain 256 MiB device copy 3x first (reference kernel for HBM counters)")
    p.add_argument("--mixed", action="store_true",
                   help="FP32 element tangents / Jacobian SpMV / Schwarz operator inside the FP64 Newton refinement")
    p.add_argument("--mixed-refined", action="store_true",
                   help="FP64 tangents and Jacobian, CG on FP32 copies inside FP64 iterative refinement")
    p.add_argument("--no-autotune", action="store_true", help="keep the heuristic SpMV lanes per node")
    p.add_argument("--stiffness-damping", type=float, default=0.0,
                   help="Material stiffness_damping_coefficient (the reference's cylinder shell scene uses 0.05)")
    p.add_argument("--mass-damping", type=float, default=0.0, help="Material mass_damping_coefficient")
    p.add_argument("--quadrature", choices=["sampled", "embedded"], default="sampled",
                   help="config 2: cut-cell rules by rejection sampling (default) or built the reference's way "
                        "(embedded tetrahedral mesh, clipping, moment fitting)")
    p.add_argument("--reorder", action="store_true",
                   help="reorder the scene's mesh (reverse Cuthill-McKee, reorder_mesh_par) before building the model")
    return p.parse_args()


def cg_kernel_bytes(model, value_bytes: int = 8) -> int:
    """Algorithmic bytes of one launch of the CG system-matrix kernel: every matrix value
    once (8 B, or 4 B in mixed precision), one node-level column index per 3x3 block (4 B per
    9 values), row offsets, node offsets, and four FP64 DOF vectors (read z, read p_old, write
    p_new, write Ap)."""
    nnz, ndof = model.nnz, model.ndof
    nv = ndof // 3
    return nnz * value_bytes + (nnz // 9) * 4 + (ndof + 1) * 4 + (nv + 1) * 4 + 4 * ndof * 8


def common_line(args, metric, value, ms, dtype, config, **extra):
    line = {"metric": metric, "value": round(value, 3), "unit": "timesteps/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic (seeded scene generator)", "config": config}
    line.update(extra)
    return line


def bench_learned(args):
    """Config 4: the embedded Hex20 cylinder with the learned coarse correction. One simulation
    harvests snapshots during its first plain steps, trains the network and continues with the
    correction enabled; a second simulation runs the same steps plain. Both time the same K
    steps after the harvest phase."""
    import numpy as np
    import torch

    from higher_order_embedded_fem_amd import fem, scenes

    sc = scenes.config_scene(2)
    model = fem.ElasticityModel(sc)
    sims = {name: fem.NewtonCgIntegrator(model, sc, newton_tolerance_factor=1e-6, mixed_precision=args.mode)
            for name in ("plain", "learned")}
    for sim in sims.values():
        if not args.no_autotune:
            sim.autotune()
    lc = fem.LearnedCoarseCorrection(sims["learned"], num_aggregates=512, seed=0)
    state = {}
    for name, sim in sims.items():
        u = torch.zeros(sc.ndof, dtype=torch.float64, device="cuda")
        v = torch.zeros_like(u)
        if name == "learned":
            lc.harvest(True, capacity=max(64, 40 * args.harvest_steps), every=16)
        for k in range(args.harvest_steps):
            sim.step(k * sc.dt, sc.dt, u, v)
        state[name] = (u, v)
    lc.harvest(False)
    t0 = time.perf_counter()
    l0, l1 = lc.train(steps=args.train_steps, learning_rate=3e-3, batch=8)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    lc.enable(True)
    res = {}
    for name, sim in sims.items():
        u, v = state[name]
        cg = newton = 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.harvest_steps, args.harvest_steps + args.steps):
            st = sim.step(k * sc.dt, sc.dt, u, v)
            cg += st["cg_iterations"]
            newton += st["newton_iterations"]
        torch.cuda.synchronize()
        res[name] = {"seconds": time.perf_counter() - t0, "cg": cg, "newton": newton, "u": u.cpu().numpy()}
        print(f"[bench] {name}: {args.steps} steps in {res[name]['seconds']:.3f}s, cg {cg}, newton {newton}",
              file=sys.stderr, flush=True)
    info = lc.info()
    d = res["learned"]["u"] - res["plain"]["u"]
    value = args.steps / res["learned"]["seconds"]
    line = common_line(
        args, "timesteps/s at 1e-6 rel Newton residual", value, 1e3 * res["learned"]["seconds"] / args.steps,
        "f32+f64" if args.mixed else "f64",
        {"workload": sc.name + " + learned coarse correction", "dofs": sc.ndof, "nnz": model.nnz,
         "harvest_steps": args.harvest_steps, "train_steps": args.train_steps, "aggregates": info["num_aggregates"],
         "precision": "mixed" if args.mixed else "fp64", "parallelism": "single GPU"},
        learned={"plain_value": round(args.steps / res["plain"]["seconds"], 3),
                 "speedup_vs_plain": round(res["plain"]["seconds"] / res["learned"]["seconds"], 4),
                 "cg_iterations_per_step": res["learned"]["cg"] / args.steps,
                 "cg_iterations_per_step_plain": res["plain"]["cg"] / args.steps,
                 "fallbacks": info["fallbacks"], "learned_solves": info["learned_solves"],
                 "harvested_pairs": info["harvested_pairs"], "train_seconds": round(t_train, 3),
                 "loss_before_after": [l0, l1],
                 "u_rel_l2_vs_plain": float(np.linalg.norm(d) / max(np.linalg.norm(res["plain"]["u"]), 1e-300))},
        roofline=None, cpu_baseline=None)
    print(json.dumps(line))


def bench_batched(args):
    """Config 5: independent scenes of uneven size stepped together (one stream and one
    captured CG batch graph per scene, host threads overlap the scenes)."""
    import numpy as np
    import torch

    from higher_order_embedded_fem_amd import batch

    t0 = time.time()
    scs = batch.batched_scenes(args.scenes)
    stepper = batch.BatchedStepper(scs, newton_tolerance_factor=1e-6, mixed_precision=args.mode,
                                   autotune=not args.no_autotune)
    print(f"[bench] {len(scs)} scenes, {sum(s.ndof for s in scs)} dofs in total, setup {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    if args.warmup:
        stepper.run(args.warmup)
    stepper.reset()
    r = stepper.run(args.steps)
    print(f"[bench] {r['scene_steps']} scene steps in {r['wall_s']:.3f}s", file=sys.stderr, flush=True)
    line = common_line(
        args, "aggregate timesteps/s at 1e-6 rel Newton residual", r["aggregate_timesteps_per_s"],
        1e3 * r["wall_s"] / args.steps, "f32+f64" if args.mixed else "f64",
        {"workload": f"{len(scs)} independent scenes (Tet10 cubes and Hex20 cylinders, DOFs log-uniform in [2e3, 2e5])",
         "total_dofs": r["total_dofs"], "precision": "mixed" if args.mixed else "fp64", "parallelism": "single GPU"},
        batched={"p50_step_latency_ms": round(r["p50_step_latency_ms"], 3),
                 "p99_step_latency_ms": round(r["p99_step_latency_ms"], 3),
                 "newton_iterations": r["newton_iterations"], "cg_iterations": r["cg_iterations"]},
        roofline=None, cpu_baseline=None)
    print(json.dumps(line))


def main():
    args = parse()
    args.mode = 2 if args.mixed_refined else (1 if args.mixed else 0)
    args.mixed = args.mode != 0
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (asked for {args.gpus}); the library has no multi-GPU path.",
              file=sys.stderr)
        sys.exit(2)
    if args.config in (4, 5):
        import torch

        assert torch.cuda.is_available(), "bench.py needs a HIP device"
        (bench_learned if args.config == 4 else bench_batched)(args)
        return
    import ctypes as C

    import numpy as np
    import torch

    from higher_order_embedded_fem_amd import _lib, fem, scenes

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    if args.calibrate_copy:
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")  # 256 MiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print("[bench] calibration copy: 3 x 268435456 bytes read + written", file=sys.stderr, flush=True)
        del src, dst
    t_setup = time.time()
    sc = (scenes.config_scene(args.config) if args.quadrature == "sampled"
          else scenes.cylinder_scene(24, scenes.HEX20, 8, quadrature="embedded"))
    if args.reorder:
        sc = scenes.reorder(sc)
    model = fem.ElasticityModel(sc)
    sim = fem.NewtonCgIntegrator(model, sc, newton_tolerance_factor=1e-6, mixed_precision=args.mode,
                                 mass_damping=args.mass_damping, stiffness_damping=args.stiffness_damping)
    L = _lib.load()
    lanes = None if args.no_autotune else sim.autotune()
    print(f"[bench] {sc.name}: {sc.num_elements} elements, {sc.ndof} dofs, nnz {model.nnz}, "
          f"setup {time.time() - t_setup:.1f}s, SpMV lanes per node (J, P) {lanes}", file=sys.stderr, flush=True)
    if lanes is not None:
        print(f"[bench] autotune us per CG iteration, rows J lanes 8/16/32, columns P lanes 8/16/32:\n"
              f"{np.array2string(sim.autotune_us, precision=2)}", file=sys.stderr, flush=True)
    precision = {0: "fp64", 1: "mixed", 2: "mixed-refined"}[args.mode]

    def run(steps, record=None):
        u = torch.zeros(sc.ndof, dtype=torch.float64, device="cuda")
        v = torch.zeros_like(u)
        tot = {"newton_iterations": 0, "cg_iterations": 0, "substeps": 0, "res_over_tol": 0.0}
        for k in range(steps):
            st = sim.step(k * sc.dt, sc.dt, u, v)
            for key in ("newton_iterations", "cg_iterations", "substeps"):
                tot[key] += st[key]
            tot["res_over_tol"] = max(tot["res_over_tol"], st["final_residual"] / st["tolerance"])
            if record is not None and k + 1 == record[0]:
                record[1] = u.cpu().numpy().copy()
        return u, tot

    run(args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    u, tot = run(args.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    # second pass over the same steps with the CG system-matrix kernel timed by kernel-timestamp
    # events (per-kernel launches instead of the replayed batch graphs)
    full_n, full_ms, all_n, all_ms = C.c_int64(), C.c_double(), C.c_int64(), C.c_double()
    prof_elapsed = None
    if not args.no_profile:
        L.hoefem_profile_enable(model.h, 1)
        torch.cuda.synchronize()
        p0 = time.perf_counter()
        run(args.steps)
        torch.cuda.synchronize()
        prof_elapsed = time.perf_counter() - p0
        L.hoefem_profile_read(model.h, C.byref(full_n), C.byref(full_ms), C.byref(all_n), C.byref(all_ms))
        L.hoefem_profile_enable(model.h, 0)
    value = args.steps / elapsed
    last_res = round(tot["res_over_tol"], 4)
    print(f"[bench] {args.steps} steps in {elapsed:.3f}s: {value:.2f} steps/s; newton {tot['newton_iterations']}, "
          f"cg {tot['cg_iterations']}", file=sys.stderr, flush=True)

    roofline = None
    if full_n.value > 0:
        avg_ms = full_ms.value / full_n.value
        nbytes = cg_kernel_bytes(model, 4 if args.mixed else 8)
        achieved = nbytes / (avg_ms * 1e-3) / 1e9
        traffic = None
        tpath = os.path.join(ROOT, "profiles", f"traffic_cg_matrix_kernel_{sc.name}_{precision}.json")
        if os.path.exists(tpath):
            traffic = json.load(open(tpath)).get("hbm_bytes_per_launch")
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "kernel": "k_cg_a (CG system-matrix kernel: p = z + beta p, Ap = J p, p.Ap)",
                    "bytes_per_launch": nbytes, "avg_launch_us": round(avg_ms * 1e3, 3),
                    "launches": full_n.value, "all_launches_avg_us": round(all_ms.value / max(all_n.value, 1) * 1e3, 3),
                    "all_launches": all_n.value,
                    "measured_in": f"a second pass over the same {args.steps} steps with per-kernel timing "
                                   f"({1e3 * prof_elapsed / args.steps:.3f} ms/step)"}

    cpu = None
    err = None
    if args.cpu_steps > 0:
        import oracle as O

        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        rec = [args.cpu_steps, None]
        run(args.cpu_steps, record=rec)  # untimed GPU run for the error against the oracle
        om = O.Model(sc)
        osim = O.Sim(om, newton_tolerance_factor=1e-6, num_threads=threads, mass_damping=args.mass_damping,
                     stiffness_damping=args.stiffness_damping)
        uo = np.zeros(sc.ndof)
        vo = np.zeros(sc.ndof)
        c0 = time.perf_counter()
        for k in range(args.cpu_steps):
            osim.step(k * sc.dt, sc.dt, uo, vo)
        c1 = time.perf_counter()
        cpu = {"value": round(args.cpu_steps / (c1 - c0), 4), "unit": "timesteps/s", "cores": threads, "kind": "port",
               "sample": f"first {args.cpu_steps} of the {sc.steps} steps of {sc.name} (same settings), "
                         f"C++ FP64 oracle with OpenMP over element colors"}
        rp, col = om.pattern()
        Mv = om.mass(sc.density)
        d = rec[1] - uo
        err = {"u_rel_l2": float(np.linalg.norm(d) / max(np.linalg.norm(uo), 1e-300)),
               "u_rel_mass_norm": float(np.sqrt(max(d @ O.spmv(rp, col, Mv, d), 0.0))
                                        / max(np.sqrt(uo @ O.spmv(rp, col, Mv, uo)), 1e-300)),
               "steps": args.cpu_steps}
        print(f"[bench] cpu baseline {cpu['value']} steps/s on {threads} threads; gpu-vs-oracle after "
              f"{args.cpu_steps} steps: {err}", file=sys.stderr, flush=True)

    line = {
        "metric": "timesteps/s at 1e-6 rel Newton residual",
        "value": round(value, 3),
        "unit": "timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if args.mixed else "f64",
        "data": "synthetic (seeded scene generator)",
        "config": {"workload": sc.name, "element": scenes.ELEMENT_NAMES[sc.element_type], "cells": sc.num_elements,
                   "dofs": sc.ndof, "nnz": model.nnz, "dt": sc.dt, "newton_rel_tol": 1e-6, "cg_rel_tol": 1e-5,
                   "preconditioner": "schwarz", "precision": precision, "spmv_lanes_per_node": lanes,
                   "parallelism": "single GPU",
                   **({"mass_damping": args.mass_damping, "stiffness_damping": args.stiffness_damping}
                      if args.mass_damping or args.stiffness_damping else {})},
        "newton_iterations_per_step": tot["newton_iterations"] / args.steps,
        "cg_iterations_per_step": tot["cg_iterations"] / args.steps,
        "error_vs_oracle": err,
        "final_residual_over_tolerance": last_res,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
