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
import csv
import glob
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
COPY_BYTES = 32 * 1024 * 1024 * 8  # the calibration copy: 256 MiB read and 256 MiB written


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--config", type=int, default=2, choices=[1, 2, 3, 4, 5],
                   help="1: Tet10 cube, 2: embedded Hex20 cylinder, 3: embedded Hex27 cylinder (48^3), "
                        "4: config 2 with the learned coarse correction, 5: batched independent scenes")
    p.add_argument("--harvest-steps", type=int, default=50, help="config 4: plain steps that harvest snapshots")
    p.add_argument("--eval-steps", type=int, default=100, help="config 4: steps timed plain and learned")
    p.add_argument("--eval-seconds", type=float, default=120.0,
                   help="config 4: wall-time cap of each evaluation run (the rate is over the steps done)")
    p.add_argument("--seed", type=int, default=7, help="config 4: seed of the pull loads")
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
    p.add_argument("--mixed-f64-tangents", action="store_true",
                   help="as --mixed, but FP64 element tangents, each Jacobian entry rounded to FP32 once")
    p.add_argument("--no-autotune", action="store_true", help="keep the heuristic SpMV lanes per node")
    p.add_argument("--stiffness-damping", type=float, default=0.0,
                   help="Material stiffness_damping_coefficient (the reference's cylinder shell scene uses 0.05)")
    p.add_argument("--mass-damping", type=float, default=0.0, help="Material mass_damping_coefficient")
    p.add_argument("--quadrature", choices=["sampled", "embedded"], default="sampled",
                   help="config 2: cut-cell rules by rejection sampling (default) or built the reference's way "
                        "(embedded tetrahedral mesh, clipping, moment fitting)")
    p.add_argument("--reorder", action="store_true",
                   help="reorder the scene's mesh (reverse Cuthill-McKee, reorder_mesh_par) before building the model")
    p.add_argument("--no-traffic", action="store_true",
                   help="skip the two rocprofv3 counter runs (FETCH_SIZE, WRITE_SIZE) that measure HBM traffic")
    p.add_argument("--traffic-probe", action="store_true",
                   help="internal: the short run profiled by the counter runs (autotune trials, calibration copy, "
                        "one step)")
    return p.parse_args()


def cg_kernel_bytes(model, value_bytes: int = 8) -> int:
    """Algorithmic bytes of one launch of the CG system-matrix kernel: every matrix value
    once (8 B, or 4 B in mixed precision), one node-level column index per 3x3 block (4 B per
    9 values), row offsets, node offsets, and four FP64 DOF vectors (read z, read p_old, write
    p_new, write Ap)."""
    nnz, ndof = model.nnz, model.ndof
    nv = ndof // 3
    return nnz * value_bytes + (nnz // 9) * 4 + (ndof + 1) * 4 + (nv + 1) * 4 + 4 * ndof * 8


def stiffness_bytes(sc, value_bytes: int = 8) -> int:
    """Algorithmic bytes of one stiffness assembly (per element: every entry of K_e into the CSR
    values, read-modify-write, with a 4-byte slot index; u gathered; geometry corners; connectivity;
    32 B per point of per-element rules)."""
    n = sc.nodes_per_element
    g = 4 if sc.element_type <= 2 else 8
    q = per_element_points(sc.stiffness_quadrature, sc.num_interior, sc.num_elements)
    per = (3 * n) ** 2 * (2 * value_bytes + 4) + 24 * n + 24 * g + 4 * n
    return int(sc.num_elements * per + 32 * q[sc.num_interior:].sum())


def force_bytes(sc) -> int:
    """Algorithmic bytes of one force assembly (per element: u gathered, geometry corners,
    connectivity, f read-modify-write; 32 B per point of per-element rules)."""
    n = sc.nodes_per_element
    g = 4 if sc.element_type <= 2 else 8
    q = per_element_points(sc.elliptic_quadrature, sc.num_interior, sc.num_elements)
    return int(sc.num_elements * (24 * n + 24 * g + 4 * n + 48 * n) + 32 * q[sc.num_interior:].sum())


def per_element_points(quad, num_interior, num_elements):
    import numpy as np

    q = np.full(num_elements, len(quad.interior_w), dtype=np.int64)
    if quad.interface_offsets is not None and num_elements > num_interior:
        q[num_interior:] = np.diff(np.asarray(quad.interface_offsets, dtype=np.int64))
    return q


def tangent_flops(sc):
    """(algorithmic, issued) flops of one set of element tangents. Algorithmic: per point
    n(n+1)/2 * 162 for the contraction plus ~1000 for dP/dF and the gradients. Issued: the
    v_mfma_f64_16x16x4 products of the tangent kernel (six component pairs, nodes padded to 16 or
    32 so a pair is 1 or 2 x 2 tiles, the three diagonal pairs without their tile above the diagonal;
    ceil(3 c / 4) k-steps per chunk of c <= 8 points, 2048 flops each)."""
    import numpy as np

    n = sc.nodes_per_element
    q = per_element_points(sc.stiffness_quadrature, sc.num_interior, sc.num_elements)
    alg = float(q.sum()) * (n * (n + 1) / 2 * 162 + 1000)
    nt = 2 if n > 16 else 1
    full, rest = q // 8, q % 8
    ksteps = full * 6 + (3 * rest + 3) // 4
    tiles = 3 * nt * nt + 3 * nt * (nt + 1) // 2
    issued = float(ksteps.sum()) * tiles * 2048
    return alg, issued


# ------------------------------------------------------------------ HBM traffic (rocprofv3 counters)
def _probe_command(args, counter, outdir):
    cmd = ["rocprofv3", "--pmc", counter, "--output-format", "csv", "-d", outdir, "-o", "run", "--",
           sys.executable, os.path.abspath(__file__), "--traffic-probe", "--config", str(args.config)]
    for flag, on in (("--mixed", args.mode == 1), ("--mixed-refined", args.mode == 2),
                     ("--mixed-f64-tangents", args.mode == 3), ("--reorder", args.reorder),
                     ("--no-autotune", args.no_autotune)):
        if on:
            cmd.append(flag)
    if args.quadrature != "sampled":
        cmd += ["--quadrature", args.quadrature]
    return cmd


def _read_counters(outdir):
    """rows (dispatch id, kernel name, counter value) of a counter run, in dispatch order"""
    rows = []
    for path in glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True):
        with open(path, newline="") as fh:
            for r in csv.DictReader(fh):
                try:
                    rows.append((int(r.get("Dispatch_Id", 0)), r["Kernel_Name"], float(r["Counter_Value"])))
                except (KeyError, ValueError):
                    continue
    rows.sort()
    return rows


def _median(xs):
    xs = sorted(xs)
    if not xs:
        return None
    m = len(xs) // 2
    return xs[m] if len(xs) % 2 else 0.5 * (xs[m - 1] + xs[m])


def _kernel_kib(rows):
    """median counter value (KiB) per kernel kind: k_cg_a by lanes per node (launches that found the
    solve finished, which move almost nothing, left out), the element tangent, the Jacobian gather (the
    block gather that directly follows a tangent launch), element forces, force gather, and the
    calibration copy"""
    out = {"cg_a": {}, "tangent": [], "jgather": [], "forces": [], "fgather": [], "copy": []}
    cga = {}
    prev = ""
    for _, name, v in rows:
        m = re.search(r"k_cg_a<(\d+), false, (double|float), false>", name)
        if m:
            cga.setdefault(int(m.group(1)), []).append(v)
        elif "k_element_stiffness_tw" in name or "k_element_stiffness<" in name:
            out["tangent"].append(v)
        elif "k_block_gather" in name and ("k_element_stiffness" in prev):
            out["jgather"].append(v)
        elif "k_element_forces" in name:
            out["forces"].append(v)
        elif "k_node_force" in name:
            out["fgather"].append(v)
        elif "copyBuffer" in name or "copy_kernel" in name:
            out["copy"].append(v)
        prev = name
    for g, vs in cga.items():
        top = max(vs)
        out["cg_a"][g] = _median([v for v in vs if v > 0.5 * top])
    for k in ("tangent", "jgather", "forces", "fgather"):
        out[k] = _median(out[k])
    out["copy"] = max(out["copy"]) if out["copy"] else None  # the 256 MiB calibration copy
    return out


def measure_traffic(args):
    """Two short counter runs of this workload under rocprofv3 (FETCH_SIZE, then WRITE_SIZE), started
    before this process touches the GPU. Each counter is scaled by what it reads for a 256 MiB device
    copy in the same run. Returns per-kernel HBM bytes per launch, or a reason string."""
    if shutil.which("rocprofv3") is None:
        return "rocprofv3 not found"
    res = {}
    env = dict(os.environ, TMPDIR="/tmp")
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        outdir = tempfile.mkdtemp(prefix=f"hoefem_{counter.lower()}_", dir="/tmp")
        t0 = time.time()
        try:
            r = subprocess.run(_probe_command(args, counter, outdir), cwd="/tmp", env=env, capture_output=True,
                               text=True, timeout=180)
            rc = r.returncode
        except subprocess.TimeoutExpired:
            rc = "timeout"
        rows = _read_counters(outdir)
        shutil.rmtree(outdir, ignore_errors=True)
        print(f"[bench] counter run {counter}: rc {rc}, {len(rows)} dispatches, {time.time() - t0:.1f}s",
              file=sys.stderr, flush=True)
        if not rows:
            return f"{counter} run produced no counters (rc {rc})"
        kib = _kernel_kib(rows)
        if not kib["copy"]:
            return f"{counter} run has no calibration copy"
        scale = COPY_BYTES / 1024.0 / kib["copy"]  # true KiB per counted KiB
        res[counter] = {"scale": scale, "kib": kib}
    return res


def traffic_bytes(tr, kind, lanes=None):
    """measured HBM bytes (read + write) per launch of one kernel kind, or None"""
    if not isinstance(tr, dict):
        return None
    tot = 0.0
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        kib = tr[counter]["kib"][kind]
        if kind == "cg_a":
            kib = kib.get(lanes) if kib else None
        if kib is None:
            return None
        tot += kib * 1024.0 * tr[counter]["scale"]
    return int(round(tot))


def cpu_info():
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count()
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return cores, model


def common_line(args, metric, value, ms, dtype, config, **extra):
    line = {"metric": metric, "value": round(value, 3), "unit": "timesteps/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic (seeded scene generator)", "config": config}
    line.update(extra)
    return line


def bench_learned(args):
    """Config 4: the embedded Hex20 cylinder of config 2, bottom cap fixed and top cap pulled by seeded
    random Dirichlet loads (a new target of 0.05-0.2 x the height every 50 steps, reached by a ramp).
    The first harvest steps run the plain Schwarz-CG and store (residual, error) snapshots of its
    solves; the network (512 aggregates, 64-dim residual projection in, 64-dim correction out) is
    trained on them; then the next steps run twice from the same state, plain and with the learned
    correction, each step to the same 1e-6 relative Newton residual."""
    import numpy as np
    import torch

    from higher_order_embedded_fem_amd import fem, scenes

    sc, pulled = scenes.pull_scene(scenes.config_scene(2))
    total = args.harvest_steps + args.eval_steps
    sched = scenes.pull_schedule(total, 1.6, seed=args.seed)
    model = fem.ElasticityModel(sc)
    sims = {name: fem.NewtonCgIntegrator(model, sc, newton_tolerance_factor=1e-6, mixed_precision=args.mode)
            for name in ("plain", "learned")}
    for sim in sims.values():
        if not args.no_autotune:
            sim.autotune()
    lc = fem.LearnedCoarseCorrection(sims["learned"], num_aggregates=512, seed=0)

    def set_pull(sim, k):
        vals = np.zeros((sc.dirichlet_nodes.size, 3))
        vals[pulled] = sched[k]
        sim.set_dirichlet_values(vals)

    u = torch.zeros(sc.ndof, dtype=torch.float64, device="cuda")
    v = torch.zeros_like(u)
    lc.harvest(True, capacity=64, every=128)
    harvest_cg = 0
    for k in range(args.harvest_steps):
        set_pull(sims["learned"], k)
        harvest_cg += sims["learned"].step(k * sc.dt, sc.dt, u, v)["cg_iterations"]
    lc.harvest(False)
    info = lc.info()
    t0 = time.perf_counter()
    l0, l1 = lc.train(steps=args.train_steps, learning_rate=3e-3, batch=8)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    print(f"[bench] harvest: {args.harvest_steps} steps, {harvest_cg} CG iterations, {info['stored_pairs']} "
          f"snapshots x {info['num_aggregates']} aggregates; training {t_train:.1f}s, loss {l0:.4e} -> {l1:.4e}",
          file=sys.stderr, flush=True)
    lc.enable(True)
    res = {}
    for name, sim in sims.items():
        uu, vv = u.clone(), v.clone()
        cg = newton = done = 0
        worst = 0.0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.harvest_steps, total):
            set_pull(sim, k)
            st = sim.step(k * sc.dt, sc.dt, uu, vv)
            cg += st["cg_iterations"]
            newton += st["newton_iterations"]
            worst = max(worst, st["final_residual"] / st["tolerance"])
            done += 1
            if done % 10 == 0:
                print(f"[bench] {name}: {done} steps, {time.perf_counter() - t0:.1f}s, cg {cg}", file=sys.stderr,
                      flush=True)
            if time.perf_counter() - t0 > args.eval_seconds:
                break  # the wall-time cap: the rate below is over the steps done
        torch.cuda.synchronize()
        res[name] = {"seconds": time.perf_counter() - t0, "cg": cg, "newton": newton, "u": uu.cpu().numpy(),
                     "worst": worst, "steps": done}
        print(f"[bench] {name}: {done} steps in {res[name]['seconds']:.3f}s, cg {cg}, newton {newton}, "
              f"largest final residual / tolerance {worst:.3f}", file=sys.stderr, flush=True)
    info = lc.info()
    d = res["learned"]["u"] - res["plain"]["u"] if res["learned"]["steps"] == res["plain"]["steps"] else None
    value = res["learned"]["steps"] / res["learned"]["seconds"]
    line = common_line(
        args, "timesteps/s at 1e-6 rel Newton residual", value, 1e3 * res["learned"]["seconds"] / res["learned"]["steps"],
        "f32+f64" if args.mixed else "f64",
        {"workload": sc.name + " (seeded random top-cap pulls) + learned coarse correction", "dofs": sc.ndof,
         "nnz": model.nnz, "harvest_steps": args.harvest_steps, "eval_steps": args.eval_steps,
         "train_steps": args.train_steps, "aggregates": info["num_aggregates"],
         "precision": "mixed" if args.mixed else "fp64", "parallelism": "single GPU"},
        learned={"plain_value": round(res["plain"]["steps"] / res["plain"]["seconds"], 3),
                 "steps_timed": [res["plain"]["steps"], res["learned"]["steps"]],
                 "speedup_vs_plain": round(value / (res["plain"]["steps"] / res["plain"]["seconds"]), 4),
                 "cg_iteration_ratio_plain_over_learned": round(
                     (res["plain"]["cg"] / res["plain"]["steps"]) / max(res["learned"]["cg"] / res["learned"]["steps"], 1e-9), 4),
                 "cg_iterations_per_step": res["learned"]["cg"] / res["learned"]["steps"],
                 "cg_iterations_per_step_plain": res["plain"]["cg"] / res["plain"]["steps"],
                 "worst_final_residual_over_tolerance": [round(res["plain"]["worst"], 4),
                                                         round(res["learned"]["worst"], 4)],
                 "fallbacks": info["fallbacks"], "learned_solves": info["learned_solves"],
                 "harvested_snapshots": info["stored_pairs"],
                 "training_pairs": info["stored_pairs"] * info["num_aggregates"],
                 "train_seconds": round(t_train, 3), "loss_before_after": [l0, l1],
                 "u_rel_l2_vs_plain": None if d is None else
                 float(np.linalg.norm(d) / max(np.linalg.norm(res["plain"]["u"]), 1e-300))},
        roofline=None, cpu_baseline=None)
    line["steps"] = res["learned"]["steps"]
    print(json.dumps(line))


def bench_batched(args):
    """Config 5: independent scenes of uneven size stepped together by hoefem_batch_step (lockstep
    Newton loops, one batched PCG per round with captured iteration batches)."""
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
                 "batch_step_ms": round(r["batch_step_ms"], 3),
                 "newton_iterations": r["newton_iterations"], "cg_iterations": r["cg_iterations"]},
        roofline=None, cpu_baseline=None)
    print(json.dumps(line))


def build_scene(args):
    from higher_order_embedded_fem_amd import scenes

    sc = (scenes.config_scene(args.config) if args.quadrature == "sampled"
          else scenes.cylinder_scene(24, scenes.HEX20, 8, quadrature="embedded"))
    return scenes.reorder(sc) if args.reorder else sc


def run_probe(args):
    """The run profiled by measure_traffic: the autotuner's trials (every lanes-per-node variant of
    the CG kernels), one 256 MiB calibration copy, and one step from rest."""
    import torch

    from higher_order_embedded_fem_amd import fem

    sc = build_scene(args)
    model = fem.ElasticityModel(sc)
    sim = fem.NewtonCgIntegrator(model, sc, newton_tolerance_factor=1e-6, mixed_precision=args.mode)
    if not args.no_autotune:
        sim.autotune()
    src = torch.ones(COPY_BYTES // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    u = torch.zeros(sc.ndof, dtype=torch.float64, device="cuda")
    v = torch.zeros_like(u)
    sim.step(0.0, sc.dt, u, v)
    torch.cuda.synchronize()


def main():
    args = parse()
    args.mode = 3 if args.mixed_f64_tangents else (2 if args.mixed_refined else (1 if args.mixed else 0))
    args.mixed = args.mode != 0
    if args.traffic_probe:
        run_probe(args)
        return
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (asked for {args.gpus}); the library has no multi-GPU path.",
              file=sys.stderr)
        sys.exit(2)
    if args.config in (4, 5):
        import torch

        assert torch.cuda.is_available(), "bench.py needs a HIP device"
        (bench_learned if args.config == 4 else bench_batched)(args)
        return
    # the counter runs go first: this process must not have touched the GPU when it starts them
    traffic = "skipped (--no-traffic)" if args.no_traffic else measure_traffic(args)
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
    sc = build_scene(args)
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
    precision = {0: "fp64", 1: "mixed", 2: "mixed-refined", 3: "mixed-f64-tangents"}[args.mode]

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
    kern = {}
    if not args.no_profile:
        L.hoefem_profile_enable(model.h, 1)
        torch.cuda.synchronize()
        p0 = time.perf_counter()
        run(args.steps)
        torch.cuda.synchronize()
        prof_elapsed = time.perf_counter() - p0
        L.hoefem_profile_read(model.h, C.byref(full_n), C.byref(full_ms), C.byref(all_n), C.byref(all_ms))
        for name, kind in (("tangent", 0), ("jgather", 1), ("forces", 2), ("fgather", 3)):
            n_, ms_ = C.c_int64(), C.c_double()
            L.hoefem_profile_kernel(model.h, kind, C.byref(n_), C.byref(ms_))
            kern[name] = (n_.value, ms_.value)
        L.hoefem_profile_enable(model.h, 0)
    value = args.steps / elapsed
    last_res = round(tot["res_over_tol"], 4)
    print(f"[bench] {args.steps} steps in {elapsed:.3f}s: {value:.2f} steps/s; newton {tot['newton_iterations']}, "
          f"cg {tot['cg_iterations']}", file=sys.stderr, flush=True)

    roofline = None
    traffic_note = traffic if isinstance(traffic, str) else (
        "rocprofv3 FETCH_SIZE and WRITE_SIZE runs of this workload started by this bench (one step from rest "
        "after the autotuner's trials), each scaled by its reading of a 256 MiB device copy in the same run "
        f"(scales {traffic['FETCH_SIZE']['scale']:.4f}, {traffic['WRITE_SIZE']['scale']:.4f})")
    if full_n.value > 0:
        avg_ms = full_ms.value / full_n.value
        nbytes = cg_kernel_bytes(model, 4 if args.mixed else 8)
        achieved = nbytes / (avg_ms * 1e-3) / 1e9
        ga = lanes[0] if lanes is not None else None
        if ga is None:
            nnz_per_node = model.nnz / 9 / max(sc.num_vertices, 1)
            ga = 32 if nnz_per_node >= 40 else (16 if nnz_per_node >= 12 else 8)
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_bytes(traffic, "cg_a", ga),
                    "traffic_measured_in": traffic_note,
                    "kernel": "k_cg_a (CG system-matrix kernel: p = z + beta p, Ap = J p, p.Ap)",
                    "bytes_per_launch": nbytes, "avg_launch_us": round(avg_ms * 1e3, 3),
                    "launches": full_n.value, "all_launches_avg_us": round(all_ms.value / max(all_n.value, 1) * 1e3, 3),
                    "all_launches": all_n.value,
                    "measured_in": f"a second pass over the same {args.steps} steps with per-kernel timing "
                                   f"({1e3 * prof_elapsed / args.steps:.3f} ms/step)"}

    assembly = None
    if kern.get("tangent", (0, 0))[0] > 0 and kern.get("jgather", (0, 0))[0] > 0:
        # element tangents + their gather into J, per assembly (one per Newton iteration)
        n_asm = kern["tangent"][0]
        t_ms = (kern["tangent"][1] + kern["jgather"][1]) / n_asm
        sb = stiffness_bytes(sc, 4 if args.mode == 1 else 8)
        alg_fl, mfma_fl = tangent_flops(sc)
        tt = traffic_bytes(traffic, "tangent")
        tg = traffic_bytes(traffic, "jgather")
        fb = force_bytes(sc)
        f_ms = ((kern["forces"][1] + kern["fgather"][1]) / kern["forces"][0]) if kern["forces"][0] else None
        ft, fg = traffic_bytes(traffic, "forces"), traffic_bytes(traffic, "fgather")
        assembly = {
            "stiffness": {
                "bound": "hbm", "achieved": round(sb / (t_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(sb / (t_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                "traffic": (tt + tg) if (tt is not None and tg is not None) else None,
                "bytes_per_assembly": sb, "avg_assembly_us": round(t_ms * 1e3, 3),
                "tangent_us": round(kern["tangent"][1] / n_asm * 1e3, 3),
                "gather_us": round(kern["jgather"][1] / kern["jgather"][0] * 1e3, 3),
                "assemblies": n_asm,
                "kernels": "element tangents (k_element_stiffness_tw, v_mfma_f64_16x16x4) + color-ordered gather "
                           "into J = M + dt^2 K (k_block_gather)",
                "mfma": {"algorithmic_gflop": round(alg_fl / 1e9, 3), "issued_gflop": round(mfma_fl / 1e9, 3),
                         "tflops_algorithmic": round(alg_fl / (kern["tangent"][1] / n_asm * 1e-3) / 1e12, 2),
                         "tflops_issued": round(mfma_fl / (kern["tangent"][1] / n_asm * 1e-3) / 1e12, 2),
                         "peak_tflops": FP64_PEAK_TFLOPS,
                         "frac_issued": round(mfma_fl / (kern["tangent"][1] / n_asm * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4)}
                if args.mode != 1 else None},
            "forces": None if f_ms is None else {
                "bound": "hbm", "achieved": round(fb / (f_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(fb / (f_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                "traffic": (ft + fg) if (ft is not None and fg is not None) else None,
                "bytes_per_assembly": fb, "avg_assembly_us": round(f_ms * 1e3, 3), "assemblies": kern["forces"][0],
                "kernels": "element forces (k_element_forces) + gather into f = M g + f_int (k_node_force)"},
        }

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
        phases = {"forces": 0.0, "stiffness": 0.0, "jacobian": 0.0, "schwarz": 0.0, "cg": 0.0}
        c0 = time.perf_counter()
        for k in range(args.cpu_steps):
            st = osim.step(k * sc.dt, sc.dt, uo, vo)
            for key, field in (("forces", "t_forces"), ("stiffness", "t_stiffness"), ("jacobian", "t_jacobian"),
                               ("schwarz", "t_precond"), ("cg", "t_cg")):
                phases[key] += st[field]
        c1 = time.perf_counter()
        nproc, model_name = cpu_info()
        cpu = {"value": round(args.cpu_steps / (c1 - c0), 4), "unit": "timesteps/s", "cores": threads, "kind": "port",
               "sample": f"first {args.cpu_steps} of the {sc.steps} steps of {sc.name} (same settings), "
                         f"C++ FP64 oracle with OpenMP over element colors",
               "phase_ms_per_step": {k: round(1e3 * v / args.cpu_steps, 2) for k, v in phases.items()},
               "nproc": nproc, "cpu_model": model_name}
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
        "assembly": assembly,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
