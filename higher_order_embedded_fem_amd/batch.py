"""Many independent scenes advanced together on one GPU (hoefem_batch_step).

The scenes' Newton loops run in lockstep from one host thread: each round issues every scene's
residual evaluations, Jacobian assemblies or line-search trials before one synchronization, and
the round's linear solves run as one batched Schwarz-PCG whose launches cover all scenes, each
scene stopping on the device when it converges; iteration batches are captured HIP graphs. Each
scene keeps its own model and tile shapes (lanes per node row of J and of P, picked by the
library's autotuner); the models are spread round robin over a few streams so the per-scene
assembly kernels of different scenes overlap. Every scene's result equals what stepping it alone
gives, bit for bit.
"""

from __future__ import annotations

import math
import time
from typing import Dict, List

import numpy as np

from . import scenes


def _cube_cells_for(ndof: float) -> int:
    return max(2, int(round(((ndof / 3.0) ** (1.0 / 3.0) - 1.0) / 2.0)))


def _cylinder_cells_for(ndof: float) -> int:
    # dofs of the hollow-cylinder Hex20 scene grow like 51.1 * cells^2.227 (fit of 8 and 24 cells)
    return max(4, int(round((ndof / 51.1) ** (1.0 / 2.227))))


def batched_scenes(count: int = 64, seed: int = 5, dof_range=(2e3, 2e5), steps: int = 10) -> List[scenes.Scene]:
    """count scenes: DOF targets log-uniform in dof_range, Tet10 cubes and Hex20 hollow
    cylinders alternating, E log-uniform in [1e5, 1e7], nu uniform in [0.3, 0.48]."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        target = math.exp(rng.uniform(math.log(dof_range[0]), math.log(dof_range[1])))
        young = math.exp(rng.uniform(math.log(1e5), math.log(1e7)))
        poisson = float(rng.uniform(0.3, 0.48))
        if k % 2 == 0:
            sc = scenes.cube_scene(_cube_cells_for(target), scenes.TET10, young=young, poisson=poisson, steps=steps)
        else:
            sc = scenes.cylinder_scene(_cylinder_cells_for(target), scenes.HEX20, 8, young=young, poisson=poisson,
                                       steps=steps, seed=seed * 1000 + k)
        sc.name = f"batch{k:02d}_{sc.name}"
        out.append(sc)
    return out


class BatchedStepper:
    def __init__(self, scene_list: List[scenes.Scene], newton_tolerance_factor: float = 1e-6,
                 mixed_precision: int = 0, autotune: bool = True, streams: int = 4, **sim_kwargs):
        import ctypes as C

        import torch

        from . import _lib, fem

        self.scenes = scene_list
        self.streams = [torch.cuda.Stream() for _ in range(max(1, min(streams, len(scene_list))))]
        self.models, self.sims, self.tiles = [], [], []
        L = _lib.load()
        for k, sc in enumerate(scene_list):
            m = fem.ElasticityModel(sc)
            _lib.check(L.hoefem_model_set_stream(m.h, self.streams[k % len(self.streams)].cuda_stream))
            s = fem.NewtonCgIntegrator(m, sc, newton_tolerance_factor=newton_tolerance_factor,
                                       mixed_precision=mixed_precision, **sim_kwargs)
            self.tiles.append(s.autotune() if autotune else None)
            self.models.append(m)
            self.sims.append(s)
        torch.cuda.synchronize()
        self.u = [torch.zeros(sc.ndof, dtype=torch.float64, device="cuda") for sc in scene_list]
        self.v = [torch.zeros_like(u) for u in self.u]
        n = len(scene_list)
        handles = (C.c_void_p * n)(*[s.h.value for s in self.sims])
        h = C.c_void_p()
        _lib.check(L.hoefem_batch_create(handles, n, C.byref(h)))
        self.h = h
        self._dt = (C.c_double * n)(*[sc.dt for sc in scene_list])
        self._u = (C.c_void_p * n)(*[u.data_ptr() for u in self.u])
        self._v = (C.c_void_p * n)(*[v.data_ptr() for v in self.v])
        self._stats = (_lib.StepStats * n)()
        self._lat = (C.c_double * n)()
        self._status = (C.c_int32 * n)()

    def __del__(self):
        from . import _lib

        if getattr(self, "h", None) is not None and _lib._lib is not None:
            _lib._lib.hoefem_batch_destroy(self.h)
            self.h = None

    def reset(self):
        import torch

        for u, v in zip(self.u, self.v):
            u.zero_()
            v.zero_()
        torch.cuda.synchronize()

    def step(self) -> Dict:
        """One time step of every scene (each with its own dt): per-scene stats, the time until each
        scene's Newton loop ended, and its status."""
        from . import _lib

        _lib.check(_lib.load().hoefem_batch_step(self.h, self._dt, self._u, self._v, self._stats, self._lat,
                                                  self._status))
        names = [k for k, _ in _lib.StepStats._fields_]
        return {
            "stats": [{k: getattr(st, k) for k in names} for st in self._stats],
            "latency_ms": list(self._lat),
            "status": list(self._status),
        }

    def run(self, steps: int) -> Dict:
        """Advance every scene by `steps` time steps; returns aggregate throughput and latency."""
        import torch

        torch.cuda.synchronize()
        t0 = time.perf_counter()
        lat, newton, cg, walls = [], 0, 0, []
        for _ in range(steps):
            t1 = time.perf_counter()
            r = self.step()
            walls.append(time.perf_counter() - t1)
            lat.extend(r["latency_ms"])
            newton += sum(s["newton_iterations"] for s in r["stats"])
            cg += sum(s["cg_iterations"] for s in r["stats"])
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        lat = np.array(lat)
        return {
            "scene_steps": len(self.scenes) * steps,
            "wall_s": wall,
            "aggregate_timesteps_per_s": len(self.scenes) * steps / wall,
            "batch_step_ms": float(np.mean(walls) * 1e3),
            "p50_step_latency_ms": float(np.percentile(lat, 50)),
            "p99_step_latency_ms": float(np.percentile(lat, 99)),
            "newton_iterations": newton,
            "cg_iterations": cg,
            "total_dofs": int(sum(sc.ndof for sc in self.scenes)),
        }
