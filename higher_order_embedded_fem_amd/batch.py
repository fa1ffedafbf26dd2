"""Many independent scenes advanced together on one GPU.

Each scene gets its own model, simulation and HIP stream; host threads step the scenes
concurrently (the C ABI releases the Python GIL), so the kernels of different scenes overlap
on the device. The SpMV tile shape (lanes per node row block) of every scene's Jacobian and
Schwarz operator is picked by the library's autotuner.
"""

from __future__ import annotations

import concurrent.futures
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
                 mixed_precision: bool = False, autotune: bool = True, workers: int = 16):
        import torch

        from . import _lib, fem

        self.scenes = scene_list
        self.streams = [torch.cuda.Stream() for _ in scene_list]
        self.models, self.sims, self.tiles = [], [], []
        L = _lib.load()
        for sc, st in zip(scene_list, self.streams):
            m = fem.ElasticityModel(sc)
            _lib.check(L.hoefem_model_set_stream(m.h, st.cuda_stream))
            s = fem.NewtonCgIntegrator(m, sc, newton_tolerance_factor=newton_tolerance_factor,
                                       mixed_precision=mixed_precision)
            self.tiles.append(s.autotune() if autotune else None)
            self.models.append(m)
            self.sims.append(s)
        torch.cuda.synchronize()
        self.u = [torch.zeros(sc.ndof, dtype=torch.float64, device="cuda") for sc in scene_list]
        self.v = [torch.zeros_like(u) for u in self.u]
        self.workers = max(1, min(workers, len(scene_list)))

    def reset(self):
        import torch

        for u, v in zip(self.u, self.v):
            u.zero_()
            v.zero_()
        torch.cuda.synchronize()

    def _scene_steps(self, k: int, steps: int):
        sc, sim = self.scenes[k], self.sims[k]
        lat, stats = [], []
        for i in range(steps):
            t0 = time.perf_counter()
            stats.append(sim.step(i * sc.dt, sc.dt, self.u[k], self.v[k]))
            lat.append(time.perf_counter() - t0)
        return lat, stats

    def run(self, steps: int) -> Dict:
        """Advance every scene by `steps` time steps; returns aggregate throughput and latency."""
        import torch

        torch.cuda.synchronize()
        t0 = time.perf_counter()
        with concurrent.futures.ThreadPoolExecutor(max_workers=self.workers) as ex:
            results = list(ex.map(lambda k: self._scene_steps(k, steps), range(len(self.scenes))))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        lat = np.array([x for r in results for x in r[0]])
        newton = sum(s["newton_iterations"] for r in results for s in r[1])
        cg = sum(s["cg_iterations"] for r in results for s in r[1])
        return {
            "scene_steps": len(self.scenes) * steps,
            "wall_s": wall,
            "aggregate_timesteps_per_s": len(self.scenes) * steps / wall,
            "p50_step_latency_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_step_latency_ms": float(np.percentile(lat, 99) * 1e3),
            "newton_iterations": newton,
            "cg_iterations": cg,
            "total_dofs": int(sum(sc.ndof for sc in self.scenes)),
        }
