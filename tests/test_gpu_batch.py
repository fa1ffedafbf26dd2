"""Batched stepping (hoefem_batch_step through batch.BatchedStepper) against each scene stepped alone.

* The 64-scene workload at its stated sizes (Tet10 cubes and embedded Hex20 hollow cylinders, DOF
  counts log-uniform in [2e3, 2e5], E log-uniform in [1e5, 1e7], nu uniform in [0.3, 0.48]; about
  2.5M DOFs in all): two batched steps, then every scene stepped alone through hoefem_sim_step with
  the same sim (so the same autotuned tile shapes) from the same start. u, v and the Newton and CG
  iteration counts are equal, bit for bit.
* A small batch with a scene whose single-step attempt fails (Newton limit 2 at dt 0.05: five
  substeps, simulation_toolbox/src/fem/newton_cg.rs:146-228) and one whose substepping gives up
  (Newton limit 1 at dt 0.2: status SUBSTEPPING): statuses, statistics and states equal the
  single-scene results, and the substepped scene matches the oracle in parity mode.
* The FP32-operator precision modes (1 and 3) on a 12-scene batch: equal to each scene alone, bit
  for bit.
"""

import numpy as np
import pytest

import oracle as O
from higher_order_embedded_fem_amd import scenes

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def run_alone(stepper, steps):
    """Every scene of the stepper stepped alone (hoefem_sim_step) from zero: states and stats."""
    from higher_order_embedded_fem_amd._lib import HoefemError

    out = []
    for k, (sc, sim) in enumerate(zip(stepper.scenes, stepper.sims)):
        u = torch.zeros(sc.ndof, dtype=torch.float64, device="cuda")
        v = torch.zeros_like(u)
        stats, status = [], 0
        for i in range(steps):
            try:
                stats.append(sim.step(i * sc.dt, sc.dt, u, v))
            except HoefemError as e:
                status = e.code
                break
        torch.cuda.synchronize()
        out.append((u, v, stats, status))
    return out


@pytest.mark.timeout(900)
def test_batched_64_scenes_equal_each_scene_alone():
    from higher_order_embedded_fem_amd import batch

    scs = batch.batched_scenes(64)
    total = sum(s.ndof for s in scs)
    print(f"64 scenes, {total} DOFs, {min(s.ndof for s in scs)}..{max(s.ndof for s in scs)} per scene", flush=True)
    assert 2e6 < total < 4e6
    stepper = batch.BatchedStepper(scs, newton_tolerance_factor=1e-6, autotune=True)
    steps = 2
    bstats = []
    for i in range(steps):
        r = stepper.step()
        assert all(s == 0 for s in r["status"]), r["status"]
        bstats.append(r["stats"])
        lat = np.array(r["latency_ms"])
        print(f"batched step {i}: latency p50 {np.percentile(lat, 50):.1f} ms, max {lat.max():.1f} ms, newton "
              f"{sum(s['newton_iterations'] for s in r['stats'])}, cg {sum(s['cg_iterations'] for s in r['stats'])}",
              flush=True)
    torch.cuda.synchronize()
    alone = run_alone(stepper, steps)
    for k, (u, v, stats, status) in enumerate(alone):
        assert status == 0, (k, status)
        assert torch.equal(u, stepper.u[k]) and torch.equal(v, stepper.v[k]), k
        for i in range(steps):
            assert stats[i] == bstats[i][k], (k, i, stats[i], bstats[i][k])


@pytest.mark.timeout(600)
@pytest.mark.parametrize("mode", [1, 3])
def test_batched_fp32_operator_equals_each_scene_alone(mode):
    """The FP32-operator modes (1: FP32 tangents; 3: FP64 tangents rounded into the FP32 Jacobian) run
    their batched solves on the FP32 CG kernels: 12 scenes of 2e3..3e4 DOFs, two batched steps, equal
    to each scene stepped alone bit for bit."""
    from higher_order_embedded_fem_amd import batch

    scs = batch.batched_scenes(12, seed=11, dof_range=(2e3, 3e4))
    stepper = batch.BatchedStepper(scs, newton_tolerance_factor=1e-6, mixed_precision=mode, autotune=True)
    steps = 2
    bstats = []
    for i in range(steps):
        r = stepper.step()
        assert all(s == 0 for s in r["status"]), r["status"]
        bstats.append(r["stats"])
    torch.cuda.synchronize()
    print(f"mode {mode}: newton {[s['newton_iterations'] for s in bstats[-1]]}, "
          f"cg {[s['cg_iterations'] for s in bstats[-1]]}", flush=True)
    alone = run_alone(stepper, steps)
    for k, (u, v, stats, status) in enumerate(alone):
        assert status == 0, (k, status)
        assert torch.equal(u, stepper.u[k]) and torch.equal(v, stepper.v[k]), k
        for i in range(steps):
            assert stats[i] == bstats[i][k], (k, i, stats[i], bstats[i][k])


@pytest.mark.timeout(600)
def test_batch_with_substepping_and_failing_scenes():
    import ctypes as C

    from higher_order_embedded_fem_amd import _lib, fem

    PARITY = dict(cg_tolerance=1e-12, cg_max_iter=20000)
    base = scenes.cube_scene(3, scenes.TET10, young=1e5)
    sub = scenes.cube_scene(3, scenes.TET10, young=1e5)
    sub.dt = 0.05  # fails with two Newton iterations in one step, succeeds with five substeps
    bad = scenes.cube_scene(3, scenes.TET10, young=1e6)
    bad.dt = 0.2  # fails at every substep count below 1000 with one Newton iteration
    cyl = scenes.cylinder_scene(8, scenes.HEX20, 8)
    scs = [base, sub, bad, cyl]
    limits = [100, 2, 1, 100]
    # one sim per scene with its own Newton limit, through the C ABI (BatchedStepper gives every
    # scene the same settings)
    L = _lib.load()
    models = [fem.ElasticityModel(sc) for sc in scs]
    sims = [fem.NewtonCgIntegrator(m, sc, newton_tolerance_factor=1e-9, max_newton_iter=lim, **PARITY)
            for m, sc, lim in zip(models, scs, limits)]
    h = C.c_void_p()
    handles = (C.c_void_p * 4)(*[s.h.value for s in sims])
    _lib.check(L.hoefem_batch_create(handles, 4, C.byref(h)))
    try:
        us = [torch.zeros(sc.ndof, dtype=torch.float64, device="cuda") for sc in scs]
        vs = [torch.zeros_like(u) for u in us]
        dts = (C.c_double * 4)(*[sc.dt for sc in scs])
        pu = (C.c_void_p * 4)(*[u.data_ptr() for u in us])
        pv = (C.c_void_p * 4)(*[v.data_ptr() for v in vs])
        stats = (_lib.StepStats * 4)()
        lat = (C.c_double * 4)()
        status = (C.c_int32 * 4)()
        rc = L.hoefem_batch_step(h, dts, pu, pv, stats, lat, status)
        torch.cuda.synchronize()
        assert rc == 10 and list(status) == [0, 0, 10, 0], (rc, list(status))  # SUBSTEPPING for scene 2
        assert b"scene 2" in L.hoefem_last_error()
        names = [k for k, _ in _lib.StepStats._fields_]
        bst = [{k: getattr(s, k) for k in names} for s in stats]
        assert bst[1]["substeps"] == 5 and bst[0]["substeps"] == 1 and bst[3]["substeps"] == 1
        # each scene alone, same sims, same start
        for k, (sc, sim) in enumerate(zip(scs, sims)):
            u = torch.zeros(sc.ndof, dtype=torch.float64, device="cuda")
            v = torch.zeros_like(u)
            if k == 2:
                with pytest.raises(_lib.HoefemError) as ei:
                    sim.step(0.0, sc.dt, u, v)
                assert ei.value.code == 10
            else:
                st = sim.step(0.0, sc.dt, u, v)
                assert st == bst[k], (k, st, bst[k])
            torch.cuda.synchronize()
            assert torch.equal(u, us[k]) and torch.equal(v, vs[k]), k
        # the substepped scene against the oracle (integrate_model_with_substeps)
        om = O.Model(sub)
        osim = O.Sim(om, newton_tolerance_factor=1e-9, max_newton_iter=2, **PARITY)
        uo, vo = np.zeros(sub.ndof), np.zeros(sub.ndof)
        so = osim.step(0.0, sub.dt, uo, vo)
        assert so["substeps"] == 5 and so["newton_iterations"] == bst[1]["newton_iterations"]
        ug = us[1].cpu().numpy()
        assert np.linalg.norm(ug - uo) <= 1e-8 * np.linalg.norm(uo)
    finally:
        L.hoefem_batch_destroy(h)


@pytest.mark.timeout(300)
def test_nonfinite_state_reported_alone_batched_and_by_oracle():
    """A NaN in a free DOF of u: the residual is NaN, the Newton test !(|F| > tol) passes, and the
    finite check after the step reports NOT_FINITE (status 11), in the oracle, hoefem_sim_step and
    hoefem_batch_step alike; the other scene of the batch is stepped normally."""
    import ctypes as C

    from higher_order_embedded_fem_amd import _lib, fem

    sc = scenes.cube_scene(2, scenes.TET10)
    free = int(np.setdiff1d(np.arange(sc.num_vertices), sc.dirichlet_nodes)[0])
    u0 = np.zeros(sc.ndof)
    u0[3 * free + 1] = np.nan
    om = O.Model(sc)
    osim = O.Sim(om)
    with pytest.raises(O.OracleError) as eo:
        osim.step(0.0, sc.dt, u0.copy(), np.zeros(sc.ndof))
    assert eo.value.code == 11
    L = _lib.load()
    models = [fem.ElasticityModel(sc), fem.ElasticityModel(sc)]
    sims = [fem.NewtonCgIntegrator(m, sc) for m in models]
    u = torch.tensor(u0, device="cuda")
    v = torch.zeros_like(u)
    with pytest.raises(_lib.HoefemError) as eg:
        sims[0].step(0.0, sc.dt, u, v)
    assert eg.value.code == 11
    h = C.c_void_p()
    _lib.check(L.hoefem_batch_create((C.c_void_p * 2)(*[s.h.value for s in sims]), 2, C.byref(h)))
    try:
        us = [torch.tensor(u0, device="cuda"), torch.zeros(sc.ndof, dtype=torch.float64, device="cuda")]
        vs = [torch.zeros_like(x) for x in us]
        status = (C.c_int32 * 2)()
        rc = L.hoefem_batch_step(h, (C.c_double * 2)(sc.dt, sc.dt), (C.c_void_p * 2)(*[x.data_ptr() for x in us]),
                                 (C.c_void_p * 2)(*[x.data_ptr() for x in vs]), None, None, status)
        torch.cuda.synchronize()
        assert rc == 11 and list(status) == [11, 0]
        assert torch.isnan(us[0]).any() and not torch.isnan(us[1]).any() and us[1].abs().max() > 0
    finally:
        L.hoefem_batch_destroy(h)
