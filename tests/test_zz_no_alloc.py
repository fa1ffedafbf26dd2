"""hoefem_pcg_noalloc (the allocation-free PCG) against hoefem_pcg through the C ABI.

Two calls in a row on one GPU, the second on a smaller model (the scratch array keeps the first
call's contents, which the second must not depend on), each compared with hoefem_pcg on the same
system: same kernels and launch shapes, so the solution and the iteration count are identical.
Also: the Jacobi variant, a zero right-hand side, the iteration limit, and the size limit.
"""

import ctypes as C

import numpy as np
import pytest

import oracle as O
from higher_order_embedded_fem_amd import _lib, scenes

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(x):
    return torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64, device="cuda")


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def system(fem, sc, seed):
    """The scene's rest-state Jacobian M + 1e-4 K (Dirichlet applied), a right-hand side and the
    scene's Schwarz operator updated with it."""
    om = O.Model(sc)
    mu, lam = scenes.lame_from_young_poisson(sc.young, sc.poisson)
    rp, col = om.pattern()
    K = om.stiffness_add(mu, lam, np.zeros(sc.ndof), np.zeros(om.nnz))
    J = O.dirichlet_csr(rp, col, (0.0 + 1e-4 * K) + 1.0 * om.mass(sc.density), sc.dirichlet_nodes)
    gm = fem.ElasticityModel(sc)
    sw = fem.SchwarzPreconditioner.from_scene(gm, sc)
    Jd = dev(J)
    sw.update(Jd)
    b = np.random.default_rng(seed).standard_normal(sc.ndof)
    return gm, sw, Jd, dev(b)


def solve(entry, gm, Jd, sw, bd, tol, max_iter):
    x = torch.zeros_like(bd)
    it = C.c_int64(-1)
    st = entry(gm.h, Jd.data_ptr(), sw.h if sw is not None else None, bd.data_ptr(), x.data_ptr(), tol, max_iter,
               C.byref(it))
    return st, it.value, host(x)


@pytest.fixture(scope="module")
def fem():
    from higher_order_embedded_fem_amd import fem as F

    return F


def test_noalloc_matches_pcg_on_two_calls_in_a_row(fem):
    L = _lib.load()
    big = system(fem, scenes.cylinder_scene(8, scenes.HEX20, 8), 1)
    small = system(fem, scenes.cube_scene(3, scenes.TET10), 2)
    assert big[0].ndof > small[0].ndof
    for gm, sw, Jd, bd in (big, small):
        for pre in (sw, None):  # Schwarz and Jacobi preconditioners
            st_r, it_r, x_r = solve(L.hoefem_pcg, gm, Jd, pre, bd, 1e-10, 5000)
            st_n, it_n, x_n = solve(L.hoefem_pcg_noalloc, gm, Jd, pre, bd, 1e-10, 5000)
            print(f"ndof {gm.ndof} {'schwarz' if pre else 'jacobi'}: {it_r} / {it_n} iterations, status {st_r} / {st_n}")
            # Jacobi on the embedded cylinder stops at the iteration limit; both entry points must
            # stop the same way with the same iterate
            assert st_n == st_r and st_r in (0, 5)
            if pre is not None:
                assert st_r == 0
            assert it_n == it_r and it_r > 0
            assert np.array_equal(x_n, x_r)


def test_noalloc_limits_and_statuses(fem):
    L = _lib.load()
    gm, sw, Jd, bd = system(fem, scenes.cube_scene(3, scenes.TET10), 3)
    # the iteration limit: the same status and partial solution as hoefem_pcg
    st_r, it_r, x_r = solve(L.hoefem_pcg, gm, Jd, sw, bd, 1e-14, 3)
    st_n, it_n, x_n = solve(L.hoefem_pcg_noalloc, gm, Jd, sw, bd, 1e-14, 3)
    assert st_r == st_n == 5 and it_r == it_n == 3 and np.array_equal(x_n, x_r)
    # a zero right-hand side gives x = 0 after a call that left other values in the scratch
    st, it, x = solve(L.hoefem_pcg_noalloc, gm, Jd, sw, torch.zeros_like(bd), 1e-10, 100)
    assert st == 0 and it == 0 and not np.any(x)
    # a model above the 262144-DOF limit is rejected before any device work
    sc = scenes.cube_scene(44, scenes.TET4)
    assert sc.ndof > 262144
    gbig = fem.ElasticityModel(sc)
    A = torch.zeros(gbig.nnz, dtype=torch.float64, device="cuda")
    b = torch.zeros(gbig.ndof, dtype=torch.float64, device="cuda")
    st, it, _ = solve(L.hoefem_pcg_noalloc, gbig, A, None, b, 1e-10, 10)
    assert st == 1 and it == 0
    assert b"262144" in L.hoefem_last_error()
