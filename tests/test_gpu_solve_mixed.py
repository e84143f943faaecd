"""Whole solves with every boundary-condition type against the oracle.

The other solve tests use the modelA stack under BC mode 3.  Here the six
faces mix convection, heat flux and isothermal conditions (one isothermal
face copies λ, ρ and cp into its halo), the materials are nine random
(λ, ρ, cp) classes, and the heat source is random: calRHS! followed by
PBiCGSTAB!/CG! with GS(5) must be bitwise the oracle in its exact-sum mode
(convergence flag, iteration count, res0, residual history and θ), steady
and transient, with one sweep launch per colour and with the two-colour
launches."""
import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="no HIP device")]

from h2_amd import _lib as L
from tests.test_gpu_kernels import make_pair, mixed_bc, put


@pytest.mark.parametrize("size", [(40, 40, 30), (13, 70, 60)])
@pytest.mark.parametrize("solver", ["pbicgstab", "cg"])
@pytest.mark.parametrize("steady", [True, False])
@pytest.mark.parametrize("chunk", [0, 7])
def test_mixed_bc_solve_bitwise(oracle, size, solver, steady, chunk):
    print("case", size, solver, "steady" if steady else "transient", "chunk", chunk, flush=True)
    dt = 0.25
    wk, c, rng = make_pair(oracle, *size, seed=11, bc_set=mixed_bc(), random_lam="classes", steady=steady, dt=dt)
    wk.set_sweep_chunk(chunk)
    shape = c.shape
    put(wk, c, "hsrc", rng.uniform(0.0, 2.0e11, shape))
    put(wk, c, "b", np.zeros(shape))
    wk.calc_rhs(dt, steady)
    with oracle.sum_mode(oracle.EXACT):
        c.cal_rhs()
        conv, itr, res0, hist = c.solve(solver, tol=1e-6, maxItr=2000, smoother=1)
    assert np.array_equal(wk.download("b"), c.arr["b"])
    s = L.SOLVER_CG if solver == "cg" else L.SOLVER_PBICGSTAB
    out = wk.solve(s, dt, 1e-6, 2000, L.SMOOTHER_GS, steady)
    print("  oracle", conv, itr, res0, flush=True)
    assert wk.num_classes() > 0
    assert out == (conv, itr, res0)
    assert np.array_equal(wk.history(), hist)
    assert np.array_equal(wk.download("theta"), c.arr["theta"])
    wk.close()


@pytest.mark.parametrize("solver", ["pbicgstab", "cg"])
def test_mixed_bc_fp32_refined(oracle, solver):
    """The FP32-refined solve on the same mixed-condition case (tol 1e-8).
    The oracle recomputes the true residual of the returned θ, which must be
    below 1e-8 of res0.  θ is compared with the oracle's FP64 CG solve to
    1e-12: on this case an FP64 PBiCGSTAB solve stopped at 1e-8 (residual
    4.0e-9) is 2.1e-4 K from it and one stopped at 1e-9 is 3.7e-5 K away, so
    the bound is 1e-3 K."""
    dt = 0.25
    wk, c, rng = make_pair(oracle, 40, 40, 30, seed=12, bc_set=mixed_bc(), random_lam="classes", steady=True, dt=dt)
    shape = c.shape
    put(wk, c, "hsrc", rng.uniform(0.0, 2.0e11, shape))
    put(wk, c, "b", np.zeros(shape))
    wk.calc_rhs(dt, True)
    with oracle.sum_mode(oracle.EXACT):
        c.cal_rhs()
        conv_o, itr_o, res0_o, _ = c.solve("cg", tol=1e-12, maxItr=5000, smoother=1)
    assert conv_o
    s = L.SOLVER_CG if solver == "cg" else L.SOLVER_PBICGSTAB
    conv, itr, res0 = wk.solve(s, dt, 1e-8, 5000, L.SMOOTHER_GS, True, precision=L.FP32_REFINED)
    print("solver", solver, "refined", conv, itr, wk.history(), flush=True)
    assert conv and res0 == res0_o
    th = wk.download("theta")
    dth = np.max(np.abs(th - c.arr["theta"]))
    c.arr["theta"][:] = th
    rel = c.calc_rk("r", "theta", "b") / res0_o
    print("  true residual", rel, "max dθ", dth, flush=True)
    assert rel < 1e-8
    assert dth < 1e-3
    wk.close()


@pytest.mark.parametrize("solver", ["pbicgstab", "cg"])
def test_mixed_bc_no_smoother_bitwise(oracle, solver):
    """Preconditioner! :none (a copy) on the mixed-condition case, bitwise
    against the oracle; CG! then forms ρ with a separate dot pass."""
    print("case no smoother", solver, flush=True)
    dt = 0.25
    wk, c, rng = make_pair(oracle, 40, 40, 30, seed=13, bc_set=mixed_bc(), random_lam="classes", steady=False, dt=dt)
    shape = c.shape
    put(wk, c, "hsrc", rng.uniform(0.0, 2.0e11, shape))
    put(wk, c, "b", np.zeros(shape))
    wk.calc_rhs(dt, False)
    with oracle.sum_mode(oracle.EXACT):
        c.cal_rhs()
        conv, itr, res0, hist = c.solve(solver, tol=1e-6, maxItr=3000, smoother=0)
    s = L.SOLVER_CG if solver == "cg" else L.SOLVER_PBICGSTAB
    out = wk.solve(s, dt, 1e-6, 3000, L.SMOOTHER_NONE, False)
    print("  oracle", conv, itr, res0, flush=True)
    assert out == (conv, itr, res0)
    assert np.array_equal(wk.history(), hist)
    assert np.array_equal(wk.download("theta"), c.arr["theta"])
    wk.close()
