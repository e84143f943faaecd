"""GPU parity of whole solves (calRHS! + PBiCGSTAB!/CG!) against the oracle.

Every element-wise operation of the device iteration is bitwise the
reference's.  Its grid sums are exactly rounded (order-independent), while
the reference sums sequentially.  So:
  * against the oracle in exact-sum mode the device solve is bitwise equal:
    same iteration count, same residual history, same θ;
  * against the oracle in the reference's sequential mode the two are
    different roundings of the same iteration; for a tightly converged solve
    θ must agree to 1e-6 K (the reference's FP64 solve of the same inputs).
"""
import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="no HIP device")]

import h2_amd
from h2_amd import _lib as L
from h2_amd.heat3ds import WorkBuffers, load_model


def device_case(NX, NY, NZ, power="modelA", form=L.COEF_AUTO):
    m = h2_amd.build_model(NX, NY, NZ, power=power)
    wk = WorkBuffers(*m.shape)
    wk.set_coef_form(form)
    load_model(wk, m)
    return wk, m


def oracle_solve(oracle, NX, NY, NZ, solver, smoother, tol, maxItr=20000, steady=True, exact=True):
    with oracle.sum_mode(oracle.EXACT if exact else oracle.SEQUENTIAL):
        c = oracle.modelA_case(NX, NY, NZ, steady=steady)
        c.cal_rhs()
        conv, itr, res0, hist = c.solve(solver, tol=tol, maxItr=maxItr, smoother=smoother)
    return c, conv, itr, res0, hist


@pytest.mark.parametrize("form", [L.COEF_AUTO, L.COEF_DIRECT])
@pytest.mark.parametrize("solver,smoother,tol", [("pbicgstab", 1, 1e-6), ("cg", 1, 1e-6), ("pbicgstab", 0, 1e-6),
                                                 ("cg", 0, 1e-4)])
def test_c1_solve_bitwise_vs_exact_oracle(oracle, solver, smoother, tol, form):
    c, conv, itr, res0, hist = oracle_solve(oracle, 40, 40, 30, solver, smoother, tol)
    wk, m = device_case(40, 40, 30, form=form)
    wk.calc_rhs(10000.0, True)
    assert np.array_equal(wk.download("b"), c.arr["b"])
    s = L.SOLVER_CG if solver == "cg" else L.SOLVER_PBICGSTAB
    conv_d, itr_d, res0_d = wk.solve(s, 10000.0, tol, 20000, smoother, True)
    assert (conv_d, itr_d, res0_d) == (conv, itr, res0)
    assert np.array_equal(wk.history(), hist)
    assert np.array_equal(wk.download("theta"), c.arr["theta"])
    wk.close()


@pytest.mark.parametrize("solver", ["pbicgstab", "cg"])
def test_c1_tight_solve_vs_reference_sums(oracle, solver):
    """Against the reference's sequential sums: converged to 1e-10, θ within 1e-6 K."""
    c, conv, itr, res0, hist = oracle_solve(oracle, 40, 40, 30, solver, 1, 1e-10, exact=False)
    wk, m = device_case(40, 40, 30)
    wk.calc_rhs(10000.0, True)
    s = L.SOLVER_CG if solver == "cg" else L.SOLVER_PBICGSTAB
    conv_d, itr_d, res0_d = wk.solve(s, 10000.0, 1e-10, 20000, 1, True)
    assert conv and conv_d
    assert abs(res0_d - res0) <= 1e-14 * res0
    assert np.max(np.abs(wk.download("theta") - c.arr["theta"])) < 1e-6
    wk.close()


def test_c1_transient_bitwise(oracle):
    """One implicit-Euler step (is_steady = false, Δt = 1e4 s as q3d sets)."""
    c, conv, itr, res0, hist = oracle_solve(oracle, 40, 40, 30, "pbicgstab", 1, 1e-8, steady=False)
    wk, m = device_case(40, 40, 30)
    wk.calc_rhs(10000.0, False)
    assert np.array_equal(wk.download("b"), c.arr["b"])
    conv_d, itr_d, res0_d = wk.solve(L.SOLVER_PBICGSTAB, 10000.0, 1e-8, 20000, 1, False)
    assert (conv_d, itr_d, res0_d) == (conv, itr, res0)
    assert np.array_equal(wk.history(), hist)
    assert np.array_equal(wk.download("theta"), c.arr["theta"])
    wk.close()


def test_solve_is_deterministic():
    out = []
    for _ in range(2):
        wk, m = device_case(40, 40, 30)
        wk.calc_rhs(10000.0, True)
        wk.solve(L.SOLVER_PBICGSTAB, 10000.0, 1e-6, 20000, 1, True)
        out.append((wk.download("theta"), wk.history()))
        wk.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_zero_residual_returns_early(oracle):
    """res0 == 0 returns (true, 0, 0) (src/NonUniform.jl:270-272)."""
    wk, m = device_case(40, 40, 30)
    wk.upload("hsrc", np.zeros(m.shape))
    wk.calc_rhs(10000.0, True)  # b = 0 with θ = θ∞ everywhere... but the sink h·A·θ makes r != 0
    th = wk.download("theta")
    # make θ = 0 inside and b = 0: then r = b - Aθ = 0 exactly
    th[1:-1, 1:-1, 1:-1] = 0.0
    th[:, :, 0] = 0.0
    wk.upload("theta", th)
    wk.fill("b", 0.0)
    conv, itr, res0 = wk.solve(L.SOLVER_PBICGSTAB, 10000.0, 1e-6, 100, 1, True)
    assert (conv, itr, res0) == (True, 0, 0.0)
    conv, itr, res0 = wk.solve(L.SOLVER_CG, 10000.0, 1e-6, 100, 1, True)
    assert (conv, itr, res0) == (True, 0, 0.0)
    wk.close()


def test_max_itr_not_converged(oracle):
    c, conv, itr, res0, hist = oracle_solve(oracle, 40, 40, 30, "pbicgstab", 1, 1e-12, maxItr=3)
    wk, m = device_case(40, 40, 30)
    wk.calc_rhs(10000.0, True)
    conv_d, itr_d, res0_d = wk.solve(L.SOLVER_PBICGSTAB, 10000.0, 1e-12, 3, 1, True)
    assert (conv_d, itr_d) == (False, 3) == (conv, itr)
    assert np.array_equal(wk.history(), hist)
    wk.close()


def test_breakdown_matches_oracle(oracle):
    """A residual so small that ρ = (r, r0) < FloatMin = 1e-37 breaks down at
    iteration 1 without converging (src/NonUniform.jl:291-294)."""
    c = oracle.modelA_case(8, 6, 30)
    wk = WorkBuffers(*c.shape)
    wk.set_grid(c.dx, c.dy, c.ZC, c.dZ)
    wk.set_bc(h2_amd.set_mode3_bc_parameters())
    for nm in ("lam", "rho", "cp"):
        wk.upload(nm, c.arr[nm])
    th = np.zeros(c.shape, order="F")
    wk.upload("theta", th)
    c.arr["theta"][:] = th
    b = np.zeros(c.shape, order="F")
    b[3, 3, 3] = 1e-20
    wk.upload("b", b)
    c.arr["b"][:] = b
    conv_o, itr_o, res0_o, _ = c.solve("pbicgstab", tol=1e-6, maxItr=10, smoother=1)
    conv_d, itr_d, res0_d = wk.solve(L.SOLVER_PBICGSTAB, 10000.0, 1e-6, 10, 1, True)
    assert (conv_o, itr_o) == (False, 1)
    assert (conv_d, itr_d) == (conv_o, itr_o)
    assert res0_d == res0_o
    wk.close()


def test_nan_is_an_error():
    wk, m = device_case(40, 40, 30)
    wk.calc_rhs(10000.0, True)
    b = wk.download("b")
    b[5, 5, 5] = np.nan
    wk.upload("b", b)
    with pytest.raises(L.H3DError) as e:
        wk.solve(L.SOLVER_PBICGSTAB, 10000.0, 1e-6, 50, 1, True)
    assert e.value.code == L.ERR_NAN
    wk.close()


def test_bad_smoother_is_an_error():
    wk, m = device_case(40, 40, 30)
    with pytest.raises(L.H3DError) as e:
        wk.solve(L.SOLVER_PBICGSTAB, 10000.0, 1e-6, 50, 7, True)
    assert e.value.code == L.ERR_ARG
    with pytest.raises(ValueError):
        h2_amd.PBiCGSTAB(wk, (m.dx, m.dy, 1.0), 10000.0, m.ZC, m.dZ, None, smoother=":jacobi")
    wk.close()


def test_fp32_refined_reaches_fp64_accuracy(oracle):
    """FP32 inner PBiCGSTAB+GS inside FP64 refinement to a true relative
    residual of 1e-8; the answer must match a tight FP64 solve."""
    wk, m = device_case(40, 40, 30)
    wk.calc_rhs(10000.0, True)
    conv, itr, res0 = wk.solve(L.SOLVER_PBICGSTAB, 10000.0, 1e-8, 20000, 1, True, precision=L.FP32_REFINED)
    hist = wk.history()
    assert conv and hist[-1] < 1e-8 and not wk.last_solve_floor()
    th32 = wk.download("theta")
    c, conv_o, itr_o, res0_o, _ = oracle_solve(oracle, 40, 40, 30, "pbicgstab", 1, 1e-12, exact=False)
    assert abs(res0 - res0_o) <= 1e-13 * res0_o
    assert np.max(np.abs(th32 - c.arr["theta"])) < 1e-5
    wk.close()


def test_fp32_refined_stops_at_rounding_floor(oracle):
    """A tolerance below what FP64 can resolve (1e-30): the refinement must
    stop at the rounding floor of its FP64 residual, report convergence
    there, and not spend the iteration budget; the answer matches a tight
    FP64 solve of the oracle."""
    wk, m = device_case(40, 40, 30)
    wk.calc_rhs(10000.0, True)
    conv, itr, res0 = wk.solve(L.SOLVER_PBICGSTAB, 10000.0, 1e-30, 20000, 1, True, precision=L.FP32_REFINED)
    hist = wk.history()
    assert conv and wk.last_solve_floor()
    assert itr < 5000 and len(hist) < 20
    assert 1e-30 < hist[-1] < 1e-10
    th32 = wk.download("theta")
    c, conv_o, itr_o, res0_o, _ = oracle_solve(oracle, 40, 40, 30, "pbicgstab", 1, 1e-12, exact=False)
    assert np.max(np.abs(th32 - c.arr["theta"])) < 1e-6
    # an FP64 solve afterwards clears the flag
    wk.fill("b", 0.0)
    wk.calc_rhs(10000.0, True)
    wk.solve(L.SOLVER_PBICGSTAB, 10000.0, 1e-6, 20000, 1, True)
    assert not wk.last_solve_floor()
    wk.close()


@pytest.mark.parametrize("form", [L.FORM_REFERENCE, L.FORM_FLUX])
def test_residual_scale_terms(form):
    """||s||, s = |b| + sum|terms of A theta|: with theta = 0 it is ||b||;
    with one unit cell and b = 0 every s entry is one term of A e_c (the
    diagonal at c, a conductance at each neighbour), so it equals ||A e_c||
    bitwise, at an interior cell and at cells on convective and isothermal
    faces; it is homogeneous and bounds ||b - A theta||."""
    import math
    wk, m = device_case(40, 40, 30)
    wk.set_operator_form(form)
    sh = m.shape
    rng = np.random.default_rng(3)
    b = np.asfortranarray(rng.normal(0.0, 1e-3, sh))
    wk.upload("b", b)
    wk.upload("theta", np.zeros(sh))
    inner = b[1:-1, 1:-1, 1:-1].ravel()
    assert wk.residual_scale("theta", "b", 10000.0, True) == math.sqrt(math.fsum(inner * inner))
    wk.fill("b", 0.0)
    for c in [(20, 20, 15), (1, 7, 1), (40, 40, 30), (5, 1, 30)]:
        e = np.zeros(sh, order="F")
        e[c] = 1.0
        wk.upload("theta", e)
        s = wk.residual_scale("theta", "b", 10000.0, True)
        r = wk.calc_rk("pcg_r", "theta", "b", 10000.0, True)
        assert s == r and s > 0
    th = np.asfortranarray(rng.uniform(290.0, 310.0, sh))
    wk.upload("theta", th)
    wk.upload("b", b)
    s1 = wk.residual_scale("theta", "b", 10000.0, False)
    assert wk.calc_rk("pcg_r", "theta", "b", 10000.0, False) <= s1
    wk.upload("theta", 2.0 * th)
    wk.upload("b", 2.0 * b)
    assert wk.residual_scale("theta", "b", 10000.0, False) == 2.0 * s1
    wk.close()


def test_q3d_drop_in(capsys):
    r = h2_amd.q3d(40, 40, 30, "pbicgstab", "gs", epsilon=1e-6, par="hip", is_steady=True)
    out = capsys.readouterr().out
    assert "Initial residual = " in out and "Converged at" in out
    assert r.isconverged and r.itr > 0
    with pytest.raises(ValueError):
        h2_amd.q3d(40, 40, 30, "pbicgstab", "gs", par="thread")
