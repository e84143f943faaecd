"""The resident Gauss-Seidel preconditioner (all five red-black sweeps of
Preconditioner!, src/NonUniform.jl:531-559, in one launch of co-resident
workgroups holding the iterate in LDS; mode 1) and the whole loops of
PBiCGSTAB! and CG! (src/NonUniform.jl:283-349, 160-223) in one such launch
(mode 2) against the
per-colour launches (mode 0): whole solves must be bitwise identical
(iteration count, residual history, θ) for both solvers, steady and
transient, FP64 and the FP32 refinement, on grids with even and odd row
lengths and tiles cut unevenly, and at the 240x240x30 headline case."""
import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="no HIP device")]

import h2_amd
from h2_amd import _lib as L
from h2_amd.heat3ds import WorkBuffers, load_model


def run(m, resident, solver, steady, precision, tol, max_itr=3000):
    wk = WorkBuffers(*m.shape)
    wk.set_resident(resident)
    load_model(wk, m)
    dt = 10000.0 if steady else 1e-3
    wk.calc_rhs(dt, steady)
    res = wk.solve(solver, dt, tol, max_itr, L.SMOOTHER_GS, steady, precision=precision)
    out = (res, wk.history(), wk.download("theta"), wk.resident_workgroups())
    wk.close()
    return out


@pytest.mark.parametrize("shape", [(40, 40, 30), (17, 9, 30), (13, 70, 60), (64, 48, 30)])
@pytest.mark.parametrize("solver", [L.SOLVER_PBICGSTAB, L.SOLVER_CG])
@pytest.mark.parametrize("steady,precision", [(True, L.FP64), (False, L.FP64), (True, L.FP32_REFINED)])
def test_resident_solves_bitwise(shape, solver, steady, precision):
    print("case resident", shape, "solver", solver, "steady", steady, "precision", precision, flush=True)
    m = h2_amd.build_model(*shape)
    b = run(m, 0, solver, steady, precision, 1e-8)
    assert b[3] == 0
    for mode in (1, 2):
        a = run(m, mode, solver, steady, precision, 1e-8)
        assert a[3] > 0, "the resident preconditioner must have run"
        assert a[0] == b[0], mode
        assert np.array_equal(a[1], b[1]), mode
        assert np.array_equal(a[2], b[2]), mode


@pytest.mark.parametrize("solver", [L.SOLVER_PBICGSTAB, L.SOLVER_CG])
@pytest.mark.parametrize("max_itr", [1, 2, 7])
def test_resident_one_launch_max_itr(solver, max_itr):
    """The one-launch loop stops at maxItr with the state, history and θ the
    per-iteration launches leave."""
    m = h2_amd.build_model(40, 40, 30)
    a = run(m, 2, solver, True, L.FP64, 1e-12, max_itr=max_itr)
    b = run(m, 0, solver, True, L.FP64, 1e-12, max_itr=max_itr)
    assert a[0] == b[0]
    assert len(a[1]) == max_itr
    assert np.array_equal(a[1], b[1])
    assert np.array_equal(a[2], b[2])


@pytest.mark.parametrize("solver,tol,mode", [(L.SOLVER_PBICGSTAB, 1e-4, 1), (L.SOLVER_PBICGSTAB, 1e-4, 2),
                                             (L.SOLVER_CG, 1e-6, 1), (L.SOLVER_CG, 1e-6, 2)])
def test_resident_headline_bitwise(solver, tol, mode):
    """modelA 240x240x30, steady, PBiCGSTAB+GS to 1e-4 and CG+GS to 1e-6."""
    print("case resident 240x240x30 solver", solver, "tol", tol, "mode", mode, flush=True)
    m = h2_amd.build_model(240, 240, 30)
    a = run(m, mode, solver, True, L.FP64, tol)
    b = run(m, 0, solver, True, L.FP64, tol)
    assert a[3] > 0
    assert a[0] == b[0]
    assert np.array_equal(a[1], b[1])
    assert np.array_equal(a[2], b[2])


def test_resident_kernel_timing_entry():
    """h3d_time_kernel ids 7 to 11 run the resident launches on their own;
    id 11 (the one-launch loop) leaves θ as it was."""
    m = h2_amd.build_model(40, 40, 30)
    with WorkBuffers(*m.shape) as wk:
        load_model(wk, m)
        wk.calc_rhs(10000.0, True)
        th0 = wk.download("theta")
        for which in (7, 8, 9, 10, 11):
            ms = wk.time_kernel(which, 5)
            assert 0.0 < ms < 50.0
        assert np.array_equal(wk.download("theta"), th0)
