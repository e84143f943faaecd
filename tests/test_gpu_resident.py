"""The resident Gauss-Seidel preconditioner (all five red-black sweeps of
Preconditioner!, src/NonUniform.jl:531-559, in one launch of co-resident
workgroups holding the iterate in LDS) against the per-colour launches:
whole solves must be bitwise identical (iteration count, residual history,
θ) for both solvers, steady and transient, FP64 and the FP32 refinement,
on grids with even and odd row lengths and tiles cut unevenly, and at the
240x240x30 headline case."""
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
    a = run(m, 1, solver, steady, precision, 1e-8)
    b = run(m, 0, solver, steady, precision, 1e-8)
    assert a[3] > 0, "the resident preconditioner must have run"
    assert b[3] == 0
    assert a[0] == b[0]
    assert np.array_equal(a[1], b[1])
    assert np.array_equal(a[2], b[2])


@pytest.mark.parametrize("solver,tol", [(L.SOLVER_PBICGSTAB, 1e-4), (L.SOLVER_CG, 1e-6)])
def test_resident_headline_bitwise(solver, tol):
    """modelA 240x240x30, steady, PBiCGSTAB+GS to 1e-4 and CG+GS to 1e-6."""
    print("case resident 240x240x30 solver", solver, "tol", tol, flush=True)
    m = h2_amd.build_model(240, 240, 30)
    a = run(m, 1, solver, True, L.FP64, tol)
    b = run(m, 0, solver, True, L.FP64, tol)
    assert a[3] > 0
    assert a[0] == b[0]
    assert np.array_equal(a[1], b[1])
    assert np.array_equal(a[2], b[2])


def test_resident_kernel_timing_entry():
    """h3d_time_kernel ids 7 and 8 run the resident launch on its own."""
    m = h2_amd.build_model(40, 40, 30)
    with WorkBuffers(*m.shape) as wk:
        load_model(wk, m)
        wk.calc_rhs(10000.0, True)
        for which in (7, 8):
            ms = wk.time_kernel(which, 5)
            assert 0.0 < ms < 50.0
