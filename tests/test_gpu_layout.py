"""The colour-split layout of the solver's Krylov vectors against the
reference's layout: whole solves must be bitwise identical (same iteration
count, residual history and θ), for both solvers and smoothers, steady and
transient, FP64 and the FP32 refinement, on grids with an even and an odd
number of stored cells per row (the odd one has padding entries)."""
import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="no HIP device")]

import h2_amd
from h2_amd import _lib as L
from h2_amd.heat3ds import WorkBuffers, load_model


def run(shape, layout, solver, smoother, steady, precision, tol):
    m = h2_amd.build_model(*shape)
    wk = WorkBuffers(*m.shape)
    wk.set_layout(layout)
    load_model(wk, m)
    wk.calc_rhs(10000.0, steady)
    res = wk.solve(solver, 10000.0, tol, 3000, smoother, steady, precision=precision)
    out = (res, wk.history(), wk.download("theta"), wk.num_classes())
    wk.close()
    return out


@pytest.mark.parametrize("shape", [(40, 40, 30), (17, 9, 30), (13, 70, 60)])
@pytest.mark.parametrize("solver,smoother", [(L.SOLVER_PBICGSTAB, L.SMOOTHER_GS), (L.SOLVER_CG, L.SMOOTHER_GS),
                                             (L.SOLVER_PBICGSTAB, L.SMOOTHER_NONE), (L.SOLVER_CG, L.SMOOTHER_NONE)])
@pytest.mark.parametrize("steady", [True, False])
def test_fp64_layouts_bitwise(shape, solver, smoother, steady):
    tol = 1e-8 if smoother == L.SMOOTHER_GS else 1e-5
    a = run(shape, L.LAYOUT_AUTO, solver, smoother, steady, L.FP64, tol)
    b = run(shape, L.LAYOUT_STANDARD, solver, smoother, steady, L.FP64, tol)
    assert a[3] > 0, "the table form (and so the colour-split layout) must be active"
    assert a[0] == b[0]
    assert np.array_equal(a[1], b[1])
    assert np.array_equal(a[2], b[2])


@pytest.mark.parametrize("shape", [(40, 40, 30), (17, 9, 30)])
@pytest.mark.parametrize("solver", [L.SOLVER_PBICGSTAB, L.SOLVER_CG])
def test_fp32_refined_layouts_bitwise(shape, solver):
    a = run(shape, L.LAYOUT_AUTO, solver, L.SMOOTHER_GS, True, L.FP32_REFINED, 1e-8)
    b = run(shape, L.LAYOUT_STANDARD, solver, L.SMOOTHER_GS, True, L.FP32_REFINED, 1e-8)
    assert a[0] == b[0]
    assert np.array_equal(a[1], b[1])
    assert np.array_equal(a[2], b[2])


def run_limited(shape, layout, solver, steady, precision, max_itr, chunk=None):
    m = h2_amd.build_model(*shape)
    wk = WorkBuffers(*m.shape)
    wk.set_layout(layout)
    if chunk is not None:
        wk.set_sweep_chunk(chunk)
    load_model(wk, m)
    dt = 10000.0 if steady else 1e-3
    wk.calc_rhs(dt, steady)
    res = wk.solve(solver, dt, 1e-12, max_itr, L.SMOOTHER_GS, steady, precision=precision)
    out = (res, wk.history(), wk.download("theta"))
    wk.close()
    return out


@pytest.mark.parametrize("solver,steady,precision", [(L.SOLVER_PBICGSTAB, True, L.FP64),
                                                     (L.SOLVER_CG, False, L.FP64),
                                                     (L.SOLVER_CG, False, L.FP32_REFINED)])
def test_large_grid_layouts_bitwise(solver, steady, precision):
    """260x260x60 has 2.13 M cells per colour half, where the colour-split
    sweep takes its large-grid forms (FP64 two cells per thread, 131
    half-row indices, so a row ends inside a pair; FP32 three rows per
    thread); 30 iterations of each path."""
    print("case 260x260x60 solver", solver, "steady", steady, "precision", precision, flush=True)
    a = run_limited((260, 260, 60), L.LAYOUT_AUTO, solver, steady, precision, 30)
    b = run_limited((260, 260, 60), L.LAYOUT_STANDARD, solver, steady, precision, 30)
    assert a[0] == b[0]
    assert np.array_equal(a[1], b[1])
    assert np.array_equal(a[2], b[2])


@pytest.mark.parametrize("nx", [257, 256])
@pytest.mark.parametrize("solver,steady,precision", [(L.SOLVER_PBICGSTAB, True, L.FP64),
                                                     (L.SOLVER_CG, False, L.FP64),
                                                     (L.SOLVER_CG, False, L.FP32_REFINED)])
def test_large_grid_odd_extent_bitwise(nx, solver, steady, precision):
    """nx x 270 x 60 (over 2.1 M cells per half): 257 (NX odd, 130
    half-row indices) and 256 (129 half-row indices), so that the last
    thread of a row holds one, two or three cells of the large-grid kernels
    (two FP64 or four FP32 cells per thread) inside the row; 30 iterations
    against the reference layout."""
    print("case %dx270x60 solver" % nx, solver, "steady", steady, "precision", precision, flush=True)
    a = run_limited((nx, 270, 60), L.LAYOUT_AUTO, solver, steady, precision, 30)
    b = run_limited((nx, 270, 60), L.LAYOUT_STANDARD, solver, steady, precision, 30)
    assert a[0] == b[0]
    assert np.array_equal(a[1], b[1])
    assert np.array_equal(a[2], b[2])


@pytest.mark.parametrize("shape,solver,steady,precision", [
    ((480, 480, 60), L.SOLVER_PBICGSTAB, True, L.FP64),
    ((480, 480, 60), L.SOLVER_CG, False, L.FP32_REFINED),
    ((960, 960, 120), L.SOLVER_PBICGSTAB, False, L.FP64)])
def test_full_size_grids_bitwise(shape, solver, steady, precision):
    """The large configurations' own grids (480x480x60 steady, 960x960x120
    transient): 10 iterations of the colour-split kernels (NC cells per
    thread) bitwise the reference layout's."""
    print("case %dx%dx%d solver" % shape, solver, "steady", steady, "precision", precision, flush=True)
    a = run_limited(shape, L.LAYOUT_AUTO, solver, steady, precision, 10)
    b = run_limited(shape, L.LAYOUT_STANDARD, solver, steady, precision, 10)
    assert a[0] == b[0]
    assert np.array_equal(a[1], b[1])
    assert np.array_equal(a[2], b[2])


@pytest.mark.parametrize("precision", [L.FP64, L.FP32_REFINED])
def test_large_grid_two_colour_sweeps_bitwise(precision):
    """The two-colour sweep launches (chunks of 30 planes) with CG!, whose
    preconditioner then starts in the swap buffer and forms ρ in a last
    per-colour pass, on 260x260x60 (three FP32 rows per thread): bitwise the
    per-colour launches and the reference's layout, 30 iterations."""
    print("case 260x260x60 CG transient chunk 30 precision", precision, flush=True)
    a = run_limited((260, 260, 60), L.LAYOUT_AUTO, L.SOLVER_CG, False, precision, 30, chunk=30)
    b = run_limited((260, 260, 60), L.LAYOUT_AUTO, L.SOLVER_CG, False, precision, 30, chunk=0)
    c = run_limited((260, 260, 60), L.LAYOUT_STANDARD, L.SOLVER_CG, False, precision, 30)
    for o in (b, c):
        assert a[0] == o[0]
        assert np.array_equal(a[1], o[1])
        assert np.array_equal(a[2], o[2])


def transient_solve(wk, th0, dt, solver, precision):
    wk.upload("theta", th0)
    wk.fill("b", 0.0)
    wk.calc_rhs(dt, False)
    res = wk.solve(solver, dt, 1e-10, 3000, L.SMOOTHER_GS, False, precision=precision)
    return res, wk.history(), wk.download("theta")


@pytest.mark.parametrize("precision", [L.FP64, L.FP32_REFINED])
@pytest.mark.parametrize("solver", [L.SOLVER_PBICGSTAB, L.SOLVER_CG])
def test_transient_dt_change_on_one_context(solver, precision):
    """A second transient solve with another Δt on the same context (whose
    captured iteration carries 1/Δt) equals the same solve on a fresh
    context and in the reference's layout, bit for bit."""
    m = h2_amd.build_model(40, 40, 30)
    out = []
    for layout, dts in ((L.LAYOUT_AUTO, (1e-3, 2e-3)), (L.LAYOUT_AUTO, (2e-3,)), (L.LAYOUT_STANDARD, (2e-3,))):
        with WorkBuffers(*m.shape) as wk:
            wk.set_layout(layout)
            load_model(wk, m)
            th0 = wk.download("theta")
            for dt in dts:
                r = transient_solve(wk, th0, dt, solver, precision)
            out.append(r)
    for o in out[1:]:
        assert out[0][0] == o[0]
        assert np.array_equal(out[0][1], o[1])
        assert np.array_equal(out[0][2], o[2])
