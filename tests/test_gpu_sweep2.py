"""Two-colour sweep launches (H3D_OPT_SWEEP_CHUNK): a solve whose
preconditioner runs sweeps 2-5 as one out-of-place launch per sweep must be
bitwise the solve with one launch per colour (which the other GPU tests
hold bitwise to the oracle): same convergence flag, iteration count, res0,
residual history and θ, for FP64 and FP32-refined solves, steady and
transient operators, chunk lengths from 1 plane to the whole column, and
grids whose extents are not multiples of the 62 x 14 tile or of the chunk."""
import ctypes

import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="no HIP device")]

import h2_amd
from h2_amd import _lib as L
from h2_amd.heat3ds import WorkBuffers, load_model


def solve(NX, NY, NZ, chunk, solver, steady, precision=L.FP64, tol=1e-6, power="modelA"):
    m = h2_amd.build_model(NX, NY, NZ, power=power)
    wk = WorkBuffers(*m.shape)
    load_model(wk, m)
    wk.set_sweep_chunk(chunk)
    dt = 10000.0 if steady else 1e-3
    wk.fill("b", 0.0)
    wk.calc_rhs(dt, steady)
    out = wk.solve(solver, dt, tol, 20000, L.SMOOTHER_GS, steady, precision=precision)
    res = (out, wk.history(), wk.download("theta"))
    wk.close()
    return res


def same(a, b):
    (oa, ha, ta), (ob, hb, tb) = a, b
    assert oa == ob
    assert np.array_equal(ha, hb)
    assert np.array_equal(ta, tb)


@pytest.mark.parametrize("chunk", [1, 4, 7, 30])
def test_fp64_pbicgstab_steady(chunk):
    print("case 40x40x30 pbicgstab steady chunk", chunk, flush=True)
    ref = solve(40, 40, 30, 0, L.SOLVER_PBICGSTAB, True)
    assert ref[0][0]
    same(ref, solve(40, 40, 30, chunk, L.SOLVER_PBICGSTAB, True))


@pytest.mark.parametrize("shape,chunk", [((41, 29, 30), 7), ((130, 15, 30), 4), ((63, 30, 30), 16),
                                         ((48, 20, 60), 13)])
def test_fp64_cg_transient_ragged(shape, chunk):
    print("case", shape, "cg transient chunk", chunk, flush=True)
    ref = solve(*shape, 0, L.SOLVER_CG, False, tol=1e-8, power="map")
    assert ref[0][0]
    same(ref, solve(*shape, chunk, L.SOLVER_CG, False, tol=1e-8, power="map"))


@pytest.mark.parametrize("solver", [L.SOLVER_CG, L.SOLVER_PBICGSTAB])
def test_fp32_refined(solver):
    print("case 40x40x30 fp32-refined chunk 6 solver", solver, flush=True)
    ref = solve(40, 40, 30, 0, solver, True, precision=L.FP32_REFINED, tol=1e-8)
    assert ref[0][0]
    same(ref, solve(40, 40, 30, 6, solver, True, precision=L.FP32_REFINED, tol=1e-8))


def test_timing_entry_points():
    m = h2_amd.build_model(40, 40, 30)
    wk = WorkBuffers(*m.shape)
    load_model(wk, m)
    wk.set_sweep_chunk(0)
    with pytest.raises(L.H3DError):
        wk.time_kernel(5, 3)
    wk.set_sweep_chunk(8)
    v = ctypes.c_int()
    assert wk.lib.h3d_get_option(wk.ctx, L.OPT_SWEEP_CHUNK, ctypes.byref(v)) == 0 and v.value == 8
    assert wk.time_kernel(5, 3) > 0 and wk.time_kernel(6, 3) > 0
    with pytest.raises(L.H3DError):
        wk.set_sweep_chunk(-2)
    wk.close()


def test_auto_chunk_on_a_deep_grid():
    """The default (-1) on an operator of 120 planes: bitwise the
    per-colour solve, FP64 and FP32-refined."""
    print("case 24x20x120 pbicgstab transient auto chunk", flush=True)
    ref = solve(24, 20, 120, 0, L.SOLVER_PBICGSTAB, False, tol=1e-8, power="map")
    assert ref[0][0]
    same(ref, solve(24, 20, 120, -1, L.SOLVER_PBICGSTAB, False, tol=1e-8, power="map"))
    ref = solve(24, 20, 120, 0, L.SOLVER_CG, False, precision=L.FP32_REFINED, tol=1e-8, power="map")
    assert ref[0][0]
    same(ref, solve(24, 20, 120, -1, L.SOLVER_CG, False, precision=L.FP32_REFINED, tol=1e-8, power="map"))
