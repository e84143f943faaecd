"""The 480x480x60 case with the seeded power map (layer stack with every
Zcase2 cell split in two): the FP64 solve to a relative residual of 1e-8
against an FP64 solve to 1e-12 (maximum temperature difference in K), and
the FP32-refined solve at the same 1e-8.  The refined solve must either
converge to the same accuracy or report that it did not (its guard keeps
the best correction; the temperature stays finite).  The FP32-refined
solve with CG! inner solves converges to the same accuracy."""
import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="no HIP device")]

import h2_amd
from h2_amd import _lib as L
from h2_amd.heat3ds import WorkBuffers, load_model


def test_c3_fp64_and_refined_against_tight_fp64():
    m = h2_amd.build_model(480, 480, 60, power="map")
    wk = WorkBuffers(*m.shape)
    load_model(wk, m)
    th0 = wk.download("theta")

    def run(prec, tol, solver=L.SOLVER_PBICGSTAB):
        wk.upload("theta", th0)
        wk.fill("b", 0.0)
        wk.calc_rhs(10000.0, True)
        conv, itr, _ = wk.solve(solver, 10000.0, tol, 20000, L.SMOOTHER_GS, True, precision=prec)
        return conv, itr, wk.last_solve_ms(), wk.download("theta")

    conv, itr, ms, ref = run(L.FP64, 1e-12)
    print("fp64 1e-12: itr %d, %.1f ms" % (itr, ms))
    assert conv
    conv, itr, ms, t64 = run(L.FP64, 1e-8)
    err64 = np.max(np.abs(t64 - ref))
    print("fp64 1e-8: itr %d, %.1f ms, max |dθ| %.3e K" % (itr, ms, err64))
    assert conv and err64 < 1e-5
    conv, itr, ms, t32 = run(L.FP32_REFINED, 1e-8)
    err32 = np.max(np.abs(t32 - ref))
    print("fp32-refined 1e-8: converged %s, inner itr %d, %.1f ms, max |dθ| %.3e K" % (conv, itr, ms, err32))
    assert np.all(np.isfinite(t32))
    if conv:
        assert err32 < 1e-5
    # CG! inner solves converge in FP32 on this case
    conv, itr, ms, tc = run(L.FP32_REFINED, 1e-8, L.SOLVER_CG)
    errc = np.max(np.abs(tc - ref))
    print("fp32-refined CG 1e-8: converged %s, inner itr %d, %.1f ms, max |dθ| %.3e K" % (conv, itr, ms, errc))
    assert conv and errc < 1e-5
    wk.close()
