"""The device path on the closed-form boundary-condition cases of
tests/closed_forms.py (heat-flux faces on z+, x-, x+, y-, y+; one transient
step with and without a uniform source; isothermal X and Y faces with two
materials): the same answers the oracle's pins hold, computed through the
C ABI (h3d_apply_bc, h3d_calc_rhs, h3d_pbicgstab), in both layouts."""
import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="no HIP device")]

from h2_amd import _lib as L
from h2_amd.heat3ds import BoundaryConditionSet, WorkBuffers, genZ, pitch
from tests import closed_forms as CF


@pytest.mark.parametrize("layout", [L.LAYOUT_AUTO, L.LAYOUT_STANDARD])
@pytest.mark.parametrize("name", CF.CASES)
def test_closed_form_on_device(name, layout):
    case = CF.CASES[name]()
    MX, MY, MZ = case.NX + 2, case.NY + 2, case.NZ + 2
    Z, ZC, dZ = genZ(case.NZ)
    dx, dy = pitch(case.NX), pitch(case.NY)
    shape = (MX, MY, MZ)
    with WorkBuffers(MX, MY, MZ) as wk:
        wk.set_layout(layout)
        wk.set_grid(dx, dy, ZC, dZ)
        wk.set_bc(BoundaryConditionSet(*case.faces))
        lam = np.ones(shape, order="F")
        lam[1:-1, 1:-1, 1:-1] = case.lam
        hs = np.zeros(shape, order="F")
        hs[1:-1, 1:-1, 1:-1] = case.Q
        wk.upload("lam", lam)
        wk.upload("rho", np.full(shape, case.rho, order="F"))
        wk.upload("cp", np.full(shape, case.cp, order="F"))
        wk.upload("hsrc", hs)
        wk.fill("theta", case.theta0)
        wk.apply_bc()
        wk.fill("b", 0.0)
        wk.calc_rhs(case.dt, case.steady)
        if case.start_is_solution:
            res = wk.calc_rk("pcg_r", "theta", "b", case.dt, case.steady)
            extra = (res, np.linalg.norm(wk.download("b")))
            th = None
        else:
            conv, itr, res0 = wk.solve(L.SOLVER_PBICGSTAB, case.dt, 1e-14, 20000, L.SMOOTHER_GS, case.steady)
            assert conv
            th = wk.download("theta")[1:-1, 1:-1, 1:-1]
            extra = None
    CF.check(case, th, (dx, dy, ZC), extra)
