"""GPU parity of the stand-alone RB-SOR (resSOR, rbsor!, solveSOR!) with the
oracle in its exactly rounded sum mode: resSOR and one rbsor! sweep's
residual and field, and whole solveSOR! runs (iteration count, residual
history, θ) bit for bit, on a small random case (to convergence, three
relaxation factors) and on the modelA stack (a fixed number of sweeps)."""
import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="no HIP device")]

import h2_amd
from h2_amd import _lib as L
from h2_amd.heat3ds import WorkBuffers, load_model, set_mode3_bc_parameters
from tests.test_oracle_sor import sor_case


def device_copy(c, bc_set):
    MX, MY, MZ = c.shape
    wk = WorkBuffers(MX, MY, MZ)
    wk.set_grid(c.grid.dx, c.grid.dy, c.ZC, c.dZ)
    wk.set_bc(bc_set)
    for name, key in (("lam", "lam"), ("rho", "rho"), ("cp", "cp"), ("theta", "theta"), ("b", "b"),
                      ("hsrc", "hsrc")):
        wk.upload(name, c.arr[key])
    return wk


@pytest.mark.parametrize("omega", [1.0, 1.5, 1.9])
def test_solve_sor_bitwise(oracle, omega):
    c = sor_case(oracle)
    wk = device_copy(c, set_mode3_bc_parameters())
    with oracle.sum_mode(1):
        assert wk.res_sor("theta", "b", omega) == c.res_sor("theta", "b", omega)
        conv_o, itr_o, res0_o, hist_o = c.solve_sor(omega, tol=1e-10, maxItr=5000)
    conv, itr, res0 = wk.solve_sor(omega, 1e-10, 5000)
    print("omega", omega, "itr", itr, itr_o, "res0", res0, res0_o)
    assert conv and conv_o and itr == itr_o and res0 == res0_o
    assert np.array_equal(wk.history(), hist_o)
    assert np.array_equal(wk.download("theta"), c.arr["theta"])
    wk.close()


def test_rbsor_sweeps_on_modelA_bitwise(oracle):
    c = oracle.modelA_case(24, 20, 30)
    c.arr["b"][:] = 0.0
    c.cal_rhs()
    m = h2_amd.build_model(24, 20, 30)
    wk = WorkBuffers(*m.shape)
    load_model(wk, m)
    wk.fill("b", 0.0)
    wk.calc_rhs(c.grid.dt, True)
    assert np.array_equal(wk.download("b"), c.arr["b"])
    with oracle.sum_mode(1):
        for sweep in range(3):
            r_dev = wk.rbsor("theta", "b", 1.7, c.grid.dt, True)
            r_orc = c.rbsor("theta", "b", 1.7)
            assert r_dev == r_orc
        assert np.array_equal(wk.download("theta"), c.arr["theta"])
        conv_o, itr_o, res0_o, hist_o = c.solve_sor(1.7, tol=1e-30, maxItr=40)
    conv, itr, res0 = wk.solve_sor(1.7, 1e-30, 40, c.grid.dt, True)
    assert not conv and not conv_o and itr == itr_o == 40 and res0 == res0_o
    assert np.array_equal(wk.history(), hist_o)
    assert np.array_equal(wk.download("theta"), c.arr["theta"])
    wk.close()


def test_sor_argument_errors():
    m = h2_amd.build_model(8, 8, 30)
    wk = WorkBuffers(*m.shape)
    load_model(wk, m)
    with pytest.raises(L.H3DError):
        wk.rbsor("theta", "theta", 1.0)
    with pytest.raises(L.H3DError):
        wk.solve_sor(0.0, 1e-6, 10)
    wk.close()


@pytest.mark.parametrize("omega", [1.0, 1.7])
def test_lexicographic_sor_bitwise(oracle, omega):
    """sor! on the device (a wavefront over the planes i + j + k = t) against
    the oracle's sequential sweep: θ bit for bit after three sweeps, and each
    sweep's residual bit for bit, exactly rounded and in the reference's
    sequential order, on the small random case and on the modelA stack."""
    c = sor_case(oracle)
    wk = device_copy(c, set_mode3_bc_parameters())
    for seq in (False, True):
        if seq:
            wk.set_sum_order(L.SUM_SEQUENTIAL)
        with oracle.sum_mode(0 if seq else 1):
            for sweep in range(3):
                r_dev = wk.sor("theta", "b", omega, c.grid.dt, True)
                r_orc = c.sor("theta", "b", omega)
                print("omega", omega, "seq", seq, "sweep", sweep, r_dev, r_orc)
                assert r_dev == r_orc
        assert np.array_equal(wk.download("theta"), c.arr["theta"])
    wk.close()


def test_lexicographic_sor_on_modelA_bitwise(oracle):
    c = oracle.modelA_case(24, 20, 30)
    c.arr["b"][:] = 0.0
    c.cal_rhs()
    m = h2_amd.build_model(24, 20, 30)
    with WorkBuffers(*m.shape) as wk:
        load_model(wk, m)
        wk.fill("b", 0.0)
        wk.calc_rhs(c.grid.dt, True)
        with oracle.sum_mode(1):
            for sweep in range(2):
                assert wk.sor("theta", "b", 1.5, c.grid.dt, True) == c.sor("theta", "b", 1.5)
        assert np.array_equal(wk.download("theta"), c.arr["theta"])
