"""The oracle's stand-alone RB-SOR (resSOR, rbsor!, solveSOR!,
src/NonUniform.jl:680-1050) pinned without the reference: solveSOR!
reaches the direct solution of the probed operator (scipy spsolve) for
several relaxation factors, resSOR vanishes at that solution, rbsor! with
ω = 1 is the Gauss-Seidel of the preconditioner (rbsor_core! colours 0 and
1), and over-relaxation converges in fewer iterations than ω = 1 on this
operator."""
import numpy as np
import pytest
from scipy.sparse.linalg import spsolve


def sor_case(oracle, NX=8, NY=7, NZ=6, seed=5):
    """A small case with random λ in [1, 10] on a uniform 0.1 mm grid, BC
    mode 3 and a random heat source: point SOR converges on it in a few
    hundred sweeps (the modelA stack, with its thin layers, needs tens of
    thousands)."""
    from h2_amd import output as HO
    Z, ZC, dZ = HO.genZ_from_faces(np.linspace(0.0, 1e-4 * NZ, NZ + 1))
    c = oracle.Case(NX + 2, NY + 2, NZ + 2, 1e-4, 1e-4, ZC, dZ, dt=1.0, steady=True)
    rng = np.random.default_rng(seed)
    c.arr["lam"][:] = rng.uniform(1.0, 10.0, c.shape)
    c.arr["theta"][:] = 300.0
    c.apply_bc()
    c.arr["hsrc"][:] = rng.uniform(0.0, 1e9, c.shape)
    c.arr["b"][:] = 0.0
    c.cal_rhs()
    return c


@pytest.fixture(scope="module")
def case(oracle):
    from oracle import adjoint as OA
    c = sor_case(oracle)
    A = OA.probe_matrix(c)
    c.calc_rk("r", "theta", "b")
    d = spsolve(A, c.arr["r"][1:-1, 1:-1, 1:-1].ravel(order="F"))
    exact = c.arr["theta"].copy(order="F")
    exact[1:-1, 1:-1, 1:-1] += d.reshape(exact[1:-1, 1:-1, 1:-1].shape, order="F")
    return c, exact, c.arr["theta"].copy(order="F")


@pytest.mark.parametrize("omega", [1.0, 1.5, 1.9])
def test_solve_sor_reaches_direct_solution(oracle, case, omega):
    c, exact, th0 = case
    c.arr["theta"][:] = th0
    conv, itr, res0, hist = c.solve_sor(omega, tol=1e-11, maxItr=20000)
    assert conv and res0 > 0 and hist[-1] < 1e-11
    assert np.max(np.abs(c.arr["theta"] - exact)) < 1e-8
    # the residual at the solution vanishes to rounding
    c.arr["p"][:] = exact
    assert c.res_sor("p", "b", omega) < 1e-9 * res0


def test_rbsor_unit_relaxation_is_the_gauss_seidel_sweep(oracle, case):
    c, exact, th0 = case
    c.arr["theta"][:] = th0
    c.arr["p"][:] = th0
    c.rbsor("theta", "b", 1.0)
    c.rbsor_core("p", "b", 0, 1.0)
    c.rbsor_core("p", "b", 1, 1.0)
    assert np.array_equal(c.arr["theta"], c.arr["p"])


def test_over_relaxation_converges_faster(oracle, case):
    c, exact, th0 = case
    its = []
    for omega in (1.0, 1.8):
        c.arr["theta"][:] = th0
        conv, itr, _, _ = c.solve_sor(omega, tol=1e-8, maxItr=20000)
        assert conv
        its.append(itr)
    assert its[1] < its[0]
