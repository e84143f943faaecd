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


def test_sor_lexicographic_sweep_is_the_matrix_splitting(oracle, case):
    """sor! (src/NonUniform.jl:762-826) as a matrix iteration: with A the
    probed operator (diagonal -dd, off-diagonal conductances), L its part
    toward neighbours earlier in the (k, j, i) loop order and U the rest, a
    sweep solves (D - ωL) x' = ((1-ω)D + ωU) x + ω(s_bc - b), s_bc the halo
    neighbours' contribution; its residual sum is Σ r², r = (dd + ω·c_minus)
    ·dp/ω with c_minus the conductances toward the x-, y- and z- neighbours
    (halo ones included).  A red-black or any other order, a dropped ω or a
    wrong residual term fails it."""
    import math
    from oracle import adjoint as OA
    from scipy.sparse import diags, tril, triu
    from scipy.sparse.linalg import spsolve_triangular
    c, exact, th0 = case
    A = OA.probe_matrix(c).tocsr()
    dd = -A.diagonal()
    Lc = tril(A, k=-1).tocsr()
    Uc = triu(A, k=1).tocsr()
    shape = c.shape
    inner = (slice(1, -1),) * 3

    def interior(a):
        return a[inner].ravel(order="F")

    # halo contributions: CalcAX! of θ with its interior zeroed (ss of the halo cells)
    hb = th0.copy(order="F")
    hb[inner] = 0.0
    ax = np.zeros(shape, order="F")
    c.calc_ax(ax, hb)
    s_bc = interior(ax)
    # conductances toward the minus-side halo neighbours: ones on the x-, y-, z- halo planes
    ones = np.zeros(shape, order="F")
    ones[0, :, :] = 1.0
    ones[:, 0, :] = 1.0
    ones[:, :, 0] = 1.0
    ones[inner] = 0.0
    c.calc_ax(ax, ones)
    c_minus = np.asarray(Lc.sum(axis=1)).ravel() + interior(ax)
    b = interior(c.arr["b"])
    x0 = interior(th0)
    for omega in (1.0, 1.6):
        D = diags(dd)
        rhs = ((1.0 - omega) * D + omega * Uc) @ x0 + omega * (s_bc - b)
        x1 = spsolve_triangular((D - omega * Lc).tocsr(), rhs, lower=True)
        dp = (x1 - x0) / omega
        r = (dd + omega * c_minus) * dp / omega
        expect = math.sqrt(math.fsum(r * r))
        c.arr["theta"][:] = th0
        with oracle.sum_mode(1):
            res = c.sor("theta", "b", omega)
        got = interior(c.arr["theta"])
        assert np.max(np.abs(got - x1)) <= 1e-12 * np.max(np.abs(x1))
        assert abs(res - expect) <= 1e-9 * expect
        # the reference's sequential sum of the same terms differs only by rounding
        c.arr["theta"][:] = th0
        res_seq = c.sor("theta", "b", omega)
        assert abs(res_seq - res) <= 1e-12 * res


def test_sor_fixed_point_and_convergence(oracle, case):
    """The exact solution is a fixed point of sor! (its residual vanishes),
    and repeated sweeps reach it."""
    c, exact, th0 = case
    c.arr["theta"][:] = exact
    res = c.sor("theta", "b", 1.5)
    assert np.max(np.abs(c.arr["theta"] - exact)) <= 1e-10 * np.max(np.abs(exact))
    c.arr["theta"][:] = th0
    res0 = c.sor("theta", "b", 1.5)
    assert res <= 1e-9 * res0
    for _ in range(2000):
        if c.sor("theta", "b", 1.5) < 1e-12 * res0:
            break
    assert np.max(np.abs(c.arr["theta"] - exact)) < 1e-8
