"""GPU parity of every libh3d operator against the oracle, through the C ABI.

Element-wise operators (boundary set-up, RHS, CalcAX!, CalcRK!'s residual
field, one red-black colour, the preconditioner, BiCG1!/Triad!/BICG2!) must
be bitwise equal to the oracle: both perform the reference's IEEE
operations in the reference's order.  The operators run in both coefficient
forms of the library (per-face harmonic means of the λ field, and the
class-table form used when λ takes few distinct values), which must agree
bitwise.  Grid sums are exactly rounded: equal to the oracle's exact-sum
mode and within 1e-13 of the reference's sequential sums.
"""
import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="no HIP device")]

import h2_amd
from h2_amd import _lib as L
from h2_amd.heat3ds import (BoundaryConditionSet, WorkBuffers, convection_bc, heat_flux_bc, isothermal_bc,
                            set_mode3_bc_parameters)

ORC_NAME = {"theta": "theta", "b": "b", "lam": "lam", "rho": "rho", "cp": "cp", "hsrc": "hsrc",
            "pcg_p": "p", "pcg_p_": "p_", "pcg_r": "r", "pcg_r0": "r0", "pcg_q": "q", "pcg_s": "s",
            "pcg_s_": "s_", "pcg_t_": "t_"}


def mixed_bc():
    # every face type, and the z- isothermal face that copies λ into the halo
    return BoundaryConditionSet(convection_bc(5.0, 300.0), heat_flux_bc(1.5e4), isothermal_bc(310.0),
                                convection_bc(12.0, 290.0), isothermal_bc(300.0), heat_flux_bc(-2.0e3))


def make_pair(oracle, NX, NY, NZ, seed, bc_set, random_lam=True, steady=True, dt=0.25, form=L.COEF_AUTO):
    """Identical random inputs on the device and in the oracle.  random_lam:
    True = a different λ, ρ and cp in every cell (the direct form),
    "classes" = (λ, ρ, cp) drawn from 9 random materials (the table form),
    False = modelA's λ, ρ and cp (the table form)."""
    rng = np.random.default_rng(seed)
    MX, MY, MZ = NX + 2, NY + 2, NZ + 2
    shape = (MX, MY, MZ)
    dx, dy = h2_amd.pitch(NX), h2_amd.pitch(NY)
    Z, ZC, dZ = h2_amd.genZ(NZ)
    if random_lam == "classes":
        mat = rng.integers(0, 9, size=shape)
        lam = rng.uniform(0.3, 400.0, 9)[mat]
        rho = rng.uniform(1000.0, 9000.0, 9)[mat]
        cp = rng.uniform(300.0, 1100.0, 9)[mat]
    elif random_lam:
        lam = rng.uniform(0.3, 400.0, shape)
        rho = rng.uniform(1000.0, 9000.0, shape)
        cp = rng.uniform(300.0, 1100.0, shape)
    else:
        m = h2_amd.build_model(NX, NY, NZ)
        lam, rho, cp = m.lam, m.rho, m.cp
    th = rng.uniform(290.0, 350.0, shape)
    wk = WorkBuffers(*shape)
    wk.set_coef_form(form)
    wk.set_grid(dx, dy, ZC, dZ)
    wk.set_bc(bc_set)
    for name, a in (("lam", lam), ("rho", rho), ("cp", cp), ("theta", th)):
        wk.upload(name, a)
    c = oracle.Case(MX, MY, MZ, dx, dy, ZC, dZ, dt=dt, steady=steady, bc=oracle.bc_from_set(bc_set))
    c.arr["lam"][:] = lam
    c.arr["rho"][:] = rho
    c.arr["cp"][:] = cp
    c.arr["theta"][:] = th
    wk.apply_bc()
    c.apply_bc()
    return wk, c, rng


def put(wk, c, name, a):
    wk.upload(name, a)
    c.arr[ORC_NAME[name]][:] = a


def assert_bitwise(dev, ref, what):
    diff = np.flatnonzero(dev.ravel(order="F") != ref.ravel(order="F"))
    assert diff.size == 0, "%s: %d cells differ, first at flat index %d: %r vs %r" % (
        what, diff.size, diff[0], dev.ravel(order="F")[diff[0]], ref.ravel(order="F")[diff[0]])


SIZES = [(8, 6, 30), (40, 40, 30), (13, 70, 60)]


@pytest.mark.parametrize("size", SIZES)
def test_apply_bc_bitwise(oracle, size):
    wk, c, _ = make_pair(oracle, *size, seed=1, bc_set=mixed_bc())
    for name in ("theta", "lam", "rho", "cp"):
        assert_bitwise(wk.download(name), c.arr[name], name)
    assert_bitwise(wk.download("mask"), c.arr["mask"], "mask")
    wk.close()


@pytest.mark.parametrize("steady", [True, False])
@pytest.mark.parametrize("size", SIZES[:2])
def test_calc_rhs_bitwise(oracle, size, steady):
    wk, c, rng = make_pair(oracle, *size, seed=2, bc_set=mixed_bc(), steady=steady)
    shape = wk.shape
    put(wk, c, "hsrc", rng.uniform(0, 1.6e11, shape) * (rng.random(shape) < 0.3))
    put(wk, c, "b", rng.normal(0, 1e-5, shape))
    wk.calc_rhs(c.grid.dt, steady)
    c.cal_rhs()
    assert_bitwise(wk.download("b"), c.arr["b"], "b")
    wk.close()


FORMS = [L.COEF_AUTO, L.COEF_DIRECT]


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("steady", [True, False])
@pytest.mark.parametrize("random_lam", [True, "classes", False])
@pytest.mark.parametrize("size", SIZES)
def test_calc_ax_bitwise(oracle, size, random_lam, steady, form):
    wk, c, rng = make_pair(oracle, *size, seed=3, bc_set=mixed_bc(), random_lam=random_lam, steady=steady,
                           form=form)
    x = rng.normal(0, 1, wk.shape)  # halo values too: the operator reads them
    put(wk, c, "pcg_p_", x)
    wk.calc_ax("pcg_q", "pcg_p_", c.grid.dt, steady)
    c.calc_ax("q", "p_")
    assert_bitwise(wk.download("pcg_q"), c.arr["q"], "ax")
    # the table form is in use exactly when (λ, ρ·cp) has few distinct pairs
    expect_table = form == L.COEF_AUTO and random_lam is not True
    assert (wk.num_classes() > 0) == expect_table
    wk.close()


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("steady", [True, False])
@pytest.mark.parametrize("size", SIZES)
def test_calc_rk(oracle, size, steady, form):
    wk, c, rng = make_pair(oracle, *size, seed=4, bc_set=mixed_bc(), steady=steady, random_lam="classes", form=form)
    put(wk, c, "b", rng.normal(0, 1e-3, wk.shape))
    res_d = wk.calc_rk("pcg_r", "theta", "b", c.grid.dt, steady)
    res_o = c.calc_rk("r", "theta", "b")
    assert_bitwise(wk.download("pcg_r"), c.arr["r"], "r")
    assert abs(res_d - res_o) <= 1e-13 * abs(res_o)
    with oracle.sum_mode(oracle.EXACT):
        assert res_d == c.calc_rk("r", "theta", "b")
    wk.close()


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("color", [0, 1])
@pytest.mark.parametrize("size", SIZES)
def test_rbsor_color_bitwise(oracle, size, color, form):
    wk, c, rng = make_pair(oracle, *size, seed=5, bc_set=set_mode3_bc_parameters(), random_lam=False, form=form)
    put(wk, c, "pcg_p", rng.normal(0, 1e-4, wk.shape))
    put(wk, c, "pcg_p_", rng.normal(0, 1e-2, wk.shape))
    wk.rbsor_color("pcg_p_", "pcg_p", color)
    c.rbsor_core("p_", "p", color)
    assert_bitwise(wk.download("pcg_p_"), c.arr["p_"], "x after colour %d" % color)
    wk.close()


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("steady", [True, False])
@pytest.mark.parametrize("smoother", [0, 1])
@pytest.mark.parametrize("size", SIZES)
def test_precond_bitwise(oracle, size, smoother, steady, form):
    wk, c, rng = make_pair(oracle, *size, seed=6, bc_set=mixed_bc(), steady=steady, random_lam="classes", form=form)
    put(wk, c, "pcg_p", rng.normal(0, 1e-4, wk.shape))
    put(wk, c, "pcg_p_", np.zeros(wk.shape))
    wk.precond("pcg_p_", "pcg_p", smoother, c.grid.dt, steady)
    c.precond("p_", "p", smoother)
    assert_bitwise(wk.download("pcg_p_"), c.arr["p_"], "preconditioned vector")
    wk.close()


@pytest.mark.parametrize("size", SIZES)
def test_dots_exactly_rounded(oracle, size):
    """Fdot1/Fdot2 on the device are the exactly rounded sums of the rounded
    products (the oracle's exact-sum mode), and within 1e-13 of the
    reference's sequential sums."""
    wk, c, rng = make_pair(oracle, *size, seed=7, bc_set=mixed_bc())
    x, y = rng.normal(0, 1, wk.shape), rng.normal(0, 1, wk.shape)
    # widen the dynamic range so the sums need more than 53 bits
    x *= 10.0 ** rng.integers(-8, 8, wk.shape)
    put(wk, c, "pcg_r", x)
    put(wk, c, "pcg_r0", y)
    MX, MY, MZ = wk.shape
    O = oracle.load()
    with oracle.sum_mode(oracle.EXACT):
        d1 = O.orc_fdot1(oracle._p(c.arr["r"]), MX, MY, MZ)
        d2 = O.orc_fdot2(oracle._p(c.arr["r"]), oracle._p(c.arr["r0"]), MX, MY, MZ)
    s1 = O.orc_fdot1(oracle._p(c.arr["r"]), MX, MY, MZ)
    assert wk.fdot1("pcg_r") == d1
    assert wk.fdot2("pcg_r", "pcg_r0") == d2
    assert abs(s1 - d1) <= 1e-13 * d1
    wk.close()


@pytest.mark.parametrize("size", SIZES[:2])
def test_blas1_bitwise(oracle, size):
    wk, c, rng = make_pair(oracle, *size, seed=8, bc_set=mixed_bc())
    MX, MY, MZ = wk.shape
    O = oracle.load()
    P = oracle._p
    for nm in ("pcg_p", "pcg_r", "pcg_q", "pcg_s", "pcg_t_", "theta"):
        put(wk, c, nm, rng.normal(0, 1, wk.shape))
    wk.bicg1("pcg_p", "pcg_r", "pcg_q", 0.37, -1.9)
    O.orc_bicg1(P(c.arr["p"]), P(c.arr["r"]), P(c.arr["q"]), 0.37, -1.9, MX, MY, MZ)
    assert_bitwise(wk.download("pcg_p"), c.arr["p"], "BiCG1!")
    wk.triad("pcg_s", "pcg_q", "pcg_r", -0.61)
    O.orc_triad(P(c.arr["s"]), P(c.arr["q"]), P(c.arr["r"]), -0.61, MX, MY, MZ)
    assert_bitwise(wk.download("pcg_s"), c.arr["s"], "Triad!")
    wk.bicg2("theta", "pcg_p", "pcg_s", 1.3, 0.7)
    O.orc_bicg2(P(c.arr["theta"]), P(c.arr["p"]), P(c.arr["s"]), 1.3, 0.7, MX, MY, MZ)
    assert_bitwise(wk.download("theta"), c.arr["theta"], "BICG2!")
    wk.close()


def test_mask_contract():
    wk = WorkBuffers(6, 5, 32)
    m = wk.download("mask")
    assert m[1:-1, 1:-1, 1:-1].min() == 1.0 and m.sum() == 4 * 3 * 30
    wk.upload("mask", m)  # the post-BC mask is accepted
    bad = m.copy()
    bad[2, 2, 2] = 0.0
    with pytest.raises(L.H3DError) as e:
        wk.upload("mask", bad)
    assert e.value.code == L.ERR_UNSUPPORTED
    wk.close()


@pytest.mark.parametrize("form", FORMS)
def test_fill_and_copy_whole_arrays(oracle, form):
    """myfill!/mycopy! (src/common.jl:95-113) cover every cell, halo
    included; a copy into λ takes effect in the next operator (its class
    tables are rebuilt), as CalcAX! on the oracle with the same λ shows."""
    print("case fill/copy form %d" % form, flush=True)
    wk, c, rng = make_pair(oracle, *SIZES[0], seed=11, bc_set=mixed_bc(), random_lam="classes", form=form)
    a = rng.normal(0, 1, wk.shape)
    wk.upload("pcg_s", a)
    wk.copy("pcg_t_", "pcg_s")
    assert_bitwise(wk.download("pcg_t_"), a, "copy")
    wk.fill("pcg_t_", 2.5)
    assert np.all(wk.download("pcg_t_") == 2.5)
    # new materials through a copy into λ
    mat = rng.integers(0, 5, size=wk.shape)
    lam2 = rng.uniform(1.0, 200.0, 5)[mat]
    wk.upload("pcg_s", lam2)
    wk.copy("lam", "pcg_s")
    c.arr["lam"][:] = lam2
    wk.apply_bc()
    c.apply_bc()
    x = rng.normal(0, 1, wk.shape)
    put(wk, c, "pcg_p_", x)
    wk.calc_ax("pcg_q", "pcg_p_", c.grid.dt, True)
    c.calc_ax("q", "p_")
    assert_bitwise(wk.download("pcg_q"), c.arr["q"], "ax after copying λ")
    with pytest.raises(L.H3DError):
        wk.copy("mask", "pcg_s")
    with pytest.raises(L.H3DError):
        wk.fill("mask", 1.0)
    wk.close()
