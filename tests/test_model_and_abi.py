"""CPU tests of the product's host side: the native model builder against
the oracle's restatement (bit-exact), the C ABI's exported symbols, and the
Python mirror of the reference's driver interface."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import h2_amd
from h2_amd import _lib as L
from h2_amd import heat3ds as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("nx,ny,nz", [(40, 40, 30), (240, 240, 30), (40, 64, 60), (17, 9, 30)])
def test_model_builder_bitwise_vs_oracle(oracle, nx, ny, nz):
    m = h2_amd.build_model(nx, ny, nz)
    c = oracle.modelA_case(nx, ny, nz)
    assert np.array_equal(m.ID, c.ID)
    assert np.array_equal(m.Z, c.Z) and np.array_equal(m.ZC, c.ZC) and np.array_equal(m.dZ, c.dZ)
    assert m.dx == c.dx and m.dy == c.dy
    lam, rho, cp = oracle.set_properties(c.ID)
    assert np.array_equal(m.lam, lam) and np.array_equal(m.rho, rho) and np.array_equal(m.cp, cp)
    assert np.array_equal(m.hsrc, c.arr["hsrc"])


def test_model_builder_thread_count_invariant():
    Z, ZC, dZ = H.genZ(30)
    a = H.fill_id(82, 82, 32, H.pitch(80), H.pitch(80), Z, threads=1)
    b = H.fill_id(82, 82, 32, H.pitch(80), H.pitch(80), Z, threads=7)
    assert np.array_equal(a, b)


def test_genz_rejects_other_counts():
    with pytest.raises(ValueError):
        H.genZ(45)


def test_power_map_properties():
    m = h2_amd.build_model(120, 120, 30)
    hs = H.power_map(m.ID, m.dx, m.dy, m.dZ, seed=7, total_watts=2.0)
    V = m.dx * m.dy * m.dZ[None, None, :]
    assert abs(float((hs * V).sum()) - 2.0) < 1e-12
    on = hs > 0
    assert np.array_equal(on, m.ID == 7)
    assert np.array_equal(hs, H.power_map(m.ID, m.dx, m.dy, m.dZ, seed=7))
    assert not np.array_equal(hs, H.power_map(m.ID, m.dx, m.dy, m.dZ, seed=8))
    # log-normal spread (sigma 0.5) plus 5x hotspots: max/median well above 1
    v = hs[on]
    assert v.max() / np.median(v) > 3.0


def header_functions():
    names = []
    for h in ("h3d.h", "h3d_model.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        names += re.findall(r"^\s*(?:int|const char \*)\s*(h3d_\w+)\s*\(", txt, re.M)
    return names


def test_every_declared_symbol_is_exported():
    declared = set(header_functions())
    assert declared == set(L.EXPORTS)
    lib = L.load()
    for name in declared:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], stdout=subprocess.PIPE, text=True).stdout
    exported = set(re.findall(r" T (h3d_\w+)", out))
    assert declared <= exported


def test_library_is_built_for_gfx950():
    """The embedded device code object targets gfx950 (MI355X)."""
    data = open(L.LIB_PATH, "rb").read()
    assert b"hipv4-amdgcn-amd-amdhsa--gfx950" in data


def test_create_without_device_reports_an_error():
    from tests.conftest import gpu_available
    if gpu_available():
        pytest.skip("a HIP device is present")
    lib = L.load()
    ctx = ctypes.c_void_p()
    rc = lib.h3d_create(ctypes.byref(ctx), 10, 10, 10, 0)
    assert rc != 0 and not ctx.value
    assert lib.h3d_create_error().decode()


def test_null_and_argument_checks():
    lib = L.load()
    assert lib.h3d_destroy(None) == 0
    assert lib.h3d_last_error(None) == b"null context"
    ctx = ctypes.c_void_p()
    assert lib.h3d_create(ctypes.byref(ctx), 2, 10, 10, 0) == L.ERR_ARG
    assert lib.h3d_create(None, 10, 10, 10, 0) == L.ERR_ARG


def test_set_bc_coef_and_mode3():
    bcs = H.set_mode3_bc_parameters()
    assert list(H.set_BC_coef(bcs)) == [5.0, 5.0, 5.0, 5.0, 0.0, 5.0]
    assert bcs.z_minus.type == L.ISOTHERMAL and bcs.z_minus.temperature == 300.0
    t, T, q, h, ta = (np.zeros(6, dtype=np.int32), np.zeros(6), np.zeros(6), np.zeros(6), np.zeros(6))
    L.load().h3d_model_mode3_bc(L.ptr(t), L.ptr(T), L.ptr(q), L.ptr(h), L.ptr(ta))
    for f, bc in enumerate(bcs.faces()):
        assert (t[f], T[f], q[f], h[f], ta[f]) == (bc.type, bc.temperature, bc.heat_flux,
                                                    bc.heat_transfer_coefficient, bc.ambient_temperature)


def test_smoother_and_par_selection():
    assert H.smoother_selector(":gs") == L.SMOOTHER_GS
    assert H.smoother_selector("none") == L.SMOOTHER_NONE
    with pytest.raises(ValueError):
        H.smoother_selector(":sor")
    with pytest.raises(ValueError):
        h2_amd.q3d(40, 40, 30, "pbicgstab", "gs", par="sequential")


def test_pointer_helper_rejects_c_order_fields():
    with pytest.raises(ValueError):
        L.ptr(np.zeros((3, 4, 5)))
    L.ptr(np.zeros((3, 4, 5), order="F"))
