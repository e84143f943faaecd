"""ctypes wrapper of the C oracle (liborc.so).  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
module, always as the checker or the CPU baseline, never as the product path.
Arrays follow the reference layout: numpy (MX, MY, MZ) in Fortran order.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liborc.so")
EXE = os.path.join(HERE, "orc_q3d")

_vp, _i, _d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


class Grid(ctypes.Structure):
    _fields_ = [("MX", _i), ("MY", _i), ("MZ", _i), ("dx", _d), ("dy", _d), ("dt", _d), ("ZC", _vp),
                ("dZ", _vp), ("HC", _d * 6), ("steady", _i)]


class BC(ctypes.Structure):
    _fields_ = [("type", _i), ("temperature", _d), ("heat_flux", _d), ("htc", _d), ("ambient", _d)]


class Work(ctypes.Structure):
    _fields_ = [(n, _vp) for n in ("theta", "b", "mask", "rho", "lam", "cp", "p", "p_", "r", "r0", "q", "s",
                                   "s_", "t_", "hsrc")]


FIELDS = ["theta", "b", "mask", "rho", "lam", "cp", "p", "p_", "r", "r0", "q", "s", "s_", "t_", "hsrc"]


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = ctypes.CDLL(LIB)
        lib.orc_pitch.restype = _d
        lib.orc_pitch.argtypes = [_i]
        lib.orc_lamf.restype = _d
        lib.orc_lamf.argtypes = [_d, _d, _d, _d]
        for name in ("orc_calc_rk", "orc_rbsor_core", "orc_fdot1", "orc_fdot2"):
            getattr(lib, name).restype = _d
        lib.orc_fdot1.argtypes = [_vp, _i, _i, _i]
        lib.orc_fdot2.argtypes = [_vp, _vp, _i, _i, _i]
        lib.orc_bicg1.argtypes = [_vp, _vp, _vp, _d, _d, _i, _i, _i]
        lib.orc_triad.argtypes = [_vp, _vp, _vp, _d, _i, _i, _i]
        lib.orc_bicg2.argtypes = [_vp, _vp, _vp, _d, _d, _i, _i, _i]
        lib.orc_fill_id = lib.orc_fillID
        lib.orc_fillID.argtypes = [_vp, _i, _i, _i, _d, _d, _vp]
        lib.orc_genZ.argtypes = [_i, _vp, _vp, _vp]
        lib.orc_set_properties.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_size_t]
        lib.orc_heat_src.argtypes = [_vp, _vp, _i, _i, _i]
        lib.orc_apply_bc.argtypes = [_vp, _vp, _vp, _vp, _vp, _i, _i, _i, _vp]
        lib.orc_set_threads.argtypes = [_i]
        lib.orc_set_sum_mode.argtypes = [_i]
        lib.orc_exact_sum.restype = _d
        lib.orc_exact_sum.argtypes = [_vp, ctypes.c_size_t]
        _lib = lib
    return _lib


def _p(a):
    """Data pointer; 3-D arrays must be in the reference's column-major layout."""
    if a.ndim == 3 and not a.flags.f_contiguous:
        raise ValueError("3-D arrays must be Fortran-ordered (column-major, i fastest)")
    if not (a.flags.c_contiguous or a.flags.f_contiguous):
        raise ValueError("array must be contiguous")
    return ctypes.c_void_p(a.ctypes.data)


def set_threads(n):
    load().orc_set_threads(int(n))


SEQUENTIAL, EXACT = 0, 1


def set_sum_mode(mode):
    """0: the reference's sequential running sums; 1: exactly rounded sums."""
    load().orc_set_sum_mode(int(mode))


class sum_mode:
    """Context manager that switches the oracle's grid-sum mode."""

    def __init__(self, mode):
        self.mode = mode

    def __enter__(self):
        self.prev = load().orc_get_sum_mode()
        set_sum_mode(self.mode)

    def __exit__(self, *a):
        set_sum_mode(self.prev)


def exact_sum(t):
    t = np.ascontiguousarray(t, dtype=np.float64)
    return load().orc_exact_sum(_p(t), t.size)


def pitch(n):
    return load().orc_pitch(n)


def genZ(nk):
    Z, ZC, dZ = np.zeros(nk + 3), np.zeros(nk + 2), np.zeros(nk + 2)
    if load().orc_genZ(nk, _p(Z), _p(ZC), _p(dZ)) != 0:
        raise ValueError("nk must be a multiple of 30")
    return Z, ZC, dZ


def fill_id(MX, MY, MZ, dx, dy, Z):
    ID = np.zeros((MX, MY, MZ), dtype=np.uint8, order="F")
    rc = load().orc_fillID(_p(ID), MX, MY, MZ, dx, dy, _p(np.ascontiguousarray(Z)))
    if rc != 0:
        raise RuntimeError("fillID failed %d" % rc)
    return ID


def set_properties(ID):
    lam, rho, cp = (np.empty(ID.shape, order="F") for _ in range(3))
    load().orc_set_properties(_p(lam), _p(rho), _p(cp), _p(ID), ID.size)
    return lam, rho, cp


def heat_src(ID):
    hs = np.zeros(ID.shape, order="F")
    load().orc_heat_src(_p(hs), _p(ID), *ID.shape)
    return hs


def mode3_bc():
    bc = (BC * 6)()
    load().orc_mode3_bc(bc)
    return bc


def bc_from_set(bc_set):
    bc = (BC * 6)()
    for f, b in enumerate(bc_set.faces()):
        bc[f] = BC(b.type, b.temperature, b.heat_flux, b.heat_transfer_coefficient, b.ambient_temperature)
    return bc


def hc_of(bc):
    HC = (_d * 6)()
    load().orc_set_bc_coef(bc, HC)
    return np.array(HC[:])


class Case:
    """WorkBuffers + metrics of the oracle, set up like q3d() (src/heat3ds.jl:193-230)."""

    def __init__(self, MX, MY, MZ, dx, dy, ZC, dZ, dt=10000.0, steady=True, bc=None):
        self.shape = (MX, MY, MZ)
        self.arr = {}
        for n in FIELDS:
            self.arr[n] = np.zeros(self.shape, order="F")
        self.arr["mask"][:] = 1.0
        self.arr["lam"][:] = 1.0
        self.arr["cp"][:] = 1.0
        self.ZC = np.ascontiguousarray(ZC, dtype=np.float64)
        self.dZ = np.ascontiguousarray(dZ, dtype=np.float64)
        self.bc = bc if bc is not None else mode3_bc()
        self.grid = Grid(MX, MY, MZ, dx, dy, dt, _p(self.ZC), _p(self.dZ), (_d * 6)(*hc_of(self.bc)),
                         1 if steady else 0)
        self.work = Work(*[_p(self.arr[n]) for n in FIELDS])

    def set_steady(self, steady, dt=None):
        self.grid.steady = 1 if steady else 0
        if dt is not None:
            self.grid.dt = dt

    def apply_bc(self):
        a = self.arr
        load().orc_apply_bc(_p(a["theta"]), _p(a["lam"]), _p(a["rho"]), _p(a["cp"]), _p(a["mask"]),
                            *self.shape, self.bc)

    def cal_rhs(self):
        load().orc_cal_rhs(ctypes.byref(self.work), ctypes.byref(self.grid), self.bc)

    def calc_ax(self, ax, x):
        load().orc_calc_ax(_p(self.arr[ax]) if isinstance(ax, str) else _p(ax),
                           _p(self.arr[x]) if isinstance(x, str) else _p(x),
                           ctypes.byref(self.work), ctypes.byref(self.grid))

    def calc_rk(self, r, th, b):
        f = load().orc_calc_rk
        f.argtypes = [_vp, _vp, _vp, _vp, _vp]
        return f(_p(self.arr[r]), _p(self.arr[th]), _p(self.arr[b]), ctypes.byref(self.work),
                 ctypes.byref(self.grid))

    def rbsor_core(self, x, b, color, omega=1.0):
        f = load().orc_rbsor_core
        f.argtypes = [_vp, _vp, _vp, _vp, _d, _i]
        return f(_p(self.arr[x]), _p(self.arr[b]), ctypes.byref(self.work), ctypes.byref(self.grid), omega, color)

    def res_sor(self, x, b, omega):
        """resSOR (src/NonUniform.jl:680-744)"""
        f = load().orc_res_sor
        f.argtypes = [_vp, _vp, _vp, _vp, _d]
        f.restype = _d
        return f(_p(self.arr[x]), _p(self.arr[b]), ctypes.byref(self.work), ctypes.byref(self.grid), omega)

    def rbsor(self, x, b, omega):
        """rbsor! (src/NonUniform.jl:936-957)"""
        f = load().orc_rbsor
        f.argtypes = [_vp, _vp, _vp, _vp, _d]
        f.restype = _d
        return f(_p(self.arr[x]), _p(self.arr[b]), ctypes.byref(self.work), ctypes.byref(self.grid), omega)

    def sor(self, x, b, omega):
        """sor! (src/NonUniform.jl:762-826): one lexicographic sweep in place"""
        f = load().orc_sor
        f.argtypes = [_vp, _vp, _vp, _vp, _d]
        f.restype = _d
        return f(_p(self.arr[x]), _p(self.arr[b]), ctypes.byref(self.work), ctypes.byref(self.grid), omega)

    def solve_sor(self, omega, tol=1e-6, maxItr=20000):
        """solveSOR! (src/NonUniform.jl:1017-1050) on θ, b: (converged, itr, res0, history)"""
        f = load().orc_solve_sor
        f.argtypes = [_vp, _vp, _d, _d, _i, ctypes.POINTER(_i), ctypes.POINTER(_d), _vp, _i]
        f.restype = _i
        itr, res0 = _i(), _d()
        hist = np.zeros(maxItr)
        conv = f(ctypes.byref(self.work), ctypes.byref(self.grid), omega, tol, maxItr, ctypes.byref(itr),
                 ctypes.byref(res0), _p(hist), maxItr)
        return bool(conv), itr.value, res0.value, hist[: itr.value].copy()

    def precond(self, xx, bb, smoother=1):
        f = load().orc_precond
        f.argtypes = [_vp, _vp, _vp, _vp, _i]
        f(_p(self.arr[xx]), _p(self.arr[bb]), ctypes.byref(self.work), ctypes.byref(self.grid), smoother)

    def solve(self, solver="pbicgstab", tol=1e-6, maxItr=20000, smoother=1, verbose=0):
        fn = load().orc_cg if solver == "cg" else load().orc_pbicgstab
        fn.argtypes = [_vp, _vp, _d, _i, _i, _i, ctypes.POINTER(_i), ctypes.POINTER(_d), _vp, _i]
        itr, res0 = _i(), _d()
        hist = np.zeros(maxItr)
        conv = fn(ctypes.byref(self.work), ctypes.byref(self.grid), tol, maxItr, smoother, verbose,
                  ctypes.byref(itr), ctypes.byref(res0), _p(hist), maxItr)
        return bool(conv), itr.value, res0.value, hist[: itr.value].copy()


def modelA_case(NX, NY, NZ, steady=True, dt=10000.0):
    """Oracle WorkBuffers after q3d()'s set-up and main()'s HeatSrc! (no RHS yet)."""
    MX, MY, MZ = NX + 2, NY + 2, NZ + 2
    dx, dy = pitch(NX), pitch(NY)
    Z, ZC, dZ = genZ(NZ)
    ID = fill_id(MX, MY, MZ, dx, dy, Z)
    c = Case(MX, MY, MZ, dx, dy, ZC, dZ, dt=dt, steady=steady)
    lam, rho, cp = set_properties(ID)
    c.arr["lam"][:] = lam
    c.arr["rho"][:] = rho
    c.arr["cp"][:] = cp
    c.arr["theta"][:] = 300.0
    c.apply_bc()
    c.arr["hsrc"][:] = heat_src(ID)
    c.ID, c.Z, c.dx, c.dy = ID, Z, dx, dy
    return c
