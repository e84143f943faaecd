"""Host-side mirror of the reference's Heat3ds driver and solver interface,
running the linear-solve path on an MI355X through libh3d.

Reference interfaces mirrored (paths into the kenoogl/H2 tree):
  WorkBuffers                          src/common.jl:49-89
  BoundaryCondition(Set), *_bc,
  create_boundary_conditions,
  set_BC_coef                          src/boundary_conditions.jl:15-83, 360-394
  set_mode3_bc_parameters, HeatSrc!    src/heat3ds.jl:27-64
  calRHS!, PBiCGSTAB!, CG!             src/NonUniform.jl:57-350
  main, q3d                            src/heat3ds.jl:107-281

Arrays are numpy arrays of shape (MX, MY, MZ) in Fortran order, i.e. the
memory layout of the reference's Julia arrays; a[i, j, k] here is the
reference's a[i+1, j+1, k+1].  The device copies are owned by a
WorkBuffers object; only scalars cross per solver iteration.
"""
import ctypes
import sys
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L

Q_SRC = 1.6e11     # src/common.jl:16
ITR_MAX = 20_000   # default maxItr of PBiCGSTAB!/CG! (src/NonUniform.jl:122,255)


# ---------------------------------------------------------------------------
# boundary conditions (src/boundary_conditions.jl:15-83)
# ---------------------------------------------------------------------------
@dataclass
class BoundaryCondition:
    type: int
    temperature: float = 0.0
    heat_flux: float = 0.0
    heat_transfer_coefficient: float = 0.0
    ambient_temperature: float = 0.0


def isothermal_bc(temperature):
    return BoundaryCondition(L.ISOTHERMAL, temperature, 0.0, 0.0, 0.0)


def heat_flux_bc(heat_flux):
    return BoundaryCondition(L.HEAT_FLUX, 0.0, heat_flux, 0.0, 0.0)


def adiabatic_bc():
    return BoundaryCondition(L.HEAT_FLUX, 0.0, 0.0, 0.0, 0.0)


def convection_bc(h, t_amb):
    return BoundaryCondition(L.CONVECTION, 0.0, 0.0, h, t_amb)


@dataclass
class BoundaryConditionSet:
    x_minus: BoundaryCondition
    x_plus: BoundaryCondition
    y_minus: BoundaryCondition
    y_plus: BoundaryCondition
    z_minus: BoundaryCondition
    z_plus: BoundaryCondition

    def faces(self):
        return [self.x_minus, self.x_plus, self.y_minus, self.y_plus, self.z_minus, self.z_plus]


def create_boundary_conditions(xm, xp, ym, yp, zm, zp):
    return BoundaryConditionSet(xm, xp, ym, yp, zm, zp)


def set_mode3_bc_parameters():
    """Mode 3 (src/heat3ds.jl:27-46): sides and top convective, h = 5 W/m^2K,
    ambient 300 K; bottom isothermal at the PCB temperature 300 K."""
    th_amb, th_pcb, ht_top, ht_side = 300.0, 300.0, 5.0, 5.0
    return create_boundary_conditions(convection_bc(ht_side, th_amb), convection_bc(ht_side, th_amb),
                                      convection_bc(ht_side, th_amb), convection_bc(ht_side, th_amb),
                                      isothermal_bc(th_pcb), convection_bc(ht_top, th_amb))


def set_BC_coef(bc_set):
    """HC[6]: h on convective faces, else 0 (src/boundary_conditions.jl:360-394)."""
    return np.array([bc.heat_transfer_coefficient if bc.type == L.CONVECTION else 0.0
                     for bc in bc_set.faces()])


# ---------------------------------------------------------------------------
# device work buffers
# ---------------------------------------------------------------------------
def _check(ctx, rc):
    if rc != L.OK:
        lib = L.load()
        msg = lib.h3d_last_error(ctx).decode() if ctx else lib.h3d_create_error().decode()
        raise L.H3DError(rc, msg)


class WorkBuffers:
    """Device-resident WorkBuffers (src/common.jl:49-89) of one h3d context.

    Fields are addressed by the reference's member names (theta, b, mask, rho,
    lam, cp, pcg_p, pcg_p_, pcg_r, pcg_r0, pcg_q, pcg_s, pcg_s_, pcg_t_, hsrc).
    """

    def __init__(self, mx, my, mz, device=0):
        self.lib = L.load()
        self.shape = (mx, my, mz)
        self.n = mx * my * mz
        ctx = ctypes.c_void_p()
        rc = self.lib.h3d_create(ctypes.byref(ctx), mx, my, mz, device)
        if rc != L.OK:
            raise L.H3DError(rc, self.lib.h3d_create_error().decode())
        self.ctx = ctx
        self.grid = None

    def close(self):
        if self.ctx:
            self.lib.h3d_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _fid(self, name):
        if isinstance(name, int):
            return name
        return L.FIELD_ID[name]

    def upload(self, name, arr):
        a = np.asarray(arr, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError("expected shape %s, got %s" % (self.shape, a.shape))
        a = np.asfortranarray(a)
        _check(self.ctx, self.lib.h3d_upload(self.ctx, self._fid(name), L.ptr(a), self.n))

    def download(self, name):
        a = np.empty(self.shape, dtype=np.float64, order="F")
        _check(self.ctx, self.lib.h3d_download(self.ctx, self._fid(name), L.ptr(a), self.n))
        return a

    def fill(self, name, value):
        _check(self.ctx, self.lib.h3d_fill(self.ctx, self._fid(name), float(value)))

    def copy(self, dst, src):
        _check(self.ctx, self.lib.h3d_copy(self.ctx, self._fid(dst), self._fid(src)))

    def set_grid(self, dx, dy, ZC, dZ):
        ZC = np.ascontiguousarray(ZC, dtype=np.float64)
        dZ = np.ascontiguousarray(dZ, dtype=np.float64)
        _check(self.ctx, self.lib.h3d_set_grid(self.ctx, float(dx), float(dy), L.ptr(ZC), L.ptr(dZ), len(ZC)))
        self.grid = (float(dx), float(dy), ZC.copy(), dZ.copy())

    def set_bc(self, bc_set):
        faces = bc_set.faces()
        t = np.array([f.type for f in faces], dtype=np.int32)
        T = np.array([f.temperature for f in faces])
        q = np.array([f.heat_flux for f in faces])
        h = np.array([f.heat_transfer_coefficient for f in faces])
        ta = np.array([f.ambient_temperature for f in faces])
        _check(self.ctx, self.lib.h3d_set_bc(self.ctx, L.ptr(t), L.ptr(T), L.ptr(q), L.ptr(h), L.ptr(ta)))

    def hc(self):
        out = np.zeros(6)
        _check(self.ctx, self.lib.h3d_get_hc(self.ctx, L.ptr(out)))
        return out

    def apply_bc(self):
        _check(self.ctx, self.lib.h3d_apply_bc(self.ctx))

    def calc_rhs(self, dt, is_steady):
        _check(self.ctx, self.lib.h3d_calc_rhs(self.ctx, float(dt), int(bool(is_steady))))

    def solve(self, solver, dt, tol, max_itr=ITR_MAX, smoother=L.SMOOTHER_GS, is_steady=True,
              precision=L.FP64, inner_tol=1e-4):
        conv, itr, res0 = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        fn = self.lib.h3d_cg if solver == L.SOLVER_CG else self.lib.h3d_pbicgstab
        rc = fn(self.ctx, float(dt), float(tol), int(max_itr), int(smoother), int(bool(is_steady)),
                int(precision), float(inner_tol), ctypes.byref(conv), ctypes.byref(itr), ctypes.byref(res0))
        _check(self.ctx, rc)
        return bool(conv.value), itr.value, res0.value

    def history(self):
        n = ctypes.c_int()
        _check(self.ctx, self.lib.h3d_get_history(self.ctx, None, 0, ctypes.byref(n)))
        out = np.zeros(n.value)
        if n.value:
            _check(self.ctx, self.lib.h3d_get_history(self.ctx, L.ptr(out), n.value, ctypes.byref(n)))
        return out

    def last_solve_ms(self):
        ms = ctypes.c_double()
        _check(self.ctx, self.lib.h3d_last_solve_ms(self.ctx, ctypes.byref(ms)))
        return ms.value

    def last_solve_floor(self):
        """True when the last FP32-refined solve converged by reaching the
        rounding floor of its FP64 residual rather than the tolerance."""
        f = ctypes.c_int()
        _check(self.ctx, self.lib.h3d_last_solve_floor(self.ctx, ctypes.byref(f)))
        return bool(f.value)

    # single operators -----------------------------------------------------
    def calc_ax(self, ax, x, dt=1.0, is_steady=True):
        _check(self.ctx, self.lib.h3d_calc_ax(self.ctx, self._fid(ax), self._fid(x), float(dt), int(bool(is_steady))))

    def calc_rk(self, r, theta, b, dt=1.0, is_steady=True):
        res = ctypes.c_double()
        _check(self.ctx, self.lib.h3d_calc_rk(self.ctx, self._fid(r), self._fid(theta), self._fid(b), float(dt),
                                              int(bool(is_steady)), ctypes.byref(res)))
        return res.value

    def residual_scale(self, theta, b, dt=1.0, is_steady=True):
        """||s||, s = |b| + sum|terms of A theta| per cell: 2**-53 times it is
        the rounding scale of the FP64 residual b - A theta."""
        out = ctypes.c_double()
        _check(self.ctx, self.lib.h3d_residual_scale(self.ctx, self._fid(theta), self._fid(b), float(dt),
                                                     int(bool(is_steady)), ctypes.byref(out)))
        return out.value

    def rbsor_color(self, x, b, color, dt=1.0, is_steady=True):
        _check(self.ctx, self.lib.h3d_rbsor_color(self.ctx, self._fid(x), self._fid(b), int(color), float(dt),
                                                  int(bool(is_steady))))

    def precond(self, xx, bb, smoother=L.SMOOTHER_GS, dt=1.0, is_steady=True):
        _check(self.ctx, self.lib.h3d_precond(self.ctx, self._fid(xx), self._fid(bb), int(smoother), float(dt),
                                              int(bool(is_steady))))

    def fdot1(self, x):
        out = ctypes.c_double()
        _check(self.ctx, self.lib.h3d_fdot1(self.ctx, self._fid(x), ctypes.byref(out)))
        return out.value

    def fdot2(self, x, y):
        out = ctypes.c_double()
        _check(self.ctx, self.lib.h3d_fdot2(self.ctx, self._fid(x), self._fid(y), ctypes.byref(out)))
        return out.value

    def bicg1(self, p, r, q, beta, omg):
        _check(self.ctx, self.lib.h3d_bicg1(self.ctx, self._fid(p), self._fid(r), self._fid(q), beta, omg))

    def triad(self, z, x, y, a):
        _check(self.ctx, self.lib.h3d_triad(self.ctx, self._fid(z), self._fid(x), self._fid(y), a))

    def bicg2(self, z, x, y, a, b):
        _check(self.ctx, self.lib.h3d_bicg2(self.ctx, self._fid(z), self._fid(x), self._fid(y), a, b))

    def set_coef_form(self, form):
        """L.COEF_AUTO (class tables when possible) or L.COEF_DIRECT."""
        _check(self.ctx, self.lib.h3d_set_option(self.ctx, L.OPT_COEF_FORM, int(form)))

    def set_layout(self, layout):
        """L.LAYOUT_AUTO (colour-split Krylov vectors with the table form) or
        L.LAYOUT_STANDARD (the reference's layout); results are identical."""
        _check(self.ctx, self.lib.h3d_set_option(self.ctx, L.OPT_LAYOUT, int(layout)))

    def set_operator_form(self, form):
        """L.FORM_REFERENCE (the reference's expression, default) or L.FORM_FLUX."""
        _check(self.ctx, self.lib.h3d_set_option(self.ctx, L.OPT_OPERATOR_FORM, int(form)))

    def set_sweep_chunk(self, chunk):
        """Red-black sweeps of the colour-split layout: 0 = one launch per
        colour, C >= 1 = both colours in one launch over z-chunks of C
        planes, -1 = chosen from the grid (default).  Bitwise identical."""
        _check(self.ctx, self.lib.h3d_set_option(self.ctx, L.OPT_SWEEP_CHUNK, int(chunk)))

    def set_resident(self, mode):
        """On grids whose iterate fits the CUs' LDS (when the sweep chunk is
        left at -1): 2 (default) = PBiCGSTAB and CG run every iteration of a
        solve in one launch of co-resident workgroups; 1 = the Gauss-Seidel
        preconditioner (with its neighbours in the iteration) runs as one
        such launch, three launches per PBiCGSTAB iteration; 0 = one launch
        per colour pass.  Bitwise identical."""
        _check(self.ctx, self.lib.h3d_set_option(self.ctx, L.OPT_RESIDENT, int(mode)))

    def resident_workgroups(self):
        """Workgroups of the last solve's resident preconditioner (0: none)."""
        v = ctypes.c_int()
        _check(self.ctx, self.lib.h3d_get_option(self.ctx, L.OPT_RESIDENT_WGS, ctypes.byref(v)))
        return v.value

    def set_sum_order(self, order):
        """L.SUM_EXACT (exactly rounded grid sums, default) or L.SUM_SEQUENTIAL
        (the reference's sequential (k, j, i) running sums: a solve then
        reproduces the reference's sequential run bit for bit, at the cost of
        one single-workgroup pass per sum)."""
        _check(self.ctx, self.lib.h3d_set_option(self.ctx, L.OPT_SUM_ORDER, int(order)))

    def num_classes(self):
        v = ctypes.c_int()
        _check(self.ctx, self.lib.h3d_get_option(self.ctx, L.OPT_NUM_CLASSES, ctypes.byref(v)))
        return v.value

    def time_kernel(self, which, reps, dt=1.0, is_steady=True):
        ms = ctypes.c_double()
        _check(self.ctx, self.lib.h3d_time_kernel(self.ctx, int(which), int(reps), float(dt), int(bool(is_steady)),
                                                  ctypes.byref(ms)))
        return ms.value

    # stand-alone RB-SOR (src/NonUniform.jl:680-1050) ---------------------
    def res_sor(self, x, b, omega, dt=1.0, is_steady=True):
        """resSOR of field x against field b."""
        r = ctypes.c_double()
        _check(self.ctx, self.lib.h3d_res_sor(self.ctx, self._fid(x), self._fid(b), float(omega), float(dt),
                                              int(bool(is_steady)), ctypes.byref(r)))
        return r.value

    def rbsor(self, x, b, omega, dt=1.0, is_steady=True):
        """rbsor!: one red-black SOR sweep of x in place; the residual RMS."""
        r = ctypes.c_double()
        _check(self.ctx, self.lib.h3d_rbsor(self.ctx, self._fid(x), self._fid(b), float(omega), float(dt),
                                            int(bool(is_steady)), ctypes.byref(r)))
        return r.value

    def sor(self, x, b, omega, dt=1.0, is_steady=True):
        """sor! (src/NonUniform.jl:762-826): one lexicographic SOR sweep of x
        in place; returns sqrt(Σ r²)."""
        res = ctypes.c_double()
        _check(self.ctx, self.lib.h3d_sor(self.ctx, self._fid(x), self._fid(b), float(omega), float(dt),
                                          int(bool(is_steady)), ctypes.byref(res)))
        return res.value

    def solve_sor(self, omega, tol, max_itr=ITR_MAX, dt=1.0, is_steady=True):
        """solveSOR! on θ and b -> (isconverged, itr, res0)."""
        conv, itr, res0 = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(self.ctx, self.lib.h3d_rbsor_solve(self.ctx, float(dt), float(omega), float(tol), int(max_itr),
                                                  int(bool(is_steady)), ctypes.byref(conv), ctypes.byref(itr),
                                                  ctypes.byref(res0)))
        return bool(conv.value), itr.value, res0.value

    # sensitivity extension (no reference counterpart) ----------------------
    def objective(self, region, beta):
        """Smooth maximum J of θ over the interior cells where region != 0;
        keeps the seed ∂J/∂θ on the device."""
        r = np.asfortranarray(np.asarray(region, dtype=np.uint8))
        if r.shape != self.shape:
            raise ValueError("expected shape %s, got %s" % (self.shape, r.shape))
        J = ctypes.c_double()
        _check(self.ctx, self.lib.h3d_objective(self.ctx, r.ctypes.data_as(ctypes.c_void_p), self.n, float(beta),
                                                ctypes.byref(J)))
        return J.value

    def adjoint_solve(self, solver, dt, tol, max_itr=ITR_MAX, smoother=L.SMOOTHER_GS, is_steady=True):
        conv, itr, res0 = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(self.ctx, self.lib.h3d_adjoint_solve(self.ctx, int(solver), float(dt), float(tol), int(max_itr),
                                                    int(smoother), int(bool(is_steady)), ctypes.byref(conv),
                                                    ctypes.byref(itr), ctypes.byref(res0)))
        return bool(conv.value), itr.value, res0.value

    def gradients(self):
        """(dJ/dλ, dJ/dhsrc) per cell for the current θ and adjoint."""
        gl = np.zeros(self.shape, order="F")
        gh = np.zeros(self.shape, order="F")
        _check(self.ctx, self.lib.h3d_gradients(self.ctx, L.ptr(gl), L.ptr(gh), self.n))
        return gl, gh

    def adjoint_fields(self):
        """(μ, seed ∂J/∂θ) as device-held."""
        mu = np.zeros(self.shape, order="F")
        seed = np.zeros(self.shape, order="F")
        _check(self.ctx, self.lib.h3d_get_adjoint(self.ctx, L.ptr(mu), L.ptr(seed), self.n))
        return mu, seed

    def adjoint_step(self, solver, dt, tol, weight, chain, max_itr=ITR_MAX, smoother=L.SMOOTHER_GS,
                     precision=L.FP64):
        """Aᵀμ_n = weight·seed + D μ_{n+1} (chain) on the transient operator."""
        conv, itr, res0 = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(self.ctx, self.lib.h3d_adjoint_step(self.ctx, int(solver), float(dt), float(tol), int(max_itr),
                                                   int(smoother), int(precision), float(weight), int(bool(chain)),
                                                   ctypes.byref(conv), ctypes.byref(itr), ctypes.byref(res0)))
        return bool(conv.value), itr.value, res0.value

    def accumulate_gradients(self, w_lambda, w_hsrc):
        _check(self.ctx, self.lib.h3d_gradients_accumulate(self.ctx, float(w_lambda), float(w_hsrc)))

    def accumulated_gradients(self, reset=False):
        """(Σ w·dJ/dλ, Σ w·dJ/dhsrc) accumulated on the device."""
        gl = np.zeros(self.shape, order="F")
        gh = np.zeros(self.shape, order="F")
        _check(self.ctx, self.lib.h3d_gradients_accumulated(self.ctx, L.ptr(gl), L.ptr(gh), self.n,
                                                            int(bool(reset))))
        return gl, gh

    def reset_gradients(self):
        _check(self.ctx, self.lib.h3d_gradients_accumulated(self.ctx, None, None, self.n, 1))

    # device-resident field copies (checkpoints) ------------------------------
    def store(self, name, slot):
        _check(self.ctx, self.lib.h3d_store_field(self.ctx, self._fid(name), int(slot)))

    def load(self, name, slot, scale=1.0):
        _check(self.ctx, self.lib.h3d_load_field(self.ctx, self._fid(name), int(slot), float(scale)))

    def release_slots(self):
        _check(self.ctx, self.lib.h3d_release_slots(self.ctx))

    def mem_info(self):
        """(bytes in use on the device, device total)."""
        used, total = ctypes.c_size_t(), ctypes.c_size_t()
        _check(self.ctx, self.lib.h3d_mem_info(self.ctx, ctypes.byref(used), ctypes.byref(total)))
        return used.value, total.value


# ---------------------------------------------------------------------------
# host model builder (libh3d's h3d_model_* functions)
# ---------------------------------------------------------------------------
def pitch(n):
    """dx = round(1.2e-3/n, digits=8) (src/heat3ds.jl:197-200)."""
    d = ctypes.c_double()
    if L.load().h3d_model_pitch(int(n), ctypes.byref(d)) != 0:
        raise ValueError("bad cell count")
    return d.value


def genZ(nk):
    """(Z[nk+3], ZC[nk+2], dZ[nk+2]) as genZ! (src/Zcoord.jl:133-166); nk = 30*r."""
    Z, ZC, dZ = np.zeros(nk + 3), np.zeros(nk + 2), np.zeros(nk + 2)
    if L.load().h3d_model_genz(int(nk), L.ptr(Z), L.ptr(ZC), L.ptr(dZ)) != 0:
        raise ValueError("NZ must be a positive multiple of 30 (Zcase2! defines 30 cells)")
    return Z, ZC, dZ


def fill_id(MX, MY, MZ, dx, dy, Z, threads=0):
    """modelA.fillID! over the whole (MX,MY,MZ) array (src/modelA.jl:545-556)."""
    ID = np.zeros((MX, MY, MZ), dtype=np.uint8, order="F")
    Z = np.ascontiguousarray(Z, dtype=np.float64)
    rc = L.load().h3d_model_fill_id(L.ptr(ID), MX, MY, MZ, dx, dy, L.ptr(Z), int(threads))
    if rc != 0:
        raise RuntimeError("fillID failed (%d): geometry outside the grid" % rc)
    return ID


def set_properties(ID):
    """modelA.setProperties! (src/modelA.jl:499-542) -> (λ, ρ, cp)."""
    lam = np.empty(ID.shape, order="F")
    rho = np.empty(ID.shape, order="F")
    cp = np.empty(ID.shape, order="F")
    L.load().h3d_model_properties(L.ptr(lam), L.ptr(rho), L.ptr(cp), L.ptr(ID), ID.size)
    return lam, rho, cp


def heat_src(ID, q_src=Q_SRC):
    """HeatSrc! (src/heat3ds.jl:55-64)."""
    hs = np.zeros(ID.shape, order="F")
    MX, MY, MZ = ID.shape
    L.load().h3d_model_heat_src(L.ptr(hs), L.ptr(ID), MX, MY, MZ, float(q_src))
    return hs


def power_map(ID, dx, dy, dZ, seed=20251203, total_watts=2.0):
    """Seeded synthetic power-grid map (log-normal cells x 16 hotspots, 2 W)."""
    hs = np.zeros(ID.shape, order="F")
    MX, MY, MZ = ID.shape
    dZ = np.ascontiguousarray(dZ, dtype=np.float64)
    rc = L.load().h3d_model_power_map(L.ptr(hs), L.ptr(ID), MX, MY, MZ, dx, dy, L.ptr(dZ), int(seed),
                                      float(total_watts))
    if rc != 0:
        raise RuntimeError("power map failed (%d)" % rc)
    return hs


@dataclass
class Model:
    NX: int
    NY: int
    NZ: int
    dx: float
    dy: float
    Z: np.ndarray
    ZC: np.ndarray
    dZ: np.ndarray
    ID: np.ndarray
    lam: np.ndarray
    rho: np.ndarray
    cp: np.ndarray
    hsrc: np.ndarray
    bc_set: BoundaryConditionSet = field(default_factory=set_mode3_bc_parameters)

    @property
    def shape(self):
        return (self.NX + 2, self.NY + 2, self.NZ + 2)


def build_model(NX, NY, NZ, power="modelA", seed=20251203, threads=0):
    """The inputs q3d() builds before main(): pitch, Z metrics, modelA
    geometry and properties, heat source (modelA's Q_src on the power grid,
    or the seeded synthetic map with power='map'), BC mode 3."""
    MX, MY, MZ = NX + 2, NY + 2, NZ + 2
    dx, dy = pitch(NX), pitch(NY)
    Z, ZC, dZ = genZ(NZ)
    ID = fill_id(MX, MY, MZ, dx, dy, Z, threads)
    lam, rho, cp = set_properties(ID)
    hs = heat_src(ID) if power == "modelA" else power_map(ID, dx, dy, dZ, seed=seed)
    return Model(NX, NY, NZ, dx, dy, Z, ZC, dZ, ID, lam, rho, cp, hs)


def load_model(wk, m, theta_init=300.0):
    """Upload a model, set θ = θ_init and apply the boundary conditions on the
    device (src/heat3ds.jl:224-230)."""
    wk.set_grid(m.dx, m.dy, m.ZC, m.dZ)
    wk.set_bc(m.bc_set)
    wk.upload("lam", m.lam)
    wk.upload("rho", m.rho)
    wk.upload("cp", m.cp)
    wk.upload("hsrc", m.hsrc)
    wk.fill("theta", theta_init)
    wk.apply_bc()


# ---------------------------------------------------------------------------
# the reference's function-level interface
# ---------------------------------------------------------------------------
_SMOOTHERS = {"none": L.SMOOTHER_NONE, "gs": L.SMOOTHER_GS}


def smoother_selector(s):
    """src/NonUniform.jl:36-41: :none or :gs, anything else is an error."""
    key = s.lstrip(":") if isinstance(s, str) else s
    if key not in _SMOOTHERS:
        raise ValueError("Unsupported smoother: %s. Use :none or :gs" % (s,))
    return _SMOOTHERS[key]


def _check_par(par):
    if par != "hip":
        raise ValueError("Invalid parallel mode %r: this backend implements par=\"hip\"" % (par,))


def calRHS(wk, dh, dt, dZ, bc_set, qsrf=None, par="hip", is_steady=False):
    """calRHS! (src/NonUniform.jl:57-91) on the device copy of b."""
    _check_par(par)
    wk.set_bc(bc_set)
    wk.calc_rhs(dt, is_steady)


def _krylov(solver, wk, dh, dt, ZC, dZ, HC, tol, maxItr, smoother, par, verbose, is_steady, precision,
            inner_tol):
    _check_par(par)
    sm = smoother_selector(smoother)
    if wk.grid is None:
        wk.set_grid(dh[0], dh[1], ZC, dZ)
    if HC is not None and not np.array_equal(np.asarray(HC, dtype=float), wk.hc()):
        raise ValueError("HC does not match the boundary-condition set of this context")
    conv, itr, res0 = wk.solve(solver, dt, tol, maxItr, sm, is_steady, precision, inner_tol)
    if verbose:
        print("Initial residual = ", repr(res0))
        for k, r in enumerate(wk.history(), 1):
            print("%10d %24.14E" % (k, r))
        if conv and itr > 0:
            print("Converged at ", itr)
    return conv, itr, res0


def PBiCGSTAB(wk, dh, dt, ZC, dZ, HC=None, tol=1e-6, maxItr=ITR_MAX, smoother="none", par="hip",
              verbose=False, is_steady=False, precision=L.FP64, inner_tol=1e-4):
    """PBiCGSTAB! (src/NonUniform.jl:248-350) -> (isconverged, itr, res0)."""
    return _krylov(L.SOLVER_PBICGSTAB, wk, dh, dt, ZC, dZ, HC, tol, maxItr, smoother, par, verbose, is_steady,
                   precision, inner_tol)


def CG(wk, dh, dt, ZC, dZ, HC=None, tol=1e-6, maxItr=ITR_MAX, smoother="none", par="hip", verbose=False,
       is_steady=False, precision=L.FP64, inner_tol=1e-4):
    """CG! (src/NonUniform.jl:115-226) -> (isconverged, itr, res0)."""
    return _krylov(L.SOLVER_CG, wk, dh, dt, ZC, dZ, HC, tol, maxItr, smoother, par, verbose, is_steady,
                   precision, inner_tol)


@dataclass
class Q3DResult:
    theta: np.ndarray
    isconverged: bool
    itr: int
    res0: float
    history: np.ndarray
    solve_ms: float        # device time of the Krylov solve (HIP events)
    main_s: float          # host wall time of calRHS! + solve (the reference's tm)
    theta_min: float
    theta_max: float
    theta_l2: float


def main(dh, dt, wk, ZC, dZ, model, solver, smoother, bc_set, par="hip", is_steady=False, itr_tol=1e-6,
         verbose=True, precision=L.FP64, log=sys.stdout):
    """main() (src/heat3ds.jl:107-160) with nt = 1: RHS, solve, θ statistics."""
    smoother_sym = "gs" if smoother == "gs" else "none"
    t0 = time.perf_counter()
    calRHS(wk, dh, dt, dZ, bc_set, None, par, is_steady=is_steady)
    fn = CG if solver == "cg" else PBiCGSTAB
    conv, itr, res0 = fn(wk, dh, dt, ZC, dZ, set_BC_coef(bc_set), tol=itr_tol, smoother=smoother_sym,
                         par=par, verbose=verbose, is_steady=is_steady, precision=precision)
    main_s = time.perf_counter() - t0
    if not conv:
        print("Warning: Solver did not converge at step 1", file=sys.stderr)
    th = wk.download("theta")
    s = th[1:-1, 1:-1, 1:-1]
    mn, mx, l2 = float(s.min()), float(s.max()), float(np.linalg.norm(s.ravel()))
    if log is not None:
        print("%d %f : θmin=%e  θmax=%e  L2 norm of θ=%e" % (1, dt, mn, mx, l2), file=log)
    return Q3DResult(th, conv, itr, res0, wk.history(), wk.last_solve_ms(), main_s, mn, mx, l2)


def q3d(NX, NY, NZ, solver="sor", smoother="", epsilon=1.0e-6, par="hip", is_steady=False, verbose=True,
        precision=L.FP64, device=0, power="modelA", seed=20251203, output_dir=None):
    """q3d() (src/heat3ds.jl:169-281) with par="hip": builds the modelA case
    on the host, runs calRHS! and the Krylov solve on the GPU and returns a
    Q3DResult.  solver "cg" runs CG!, anything else PBiCGSTAB! (as the
    reference's main() does); smoother "gs" selects the red-black
    Gauss-Seidel preconditioner, anything else none.  With output_dir the
    reference's text outputs are written there (write_outputs)."""
    _check_par(par)
    if verbose:
        print("Analysis mode: %s" % ("Steady-state" if is_steady else "Transient"))
    m = build_model(NX, NY, NZ, power=power, seed=seed)
    dt = 10000.0  # src/heat3ds.jl:225
    wk = WorkBuffers(*m.shape, device=device)
    try:
        load_model(wk, m)
        if verbose:
            print("(%d, %d, %d)  Itr.ε= %g" % (m.shape + (epsilon,)))
        res = main((m.dx, m.dy, 1.0), dt, wk, m.ZC, m.dZ, m, solver, smoother, m.bc_set, par=par,
                   is_steady=is_steady, itr_tol=epsilon, verbose=verbose, precision=precision,
                   log=sys.stdout if verbose else None)
        if output_dir is not None:
            write_outputs(output_dir, res, m, solver, smoother, epsilon, dt)
        return res
    finally:
        wk.close()


def write_outputs(output_dir, res, m, solver, smoother, epsilon, dt):
    """The text outputs of q3d (src/heat3ds.jl:119-277): log.txt (header,
    residual lines, θ statistics), the convergence CSV of a Krylov solver
    (convergence_<solver>_<NX>x<NY>x<NZ>[_<smoother>].csv) and the two
    Z-line CSVs temp3Z_ctr.csv (x = y = 0.6 mm) and temp3Z_tsv.csv
    (x = y = 0.4 mm).  The plots are not drawn."""
    import os
    from . import output as O
    os.makedirs(output_dir, exist_ok=True)
    dh = (m.dx, m.dy, 1.0)
    log = os.path.join(output_dir, "log.txt")
    O.write_log(log, O.conditions(m.shape, dh, solver, smoother, epsilon),
                [(res.history, O.theta_stats_line(1, dt, res.theta))])
    if solver in ("pbicgstab", "cg"):
        name = "convergence_%s_%dx%dx%d" % (solver, m.NX, m.NY, m.NZ) + ("_%s" % smoother if smoother else "")
        conv = O.ConvergenceData(solver, smoother)
        O.parse_residuals_from_log(conv, log)
        O.export_convergence_csv(conv, os.path.join(output_dir, name + ".csv"))
    for xy, tag in ((0.6e-3, "ctr"), (0.4e-3, "tsv")):
        z, d = O.line_z(res.theta, dh, m.Z, xy, xy)
        O.export_zline_csv(z, d, os.path.join(output_dir, "temp3Z_%s" % tag))
