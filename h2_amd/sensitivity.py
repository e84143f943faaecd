"""Steady-state sensitivities of the hottest silicon temperature.

No counterpart in the reference: an extension of its solve path
(src/NonUniform.jl PBiCGSTAB!/CG! on the operator of CalcAX!, RHS of
calRHS!) by the adjoint method.

Objective: the smooth maximum of θ over the silicon cells (IDs 2 and 7 of
modelA, src/modelA.jl:96-131),
    J = m + log(Σ_{c∈Si} exp(β(θ_c - m))) / β,   m = max θ,   β = 10 K⁻¹.
Parameters:
  * 13 conductivity scales s: λ_c = s_g λ⁰_c for the interior cells of
    group g, where groups 0..11 are the z-slabs between modelA's marks
    zm0..zm12 (src/modelA.jl:10-30, by cell centre) for every material but
    copper, and group 12 is the copper (ID 1) cells.  The boundary halo
    follows λ through apply_boundary_conditions! (isothermal faces copy it).
  * the per-cell heat source hsrc (W/m³).
Gradients: solve A θ = b, then Aᵀ μ = ∂J/∂θ (A is symmetric), then
    dJ/dλ_c and dJ/dhsrc_c = -μ_c V_c on the device (h3d_gradients),
    dJ/ds_g = Σ_{c∈g} λ⁰_c dJ/dλ_c.
`central_difference` re-solves perturbed problems for a check.
"""
import numpy as np

from . import _lib as L
from .heat3ds import WorkBuffers, load_model

BETA = 10.0                 # 1 / (0.1 K)
SILICON_IDS = (2, 7)        # silicon, power grid (active layer)
COPPER_ID = 1
NGROUPS = 13
PG_DEPTH = 5.0e-6           # pg_dpth, src/modelA.jl
# z marks zm0..zm12 (src/modelA.jl:18-30)
ZM = np.array([0.0, 0.05e-3, 0.1e-3, 0.2e-3 - PG_DEPTH, 0.2e-3, 0.25e-3, 0.35e-3 - PG_DEPTH, 0.35e-3, 0.4e-3,
               0.5e-3 - PG_DEPTH, 0.5e-3, 0.55e-3, 0.6e-3])


def silicon_region(ID):
    """uint8 region of the interior silicon cells (IDs 2 and 7)."""
    r = np.isin(ID, SILICON_IDS).astype(np.uint8)
    r[0, :, :] = r[-1, :, :] = 0
    r[:, 0, :] = r[:, -1, :] = 0
    r[:, :, 0] = r[:, :, -1] = 0
    return np.asfortranarray(r)


def layer_groups(m):
    """Group index per cell: -1 on the halo, 12 for copper, else the z-slab."""
    NX, NY, NZ = m.NX, m.NY, m.NZ
    slab = np.clip(np.searchsorted(ZM, m.ZC[1:NZ + 1], side="right") - 1, 0, NGROUPS - 2)
    g = np.full(m.ID.shape, -1, dtype=np.int64, order="F")
    g[1:NX + 1, 1:NY + 1, 1:NZ + 1] = slab[None, None, :]
    g[(m.ID == COPPER_ID) & (g >= 0)] = NGROUPS - 1
    return g


def scaled_lambda(m, groups, scales):
    lam = np.array(m.lam, order="F", copy=True)
    for k in range(NGROUPS):
        if scales[k] != 1.0:
            sel = groups == k
            lam[sel] = m.lam[sel] * scales[k]
    return lam


def delta_J(theta_p, theta_m, region, beta=BETA):
    """J(θ+) - J(θ-) without cancellation: both sums with one shift,
    S+ - S- = Σ w- expm1(β(θ+ - θ-)), ΔJ = log1p((S+ - S-)/S-)/β."""
    sel = np.asarray(region, dtype=bool)
    tp, tm = theta_p[sel], theta_m[sel]
    ref = max(tp.max(), tm.max())
    wm = np.exp(beta * (tm - ref))
    return np.log1p(np.sum(wm * np.expm1(beta * (tp - tm))) / np.sum(wm)) / beta


def halo_lambda(m, lam):
    """λ with the halo apply_boundary_conditions! leaves: the interior
    neighbour's λ on isothermal faces, 0 on flux and convective faces
    (src/boundary_conditions.jl:94-313; faces x-,x+,y-,y+,z-,z+ in order)."""
    lam = np.array(lam, order="F", copy=True)
    sl = [(0, 1), (-1, -2)]
    for f, bc in enumerate(m.bc_set.faces()):
        ax, side = f // 2, f % 2
        dst, src = [slice(None)] * 3, [slice(None)] * 3
        dst[ax], src[ax] = sl[side][0], sl[side][1]
        lam[tuple(dst)] = lam[tuple(src)] if bc.type == L.ISOTHERMAL else 0.0
    return lam


def changed_box(lam_p, lam_0):
    """Per axis the index range [lo, hi) (halo coordinates) of the cells
    where the two λ fields differ, widened by one cell on each side: only
    faces with a cell in it change their conductance.  None when equal."""
    diff = np.asarray(lam_p) != np.asarray(lam_0)
    if not diff.any():
        return None
    box = []
    for ax in range(3):
        idx = np.nonzero(diff.any(axis=tuple(a for a in range(3) if a != ax)))[0]
        box.append((max(int(idx[0]) - 1, 0), min(int(idx[-1]) + 2, diff.shape[ax])))
    return box


def conductance_change(m, lam_p, lam_0, box=True):
    """Per face direction, G·(λf(λp) - λf(λ0)) over the interior cells:
    the change of the flux-form face conductances (G the face's area over
    distance, λf the harmonic mean with factor 2 at the halo).  Each entry
    is (nb, sel, dG): nb the neighbour's slice of the full array, sel the
    interior indices (0-based within the interior) of the nonzero changes
    and dG their values.  With box, the changes are formed only over the
    cells of changed_box of the halo λ fields (the same values: every face
    outside it is unchanged)."""
    NX, NY, NZ = m.NX, m.NY, m.NZ
    n = (NX, NY, NZ)
    I = (slice(1, NX + 1), slice(1, NY + 1), slice(1, NZ + 1))
    lp, l0 = halo_lambda(m, lam_p), halo_lambda(m, lam_0)
    if box:
        bx = changed_box(lp, l0)
        if bx is None:
            return []
        # the interior cells of the box, in halo coordinates
        s = [max(1, bx[a][0]) for a in range(3)]
        e = [min(n[a] + 1, bx[a][1]) for a in range(3)]
        if any(e[a] <= s[a] for a in range(3)):
            return []
    else:
        s, e = [1, 1, 1], [NX + 1, NY + 1, NZ + 1]
    C = tuple(slice(s[a], e[a]) for a in range(3))
    shp = tuple(e[a] - s[a] for a in range(3))
    dz = m.dZ[None, None, s[2]:e[2]]
    zc = m.ZC
    out = []
    for ax in range(3):
        for d in (-1, 1):
            nb = list(I)
            nb[ax] = slice(I[ax].start + d, I[ax].stop + d)
            nb = tuple(nb)
            cn = list(C)
            cn[ax] = slice(C[ax].start + d, C[ax].stop + d)
            cn = tuple(cn)
            inner = np.ones(shp, dtype=bool)
            # the face on the domain's edge (the halo neighbour), if in C
            edge_at = 1 if d < 0 else n[ax]
            if s[ax] <= edge_at < e[ax]:
                edge = [slice(None)] * 3
                edge[ax] = edge_at - s[ax]
                inner[tuple(edge)] = False
            fac = np.where(inner, 1.0, 2.0)
            if ax == 0:
                G = m.dy * dz / m.dx
            elif ax == 1:
                G = m.dx * dz / m.dy
            else:
                k = np.arange(s[2], e[2])
                G = (m.dx * m.dy / np.abs(zc[k + d] - zc[k]))[None, None, :]

            def lamf(a, b):
                with np.errstate(invalid="ignore", divide="ignore"):
                    v = 2.0 * a * b / (a + b) * fac
                return np.where(a + b > 0, v, 0.0)

            dG = G * (lamf(lp[cn], lp[C]) - lamf(l0[cn], l0[C]))
            sel = np.nonzero(dG)
            vals = dG[sel]
            sel = tuple(sel[a] + (s[a] - 1) for a in range(3))
            out.append((nb, sel, vals))
    return out


def apply_conductance_change(m, change, theta):
    """(A(λp) - A(λ0))·θ over the interior for a conductance_change."""
    NX, NY, NZ = m.NX, m.NY, m.NZ
    I = (slice(1, NX + 1), slice(1, NY + 1), slice(1, NZ + 1))
    out = np.zeros(theta.shape, order="F")
    oi = out[I]
    ti = theta[I]
    for nb, sel, dG in change:
        oi[sel] += dG * (theta[nb][sel] - ti[sel])
    return out


def operator_change(m, lam_p, lam_0, theta):
    """(A(λp) - A(λ0))·θ over the interior in the flux form: per face,
    G·(λf(λp) - λf(λ0))·(θ_n - θ_c).  Convective and transient terms do
    not depend on λ."""
    return apply_conductance_change(m, conductance_change(m, lam_p, lam_0), theta)


def delta_J_split(theta0, d_p, d_m, region, beta=BETA):
    """J(θ0 + d_p) - J(θ0 + d_m) from the changes themselves: the
    exponents' differences are β(d_p - d_m), never a difference of two
    temperatures near 300 K."""
    sel = np.asarray(region, dtype=bool)
    t0, dp, dm = theta0[sel], d_p[sel], d_m[sel]
    tm = t0 + dm
    wm = np.exp(beta * (tm - tm.max()))
    return np.log1p(np.sum(wm * np.expm1(beta * (dp - dm))) / np.sum(wm)) / beta


class SteadySensitivity:
    """Forward solve, objective and adjoint gradients on one device context."""

    def __init__(self, m, beta=BETA, region=None, solver=L.SOLVER_PBICGSTAB, smoother=L.SMOOTHER_GS, tol=1e-12,
                 max_itr=20000, dt=10000.0):
        self.m = m
        self.beta = beta
        self.region = silicon_region(m.ID) if region is None else np.asfortranarray(region, dtype=np.uint8)
        self.groups = layer_groups(m)
        self.solver, self.smoother, self.tol, self.max_itr, self.dt = solver, smoother, tol, max_itr, dt
        self.wk = WorkBuffers(*m.ID.shape)
        load_model(self.wk, m)
        self.theta_init = self.wk.download("theta")
        self.scales = np.ones(NGROUPS)
        self.hsrc = np.array(m.hsrc, order="F", copy=True)
        self.volume = np.zeros(m.ID.shape, order="F")
        self.volume[1:-1, 1:-1, 1:-1] = m.dx * m.dy * m.dZ[None, None, 1:-1]
        self.last = None
        self._solved = False    # the device holds the forward solution of the current parameters
        self._theta0 = None

    def close(self):
        self.wk.close()

    def set_params(self, scales=None, hsrc=None):
        self._solved = False
        if scales is not None:
            self.scales = np.asarray(scales, dtype=np.float64).copy()
        if hsrc is not None:
            self.hsrc = np.array(hsrc, dtype=np.float64, order="F", copy=True)
        self.wk.upload("lam", scaled_lambda(self.m, self.groups, self.scales))
        self.wk.upload("hsrc", self.hsrc)

    def forward(self, scales=None, hsrc=None, theta_start=None):
        """Solve A θ = b for the parameters and return J.  The right-hand
        side is formed at the initial θ (= θ∞, as q3d does); the solve then
        starts from theta_start when given (same system, warm start)."""
        self.set_params(scales, hsrc)
        self.wk.upload("theta", self.theta_init)
        self.wk.apply_bc()
        self.wk.fill("b", 0.0)
        self.wk.calc_rhs(self.dt, True)
        if theta_start is not None:
            self.wk.upload("theta", theta_start)
        conv, itr, res0 = self.wk.solve(self.solver, self.dt, self.tol, self.max_itr, self.smoother, True)
        self.last = {"converged": conv, "itr": itr, "res0": res0, "solve_ms": self.wk.last_solve_ms()}
        self._solved = True
        self._theta0 = None       # host copy of the solution, formed on first use
        return self.wk.objective(self.region, self.beta)

    def theta(self):
        return self.wk.download("theta")

    def adjoint(self):
        """Gradients of the last forward solve's J."""
        if not self._solved:
            raise RuntimeError("adjoint() needs forward() after the last parameter change or difference solve")
        conv, itr, res0 = self.wk.adjoint_solve(self.solver, self.dt, self.tol, self.max_itr, self.smoother, True)
        adj_ms = self.wk.last_solve_ms()
        glam, ghs = self.wk.gradients()
        base = self.m.lam
        dscale = np.array([np.sum(base[self.groups == k] * glam[self.groups == k]) for k in range(NGROUPS)])
        return {"dJ_dscale": dscale, "dJ_dlambda": glam, "dJ_dhsrc": ghs,
                "adjoint": {"converged": conv, "itr": itr, "res0": res0, "solve_ms": adj_ms}}

    def halo_lambda(self, lam):
        return halo_lambda(self.m, lam)

    def operator_change(self, lam_p, lam_0, theta):
        return operator_change(self.m, lam_p, lam_0, theta)

    def _solve_change(self, rhs, lam, solver=None):
        """δ of A(λ) δ = rhs with a zero halo (the solve's tolerance is
        relative to this right-hand side).  The system is homogeneous in δ,
        so it is solved for a power-of-two multiple of rhs of order one:
        PBiCGSTAB!'s breakdown test compares (r, r0) with the absolute
        FloatMin = 1e-37 (src/common.jl:12), which a converging solve of a
        tiny right-hand side would reach."""
        wk = self.wk
        self._solved = False
        scale = 2.0 ** (-np.round(np.log2(np.max(np.abs(rhs)))))
        wk.upload("lam", lam)
        wk.upload("theta", self.theta_init)
        wk.apply_bc()
        wk.fill("theta", 0.0)
        wk.upload("b", rhs * scale)
        conv, itr, res0 = wk.solve(self.solver if solver is None else solver, self.dt, self.tol, self.max_itr,
                                   self.smoother, True)
        if not conv:
            raise RuntimeError("perturbation solve did not converge (itr %d)" % itr)
        return wk.download("theta") / scale

    def central_difference(self, kind, index, rel_step=1e-6, theta0=None):
        """Central difference of J in one parameter: kind 'scale' (group
        index) or 'hsrc' (cell index tuple), relative step rel_step.
        Each perturbed problem A± θ± = b± is solved exactly as
        θ± = θ0 + δ±, A± δ± = (b± - b) - (A± - A)θ0 from δ = 0, θ0 the
        unperturbed solution: the iterative solve's tolerance then applies
        to the change instead of the temperature, and the right-hand side
        holds conductance differences times temperature differences."""
        if theta0 is None:
            if self._theta0 is None:
                if not self._solved:
                    raise RuntimeError("central_difference() needs forward() first (or theta0)")
                self._theta0 = self.theta()
            theta0 = self._theta0
        lam0 = scaled_lambda(self.m, self.groups, self.scales)
        thetas = []
        for sign in (1.0, -1.0):
            if kind == "scale":
                sc = self.scales.copy()
                h = rel_step * sc[index]
                sc[index] += sign * h
                lam = scaled_lambda(self.m, self.groups, sc)
                rhs = -self.operator_change(lam, lam0, theta0)
                solver = None
            else:
                h = rel_step * self.hsrc[index]
                lam = lam0
                rhs = np.zeros(theta0.shape, order="F")
                rhs[index] = -self.volume[index] * (sign * h)  # b = -(hsrc·V + ...)
                # A one-cell right-hand side breaks PBiCGSTAB! down when the
                # cell is updated by the smoother's last colour: after it the
                # preconditioned residual vanishes there, and so does
                # (r, r0) with r0 the same one-cell vector.  CG! has no
                # shadow vector; (p, A p) cannot vanish for a definite A.
                solver = L.SOLVER_CG
            thetas.append(theta0 + self._solve_change(rhs, lam, solver))
        self.set_params(self.scales, self.hsrc)
        return delta_J(thetas[0], thetas[1], self.region, self.beta) / (2.0 * h)


try:
    import torch

    class SmoothMaxTemperature(torch.autograd.Function):
        """J(scales, hsrc) on the device with its adjoint gradient as the
        backward pass: the torch analogue of a ChainRulesCore rrule.
        Call as SmoothMaxTemperature.apply(scales, hsrc, sens) with sens a
        SteadySensitivity; scales has 13 entries, hsrc the model's shape."""

        @staticmethod
        def forward(ctx, scales, hsrc, sens):
            J = sens.forward(scales.detach().cpu().double().numpy(),
                             np.asfortranarray(hsrc.detach().cpu().double().numpy()))
            ctx.sens = sens
            return torch.tensor(J, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gJ):
            g = ctx.sens.adjoint()
            return (gJ * torch.from_numpy(g["dJ_dscale"]),
                    gJ * torch.from_numpy(np.ascontiguousarray(g["dJ_dhsrc"])), None)
except ImportError:  # torch is optional for the library itself
    pass
