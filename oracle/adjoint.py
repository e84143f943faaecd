"""CPU reference of the sensitivity extension (smooth-max objective and its
adjoint gradients).  Test infrastructure only: nothing in the product
imports it.

Nothing here restates the device's derivative formulas.  The operator
comes from the oracle's own CalcAX! (orc_calc_ax, src/NonUniform.jl:454-511):
  * the interior matrix A is recovered exactly by probing, 27 applications
    to vectors with ones on a period-3 lattice (a 7-point stencil's
    neighbours fall in distinct residue classes);
  * ∂(Aθ)/∂s is a central difference of the operator application itself;
  * the forward and adjoint solves are direct (scipy spsolve);
  * full finite differences re-solve the perturbed problem.
The objective and its seed are the closed forms of the log-sum-exp.
"""
import numpy as np
from scipy.sparse import coo_matrix
from scipy.sparse.linalg import spsolve

from . import oracle as O

SILICON_IDS = (2, 7)


def probe_matrix(c):
    """Sparse interior matrix of the oracle's CalcAX! for the case c."""
    MX, MY, MZ = c.shape
    NX, NY, NZ = MX - 2, MY - 2, MZ - 2
    n = NX * NY * NZ
    idx = np.arange(n).reshape((NX, NY, NZ), order="F")
    ii, jj, kk = np.meshgrid(np.arange(NX), np.arange(NY), np.arange(NZ), indexing="ij")
    rows, cols, vals = [], [], []
    x = np.zeros(c.shape, order="F")
    ax = np.zeros(c.shape, order="F")
    for a in range(3):
        for b in range(3):
            for d in range(3):
                sel = (ii % 3 == a) & (jj % 3 == b) & (kk % 3 == d)
                x[:] = 0.0
                x[1:-1, 1:-1, 1:-1][sel] = 1.0
                c.calc_ax(ax, x)
                y = ax[1:-1, 1:-1, 1:-1]
                # every row's entry in this probe belongs to the unique probed
                # cell among itself and its six neighbours
                for di, dj, dk in ((0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
                    ci, cj, ck = ii + di, jj + dj, kk + dk
                    ok = (ci >= 0) & (ci < NX) & (cj >= 0) & (cj < NY) & (ck >= 0) & (ck < NZ)
                    ok &= (ci % 3 == a) & (cj % 3 == b) & (ck % 3 == d)
                    rows.append(idx[ok])
                    cols.append(idx[ci[ok], cj[ok], ck[ok]])
                    vals.append(y[ok])
    A = coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n)).tocsr()
    A.eliminate_zeros()
    return A


def smooth_max(theta, region, beta):
    """J = m + log Σ exp(β(θ - m))/β over the region and its gradient."""
    sel = np.asarray(region, dtype=bool)
    t = theta[sel]
    m = t.max()
    w = np.exp(beta * (t - m))
    S = w.sum()
    g = np.zeros(theta.shape, order="F")
    g[sel] = w / S
    return m + np.log(S) / beta, g


def silicon_region(ID):
    r = np.isin(ID, SILICON_IDS).astype(np.uint8)
    r[0], r[-1] = 0, 0
    r[:, 0], r[:, -1] = 0, 0
    r[:, :, 0], r[:, :, -1] = 0, 0
    return r


class SteadyProblem:
    """A modelA case of the oracle with λ = s_g λ⁰ per group and heat source hsrc."""

    def __init__(self, NX, NY, NZ, groups, hsrc=None):
        self.base = O.modelA_case(NX, NY, NZ)
        self.lam0 = self.base.arr["lam"].copy(order="F")
        self.groups = groups
        self.hsrc = (self.base.arr["hsrc"] if hsrc is None else hsrc).copy(order="F")
        self.region = silicon_region(self.base.ID)

    def case(self, scales, hsrc=None):
        """Fresh oracle case with the parameters applied (BCs re-applied)."""
        c = O.modelA_case(*[s - 2 for s in self.base.shape])
        lam = self.lam0.copy(order="F")
        for k, s in enumerate(scales):
            sel = self.groups == k
            lam[sel] = self.lam0[sel] * s
        c.arr["lam"][:] = lam
        c.arr["theta"][:] = 300.0
        c.apply_bc()
        c.arr["hsrc"][:] = self.hsrc if hsrc is None else hsrc
        c.arr["b"][:] = 0.0
        c.cal_rhs()
        return c

    def solve(self, c):
        """θ of the affine system b - CalcAX!(θ) = 0 (direct solve)."""
        A = probe_matrix(c)
        c.calc_rk("r", "theta", "b")
        r = c.arr["r"]
        th = c.arr["theta"].copy(order="F")
        d = spsolve(A, r[1:-1, 1:-1, 1:-1].ravel(order="F"))
        th[1:-1, 1:-1, 1:-1] += d.reshape(th[1:-1, 1:-1, 1:-1].shape, order="F")
        return th, A

    def objective(self, scales, beta, hsrc=None):
        c = self.case(scales, hsrc)
        th, A = self.solve(c)
        J, g = smooth_max(th, self.region, beta)
        return J, th, g, A, c

    def operator_derivative(self, th, scales, k, h):
        """∂(CalcAX!(θ))/∂s_k by central differences of the operator
        application at relative steps h and 2h, Richardson-extrapolated."""
        ax = np.zeros(th.shape, order="F")

        def cd(step):
            d = []
            for sign in (1.0, -1.0):
                sc = np.array(scales, dtype=float)
                sc[k] += sign * step * sc[k]
                cp = self.case(sc)
                tt = th.copy(order="F")
                # the halo of θ as the perturbed case sets it
                tt[0], tt[-1] = cp.arr["theta"][0], cp.arr["theta"][-1]
                tt[:, 0], tt[:, -1] = cp.arr["theta"][:, 0], cp.arr["theta"][:, -1]
                tt[:, :, 0], tt[:, :, -1] = cp.arr["theta"][:, :, 0], cp.arr["theta"][:, :, -1]
                cp.calc_ax(ax, tt)
                d.append(ax.copy(order="F"))
            return (d[0] - d[1]) / (2.0 * step * scales[k])

        return (4.0 * cd(h) - cd(2.0 * h)) / 3.0

    def adjoint(self, scales, beta, h=1e-3):
        """J, θ, μ and the gradients: dJ/ds_g = -μᵀ ∂(Aθ)/∂s_g with the
        operator derivative by extrapolated differences (step h: the
        rounding of operators applied to θ ~ 300 K limits smaller steps),
        dJ/dhsrc = -μ V."""
        J, th, g, A, c = self.objective(scales, beta)
        mu_int = spsolve(A.T.tocsr(), g[1:-1, 1:-1, 1:-1].ravel(order="F"))
        mu = np.zeros(th.shape, order="F")
        mu[1:-1, 1:-1, 1:-1] = mu_int.reshape(mu[1:-1, 1:-1, 1:-1].shape, order="F")
        dscale = np.zeros(len(scales))
        for k in range(len(scales)):
            dscale[k] = -np.sum(mu * self.operator_derivative(th, scales, k, h))
        V = np.zeros(th.shape, order="F")
        V[1:-1, 1:-1, 1:-1] = c.dx * c.dy * self.base.dZ[None, None, 1:-1]
        return {"J": J, "theta": th, "mu": mu, "seed": g, "dJ_dscale": dscale, "dJ_dhsrc": -mu * V}


def probe_rhs_derivative(c):
    """∂b/∂θ_old of the oracle's calRHS! (orc_cal_rhs, src/NonUniform.jl:85-90
    and src/RHS.jl) as a sparse interior matrix.  b is affine in the θ it
    is formed at, so 27 lattice probes of b(θ = 2²⁰·e) - b(θ = 0) with hsrc
    = 0 recover it (the large probe keeps θ∞ - θ free of rounding)."""
    MX, MY, MZ = c.shape
    NX, NY, NZ = MX - 2, MY - 2, MZ - 2
    n = NX * NY * NZ
    idx = np.arange(n).reshape((NX, NY, NZ), order="F")
    ii, jj, kk = np.meshgrid(np.arange(NX), np.arange(NY), np.arange(NZ), indexing="ij")
    saved = {f: c.arr[f].copy(order="F") for f in ("theta", "b", "hsrc")}
    big = 2.0 ** 20

    def rhs_at(x):
        c.arr["theta"][:] = 0.0
        c.arr["theta"][1:-1, 1:-1, 1:-1] = x
        c.arr["hsrc"][:] = 0.0
        c.arr["b"][:] = 0.0
        c.cal_rhs()
        return c.arr["b"][1:-1, 1:-1, 1:-1].copy()

    b0 = rhs_at(np.zeros((NX, NY, NZ)))
    rows, cols, vals = [], [], []
    for a in range(3):
        for b in range(3):
            for d in range(3):
                sel = (ii % 3 == a) & (jj % 3 == b) & (kk % 3 == d)
                y = (rhs_at(np.where(sel, big, 0.0)) - b0) / big
                for di, dj, dk in ((0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
                    ci, cj, ck = ii + di, jj + dj, kk + dk
                    ok = (ci >= 0) & (ci < NX) & (cj >= 0) & (cj < NY) & (ck >= 0) & (ck < NZ)
                    ok &= (ci % 3 == a) & (cj % 3 == b) & (ck % 3 == d)
                    rows.append(idx[ok])
                    cols.append(idx[ci[ok], cj[ok], ck[ok]])
                    vals.append(y[ok])
    for f, v in saved.items():
        c.arr[f][:] = v
    D = coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n)).tocsr()
    D.eliminate_zeros()
    return D


class TransientProblem(SteadyProblem):
    """N implicit-Euler steps of Δt from 300 K on a modelA case with
    hsrc_n = f_n·hsrc, f_n = 1 + amp·sin(2π nΔt/period), b cleared before
    every calRHS!, J_total = Σ_n Δt·J(θⁿ).  Direct solves; the adjoint
    recursion Aᵀμ_n = Δt g_n + Dᵀ μ_{n+1} uses the probed A and D."""

    def __init__(self, NX, NY, NZ, groups, dt, nsteps, amp=0.5, period=20e-3, hsrc=None):
        super().__init__(NX, NY, NZ, groups, hsrc)
        self.dt, self.nsteps, self.amp, self.period = dt, nsteps, amp, period

    def factor(self, n):
        return 1.0 + self.amp * np.sin(2.0 * np.pi * (n * self.dt) / self.period)

    def case(self, scales, hsrc=None):
        NX, NY, NZ = [s - 2 for s in self.base.shape]
        c = O.modelA_case(NX, NY, NZ, steady=False, dt=self.dt)
        lam = self.lam0.copy(order="F")
        for k, s in enumerate(scales):
            sel = self.groups == k
            lam[sel] = self.lam0[sel] * s
        c.arr["lam"][:] = lam
        c.arr["theta"][:] = 300.0
        c.apply_bc()
        return c

    def run(self, scales, beta, hsrc=None):
        """J_total, the states θ⁰..θᴺ, the seeds, A, D and the case."""
        base = self.hsrc if hsrc is None else hsrc
        c = self.case(scales)
        A = probe_matrix(c)
        D = probe_rhs_derivative(c)
        thetas, seeds, J = [c.arr["theta"].copy(order="F")], [None], 0.0
        for n in range(1, self.nsteps + 1):
            c.arr["hsrc"][:] = self.factor(n) * base
            c.arr["b"][:] = 0.0
            c.cal_rhs()
            c.calc_rk("r", "theta", "b")
            d = spsolve(A, c.arr["r"][1:-1, 1:-1, 1:-1].ravel(order="F"))
            c.arr["theta"][1:-1, 1:-1, 1:-1] += d.reshape(c.arr["theta"][1:-1, 1:-1, 1:-1].shape, order="F")
            Jn, g = smooth_max(c.arr["theta"], self.region, beta)
            J += self.dt * Jn
            thetas.append(c.arr["theta"].copy(order="F"))
            seeds.append(g)
        return J, thetas, seeds, A, D, c

    def adjoint(self, scales, beta, h=1e-3, groups=None):
        """J_total and its gradients by the adjoint recursion; the operator
        derivative per step as in SteadyProblem.adjoint.  groups limits
        the conductivity groups (all by default)."""
        J, thetas, seeds, A, D, c = self.run(scales, beta)
        AT, DT = A.T.tocsr(), D.T.tocsr()
        shape = thetas[0].shape
        groups = range(len(scales)) if groups is None else groups
        V = np.zeros(shape, order="F")
        V[1:-1, 1:-1, 1:-1] = c.dx * c.dy * self.base.dZ[None, None, 1:-1]
        dscale = np.zeros(len(scales))
        ghs = np.zeros(shape, order="F")
        mus = [None] * (self.nsteps + 2)
        nxt = np.zeros(AT.shape[0])
        for n in range(self.nsteps, 0, -1):
            rhs = self.dt * seeds[n][1:-1, 1:-1, 1:-1].ravel(order="F") + DT @ nxt
            nxt = spsolve(AT, rhs)
            mu = np.zeros(shape, order="F")
            mu[1:-1, 1:-1, 1:-1] = nxt.reshape(mu[1:-1, 1:-1, 1:-1].shape, order="F")
            mus[n] = mu
            for k in groups:
                dscale[k] -= np.sum(mu * self.operator_derivative(thetas[n], scales, k, h))
            ghs -= self.factor(n) * mu * V
        return {"J": J, "thetas": thetas, "mus": mus, "dJ_dscale": dscale, "dJ_dhsrc": ghs, "D": D, "A": A}

    def objective(self, scales, beta, hsrc=None):
        return self.run(scales, beta, hsrc)[0]
