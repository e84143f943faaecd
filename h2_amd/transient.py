"""Implicit-Euler runs of the solve path and the adjoint of a time-integrated
smooth maximum temperature, with θ checkpoints and recompute.

The forward step follows the reference's time loop (src/heat3ds.jl:123-139):
t_n = n·Δt, calRHS! at θⁿ⁻¹, then PBiCGSTAB!/CG! with is_steady = false.
Two intended differences:
  * b is cleared before every calRHS! (the reference never clears it and
    only ever runs nt = 1: calRHS! accumulates into b, src/RHS.jl, and
    src/NonUniform.jl:85-90 folds the old b into the new one);
  * the heat source varies in time: hsrc_n = f_n·hsrc_base with
    f_n = 1 + amp·sin(2π t_n / period).

Objective (no reference counterpart): J_total = Σ_{n=1..N} Δt·J(θⁿ), J the
smooth maximum of θ over the silicon cells (sensitivity.py).
Adjoint: θⁿ solves A θⁿ = D θⁿ⁻¹ + c_n with D = ∂b/∂θ_old diagonal, so
    Aᵀ μ_n = Δt·∂J/∂θ(θⁿ) + D μ_{n+1},  μ_{N+1} = 0,
    dJ_total/dλ = Σ_n (per-cell gradient of θⁿ, μ_n),
    dJ_total/dhsrc_base = Σ_n f_n·(-μ_n V).
The backward pass needs θⁿ in reverse order.  The forward keeps θ every
`checkpoint_every` steps in device slots; the backward recomputes each
segment from its checkpoint into segment slots and walks it backwards.
Solves are deterministic (fixed reduction order), so a recomputed θⁿ is
bit for bit the forward's.
"""
import argparse
import json
import math
import sys
import time

import numpy as np

from . import _lib as L
from .heat3ds import WorkBuffers, load_model
from .sensitivity import (BETA, NGROUPS, apply_conductance_change, conductance_change, delta_J_split,
                          layer_groups, scaled_lambda, silicon_region)

SLOT_HSRC = 0      # hsrc_base
SLOT_CKPT = 1      # θ⁰, θᵏ, θ²ᵏ, ..., θᴺ from here on


class TransientRun:
    """N implicit-Euler steps of Δt on one device context, the objective
    J_total and its adjoint gradients."""

    def __init__(self, m, dt=1.0e-3, nsteps=100, checkpoint_every=10, amp=0.5, period=20.0e-3, beta=BETA,
                 region=None, solver=L.SOLVER_PBICGSTAB, smoother=L.SMOOTHER_GS, tol=1e-8, precision=L.FP64,
                 adjoint_precision=None, max_itr=20000, device=0, verbose=False, true_residual=False):
        if nsteps < 1 or checkpoint_every < 1:
            raise ValueError("nsteps and checkpoint_every must be >= 1")
        self.m = m
        self.dt, self.nsteps, self.every = float(dt), int(nsteps), int(checkpoint_every)
        self.amp, self.period, self.beta = float(amp), float(period), float(beta)
        self.region = silicon_region(m.ID) if region is None else np.asfortranarray(region, dtype=np.uint8)
        self.groups = layer_groups(m)
        self.solver, self.smoother, self.tol, self.max_itr = solver, smoother, float(tol), int(max_itr)
        self.precision = precision
        self.adjoint_precision = precision if adjoint_precision is None else adjoint_precision
        self.wk = WorkBuffers(*m.ID.shape, device=device)
        load_model(self.wk, m)
        self.theta_init = self.wk.download("theta")
        self.scales = np.ones(NGROUPS)
        self.hsrc = np.array(m.hsrc, order="F", copy=True)
        self.nseg = (self.nsteps + self.every - 1) // self.every
        self.slot_seg = SLOT_CKPT + self.nseg + 1
        self.stats = {}
        self.verbose = verbose
        # after each forward step: the true FP64 residual ||b - Aθ|| relative
        # to the step's res0, and in units of its rounding scale 2^-53·||s||
        self.true_residual = bool(true_residual)
        self._setup_count = 0     # bumped whenever the device fields are set up again
        self._ckpt_setup = -1     # the set-up whose forward run left checkpoints

    def _log(self, *a):
        if self.verbose:
            print(*a, flush=True)

    def close(self):
        self.wk.close()

    def factor(self, n):
        """f_n = 1 + amp·sin(2π t_n / period), t_n = n·Δt."""
        return 1.0 + self.amp * math.sin(2.0 * math.pi * (n * self.dt) / self.period)

    def _setup(self, scales, hsrc):
        if scales is not None:
            self.scales = np.asarray(scales, dtype=np.float64).copy()
        if hsrc is not None:
            self.hsrc = np.array(hsrc, dtype=np.float64, order="F", copy=True)
        wk = self.wk
        self._setup_count += 1
        wk.upload("lam", scaled_lambda(self.m, self.groups, self.scales))
        wk.upload("theta", self.theta_init)
        wk.apply_bc()
        wk.upload("hsrc", self.hsrc)
        wk.store("hsrc", SLOT_HSRC)

    def _step(self, n, record=False):
        """θⁿ⁻¹ (in θ) → θⁿ; returns J(θⁿ) and the solve's iteration count."""
        wk = self.wk
        wk.load("hsrc", SLOT_HSRC, self.factor(n))
        wk.fill("b", 0.0)
        wk.calc_rhs(self.dt, False)
        conv, itr, res0 = wk.solve(self.solver, self.dt, self.tol, self.max_itr, self.smoother, False,
                                   precision=self.precision)
        if not conv:
            raise RuntimeError("step %d did not converge (%d iterations)" % (n, itr))
        if record:
            rt = wk.calc_rk("pcg_r", "theta", "b", self.dt, False)
            ulp = math.ldexp(wk.residual_scale("theta", "b", self.dt, False), -53)
            self.stats["true_residual"].append(
                [n, rt / res0 if res0 > 0 else 0.0, rt / ulp if ulp > 0 else 0.0, bool(wk.last_solve_floor())])
        return wk.objective(self.region, self.beta), itr

    def forward(self, scales=None, hsrc=None, checkpoints=True):
        """Runs the N steps from the initial field; returns J_total."""
        self._setup(scales, hsrc)
        wk = self.wk
        if checkpoints:
            wk.store("theta", SLOT_CKPT)
        J, iters = 0.0, []
        self.stats["true_residual"] = []
        t0 = time.perf_counter()
        for n in range(1, self.nsteps + 1):
            Jn, itr = self._step(n, record=self.true_residual)
            J += self.dt * Jn
            iters.append(itr)
            if self.true_residual:
                _, rel, ulps, at_floor = self.stats["true_residual"][-1]
                self._log("forward step %d J %.9f itr %d true res %.3e (%.2f u|s|%s)" %
                          (n, Jn, itr, rel, ulps, ", at floor" if at_floor else ""))
            else:
                self._log("forward step %d J %.9f itr %d" % (n, Jn, itr))
            if checkpoints and (n % self.every == 0 or n == self.nsteps):
                wk.store("theta", SLOT_CKPT + (n + self.every - 1) // self.every)
        self.stats["forward_s"] = time.perf_counter() - t0
        self.stats["forward_iters"] = iters
        self.stats["mem_forward"] = wk.mem_info()[0]
        self.J = J
        self._ckpt_setup = self._setup_count if checkpoints else -1
        return J

    def adjoint(self):
        """dJ_total/dscale (13 groups), dJ_total/dλ and dJ_total/dhsrc_base
        per cell, for the last forward run (which must have kept its
        checkpoints)."""
        if self._ckpt_setup != self._setup_count:
            raise RuntimeError("adjoint() needs forward(checkpoints=True) after the last parameter change, "
                               "states() or central_difference()")
        wk = self.wk
        wk.reset_gradients()
        chain = False
        recompute_iters, adjoint_iters = [], []
        t0 = time.perf_counter()
        for s in reversed(range(self.nseg)):
            n0, n1 = s * self.every, min(self.nsteps, (s + 1) * self.every)
            # θ^{n0+1} .. θ^{n1-1} from the checkpoint θ^{n0}; θ^{n1} is a checkpoint
            wk.load("theta", SLOT_CKPT + s)
            for n in range(n0 + 1, n1):
                _, itr = self._step(n)
                recompute_iters.append(itr)
                wk.store("theta", self.slot_seg + (n - n0 - 1))
            for n in range(n1, n0, -1):
                wk.load("theta", SLOT_CKPT + s + 1 if n == n1 else self.slot_seg + (n - n0 - 1))
                wk.objective(self.region, self.beta)
                conv, itr, _ = wk.adjoint_step(self.solver, self.dt, self.tol, self.dt, chain, self.max_itr,
                                               self.smoother, self.adjoint_precision)
                if not conv:
                    raise RuntimeError("adjoint step %d did not converge (%d iterations)" % (n, itr))
                adjoint_iters.append(itr)
                wk.accumulate_gradients(1.0, self.factor(n))
                self._log("adjoint step %d itr %d" % (n, itr))
                chain = True
        glam, ghs = wk.accumulated_gradients()
        self.stats["adjoint_s"] = time.perf_counter() - t0
        self.stats["recompute_iters"] = recompute_iters
        self.stats["adjoint_iters"] = adjoint_iters
        self.stats["mem_peak"] = wk.mem_info()[0]
        base = self.m.lam
        dscale = np.array([np.sum(base[self.groups == k] * glam[self.groups == k]) for k in range(NGROUPS)])
        return {"dJ_dscale": dscale, "dJ_dlambda": glam, "dJ_dhsrc": ghs}

    def states(self, scales=None, hsrc=None):
        """θ⁰..θᴺ of a forward run, on the host."""
        self._setup(scales, hsrc)
        out = [self.wk.download("theta")]
        for n in range(1, self.nsteps + 1):
            self._step(n)
            out.append(self.wk.download("theta"))
        return out

    def _rhs_change(self, d):
        """D·d for the diagonal D = ∂b/∂θ_old, from the device's calRHS!
        itself: b(θ = s·d) - b(θ = 0) with hsrc = 0, s a power of two that
        makes s·d of order 2²⁰, so that θ∞ - s·d keeps d's digits."""
        amax = np.max(np.abs(d))
        if amax == 0.0:
            return np.zeros(d.shape, order="F")
        s = 2.0 ** (20 - np.round(np.log2(amax)))
        wk = self.wk
        wk.fill("hsrc", 0.0)
        out = []
        for x in (s * d, None):
            if x is None:
                wk.fill("theta", 0.0)
            else:
                wk.upload("theta", x)
            wk.fill("b", 0.0)
            wk.calc_rhs(self.dt, False)
            out.append(wk.download("b"))
        return (out[0] - out[1]) / s

    def _solve_change(self, rhs, lam, precision=L.FP64):
        """δ of A(λ) δ = rhs (transient operator, zero halo), solved for a
        power-of-two multiple of rhs of order one, in the given precision
        (FP64, or FP32 inner solves refined to the same FP64 tolerance)."""
        wk = self.wk
        self._setup_count += 1
        scale = 2.0 ** (-np.round(np.log2(np.max(np.abs(rhs)))))
        wk.upload("lam", lam)
        wk.upload("theta", self.theta_init)
        wk.apply_bc()
        wk.fill("theta", 0.0)
        wk.upload("b", rhs * scale)
        conv, itr, _ = wk.solve(self.solver, self.dt, self.tol, self.max_itr, self.smoother, False,
                                precision=precision)
        if not conv:
            raise RuntimeError("perturbation solve did not converge (itr %d)" % itr)
        return wk.download("theta") / scale

    def central_difference(self, k, rel_step=1e-6, thetas=None, precision=L.FP64):
        """Central difference of J_total in the scale of group k at relative
        step rel_step.  Each perturbed run θ±ⁿ = θⁿ + δ±ⁿ is computed
        through its change:
            A± δⁿ = D δⁿ⁻¹ - (A± - A) θⁿ,  δ⁰ = 0
        (b is affine in θⁿ⁻¹ with slope D and does not depend on λ), with
        (A± - A) θⁿ in the flux form on the host, D from calRHS! on the
        device and FP64 flux-form solves; the J differences come from the
        changes (delta_J_split).  The iterative solves' tolerance then
        applies to δ instead of θ.  thetas: θ⁰..θᴺ of the unperturbed run
        (recomputed when not given); precision: that of the change solves."""
        if thetas is None:
            thetas = self.states()
        m, sc0 = self.m, self.scales.copy()
        lam0 = scaled_lambda(m, self.groups, sc0)
        h = rel_step * sc0[k]
        self.wk.set_operator_form(L.FORM_FLUX)
        try:
            deltas = []
            for sign in (1.0, -1.0):
                sc = sc0.copy()
                sc[k] += sign * h
                lam = scaled_lambda(m, self.groups, sc)
                change = conductance_change(m, lam, lam0)
                d, ds = np.zeros(thetas[0].shape, order="F"), [None]
                for n in range(1, self.nsteps + 1):
                    rhs = self._rhs_change(d) - apply_conductance_change(m, change, thetas[n])
                    d = self._solve_change(rhs, lam, precision)
                    ds.append(d)
                    self._log("difference group %d sign %+d step %d" % (k, int(sign), n))
                deltas.append(ds)
        finally:
            self.wk.set_operator_form(L.FORM_REFERENCE)
        self.scales = sc0
        dJ = sum(self.dt * delta_J_split(thetas[n], deltas[0][n], deltas[1][n], self.region, self.beta)
                 for n in range(1, self.nsteps + 1))
        return dJ / (2.0 * h)


def main(argv=None):
    """Forward and adjoint of an implicit-Euler run with timings, device
    memory and optional central differences, as one JSON line."""
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--nx", type=int, default=960)
    ap.add_argument("--ny", type=int, default=960)
    ap.add_argument("--nz", type=int, default=120)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--every", type=int, default=10)
    ap.add_argument("--dt", type=float, default=1.0e-3)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--precision", choices=("fp64", "fp32"), default="fp32")
    ap.add_argument("--solver", choices=("pbicgstab", "cg"), default="pbicgstab")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--true-residual", action="store_true",
                    help="log each step's true FP64 residual and its rounding scale")
    ap.add_argument("--fd-groups", type=str, default="", help="comma-separated groups for central differences")
    ap.add_argument("--fd-step", type=float, default=1e-6)
    ap.add_argument("--out", type=str, default="")
    a = ap.parse_args(argv)
    from .heat3ds import build_model
    t0 = time.perf_counter()
    m = build_model(a.nx, a.ny, a.nz, power="map")
    print("model built in %.1f s" % (time.perf_counter() - t0), flush=True)
    prec = L.FP32_REFINED if a.precision == "fp32" else L.FP64
    solver = L.SOLVER_CG if a.solver == "cg" else L.SOLVER_PBICGSTAB
    run = TransientRun(m, dt=a.dt, nsteps=a.steps, checkpoint_every=a.every, tol=a.tol, precision=prec,
                       solver=solver, verbose=True, true_residual=a.true_residual)
    used0, total = run.wk.mem_info()
    J = run.forward()
    out = {"grid": [a.nx, a.ny, a.nz], "steps": a.steps, "dt": a.dt, "checkpoint_every": a.every, "tol": a.tol,
           "precision": a.precision, "solver": a.solver, "J_total": J, "forward_s": run.stats["forward_s"],
           "forward_iters": run.stats["forward_iters"], "device_mem_before_run": used0,
           "device_mem_after_forward": run.stats["mem_forward"], "device_total": total}
    if a.true_residual:
        out["true_residual"] = run.stats["true_residual"]
    if not a.no_adjoint:
        g = run.adjoint()
        out.update({"adjoint_s": run.stats["adjoint_s"], "adjoint_iters": run.stats["adjoint_iters"],
                    "recompute_iters": run.stats["recompute_iters"], "device_mem_peak": run.stats["mem_peak"],
                    "dJ_dscale": g["dJ_dscale"].tolist()})
        fd, thetas = {}, None
        for k in [int(v) for v in a.fd_groups.split(",") if v.strip()]:
            if thetas is None:
                thetas = run.states()
            fd[k] = run.central_difference(k, a.fd_step, thetas)
            print("group %d adjoint %.12e fd %.12e" % (k, g["dJ_dscale"][k], fd[k]), flush=True)
        if fd:
            out["fd"] = {str(k): {"adjoint": g["dJ_dscale"][k], "fd": v,
                                  "rel": abs(g["dJ_dscale"][k] - v) / abs(v)} for k, v in fd.items()}
    run.close()
    line = json.dumps(out)
    print(line, flush=True)
    if a.out:
        with open(a.out, "w") as f:
            f.write(line + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
