"""Condition estimates of the solve's operator (host side, numpy/scipy).

The FP32 forms of the solve (H3D_FP32_REFINED) depend on how ill-conditioned
the operator is.  This module assembles the interior operator of a model
(the matrix CalcAX! applies, src/NonUniform.jl:454-511, with the boundary
conditions of BC mode 3 folded into its diagonal, src/boundary_conditions.jl,
src/common.jl:35) as the symmetric positive definite matrix -A, and
estimates its extreme eigenvalues by Lanczos through the coefficients of
conjugate gradients (the tridiagonal matrix T_k built from CG's α and β has
the Ritz values of the k-step Krylov space), for -A and for its Jacobi
scaling D^-1/2 (-A) D^-1/2.

    python -m h2_amd.spectrum --nx 240 --nz 30 --iters 3000
    python -m h2_amd.spectrum --device --nx 480 --nz 60 --iters 20000

(the second on a GPU: preconditioned CG with the library's own kernels)

prints, every few hundred iterations, the smallest and largest Ritz values
and their ratio (a lower bound of κ that rises to κ as the smallest Ritz
value converges), and κ·2^-24, the product that decides whether FP32
vectors can resolve a correction to the solution.
"""
import argparse
import sys
import time

import numpy as np


def _lamf(a, b, factor):
    """λf(a, b, ma, mb) = 2ab/(a+b)·(2 - div(ma+mb, 2)) (src/common.jl:35),
    factor = 1 for an interior neighbour, 2 for a masked (halo) one."""
    with np.errstate(invalid="ignore", divide="ignore"):
        v = 2.0 * a * b / (a + b) * factor
    return np.where(a + b > 0.0, v, 0.0)


def assemble(m, steady=True, dt=1.0):
    """-A over the interior cells (column-major, i fastest) of model m, BC
    mode 3 (convective x, y and top faces with h = 5, isothermal bottom), as
    a scipy CSR matrix, and its diagonal."""
    import scipy.sparse as sp
    NX, NY, NZ = m.NX, m.NY, m.NZ
    dx, dy = m.dx, m.dy
    lam = np.asarray(m.lam, dtype=np.float64)
    L0 = lam[1:-1, 1:-1, 1:-1]
    # halo λ after apply_boundary_conditions!: convective faces 0, the
    # isothermal bottom a copy of the interior plane above it
    lh = lam.copy(order="F")
    lh[0, :, :] = 0.0
    lh[-1, :, :] = 0.0
    lh[:, 0, :] = 0.0
    lh[:, -1, :] = 0.0
    lh[:, :, -1] = 0.0
    lh[1:-1, 1:-1, 0] = lh[1:-1, 1:-1, 1]
    HC = np.array([5.0, 5.0, 5.0, 5.0, 0.0, 5.0])
    dz = np.asarray(m.dZ)[1:-1][None, None, :]
    ZC = np.asarray(m.ZC)
    dzm = (ZC[1:-1] - ZC[:-2])[None, None, :]
    dzp = (ZC[2:] - ZC[1:-1])[None, None, :]

    def side(nb, interior_mask):
        return _lamf(nb, L0, np.where(interior_mask, 1.0, 2.0))

    ii = np.arange(1, NX + 1)[:, None, None]
    jj = np.arange(1, NY + 1)[None, :, None]
    kk = np.arange(1, NZ + 1)[None, None, :]
    cxm = side(lh[:-2, 1:-1, 1:-1], ii > 1) * dy * dz / dx
    cxp = side(lh[2:, 1:-1, 1:-1], ii < NX) * dy * dz / dx
    cym = side(lh[1:-1, :-2, 1:-1], jj > 1) * dx * dz / dy
    cyp = side(lh[1:-1, 2:, 1:-1], jj < NY) * dx * dz / dy
    czm = side(lh[1:-1, 1:-1, :-2], kk > 1) * dx * dy / dzm
    czp = side(lh[1:-1, 1:-1, 2:], kk < NZ) * dx * dy / dzp
    conv = (HC[0] * dy * dz * (ii == 1) + HC[1] * dy * dz * (ii == NX) + HC[2] * dx * dz * (jj == 1) +
            HC[3] * dx * dz * (jj == NY) + HC[4] * dx * dy * (kk == 1) + HC[5] * dx * dy * (kk == NZ))
    d = cxm + cxp + cym + cyp + czm + czp + conv
    if not steady:
        rho = np.asarray(m.rho)[1:-1, 1:-1, 1:-1]
        cp = np.asarray(m.cp)[1:-1, 1:-1, 1:-1]
        d = d + rho * cp * dx * dy * dz / dt
    n = NX * NY * NZ
    f = lambda a: np.asarray(a, dtype=np.float64).ravel(order="F")  # noqa: E731
    # off-diagonals: -c toward interior neighbours only (halo couplings are
    # in the diagonal), at offsets 1, NX, NX·NY of the column-major order
    up_x = np.where(ii < NX, cxp, 0.0)
    up_y = np.where(jj < NY, cyp, 0.0)
    up_z = np.where(kk < NZ, czp, 0.0)
    ux, uy, uz = f(up_x)[:-1], f(up_y)[:-NX], f(up_z)[:-NX * NY]
    A = sp.diags([f(d), -ux, -ux, -uy, -uy, -uz, -uz], [0, 1, -1, NX, -NX, NX * NY, -NX * NY],
                 shape=(n, n), format="csr")
    return A, f(d)


def cg_lanczos(A, iters, seed=1, every=250, log=print):
    """Ritz values of A (symmetric positive definite) from k CG steps on a
    random right-hand side; yields (k, smallest, largest) every `every`
    steps."""
    from scipy.linalg import eigvalsh_tridiagonal
    rng = np.random.default_rng(seed)
    n = A.shape[0]
    b = rng.standard_normal(n)
    x = np.zeros(n)
    r = b.copy()
    p = r.copy()
    rr = r @ r
    alphas, betas = [], []
    out = []
    for k in range(1, iters + 1):
        q = A @ p
        a = rr / (p @ q)
        x += a * p
        r -= a * q
        rr_new = r @ r
        bta = rr_new / rr
        alphas.append(a)
        betas.append(bta)
        rr = rr_new
        p = r + bta * p
        if k % every == 0 or k == iters:
            al, be = np.array(alphas), np.array(betas)
            diag = 1.0 / al
            diag[1:] += be[:-1] / al[:-1]
            off = np.sqrt(be[:-1]) / al[:-1]
            w = eigvalsh_tridiagonal(diag, off)
            out.append((k, w[0], w[-1]))
            log("  k %6d  ritz min %.6e  max %.6e  ratio %.4e  rel res %.3e"
                % (k, w[0], w[-1], w[-1] / w[0], np.sqrt(rr) / np.linalg.norm(b)))
    return out


def _ritz(alphas, betas):
    """Extreme eigenvalues of the Lanczos tridiagonal matrix that k CG steps
    with coefficients α_1..α_k, β_1..β_k build: diagonal 1/α_i +
    β_{i-1}/α_{i-1}, off-diagonal sqrt(β_i)/α_i."""
    from scipy.linalg import eigvalsh_tridiagonal
    al, be = np.asarray(alphas, dtype=np.float64), np.asarray(betas, dtype=np.float64)
    # CG on the negative definite A itself has α < 0: the Ritz values of -A
    al = np.abs(al)
    diag = 1.0 / al
    diag[1:] += be[:-1] / al[:-1]
    off = np.sqrt(be[:-1]) / al[:-1]
    n = len(diag)
    lo = eigvalsh_tridiagonal(diag, off, select="i", select_range=(0, 0))[0]
    hi = eigvalsh_tridiagonal(diag, off, select="i", select_range=(n - 1, n - 1))[0]
    return lo, hi


def device_pcg_lanczos(wk, iters, smoother, seed=1, every=250, log=print, dt=1.0, is_steady=True, rtol=1e-13):
    """Ritz values of M^-1 A from preconditioned CG run on the device with
    the library's own CalcAX!, Preconditioner! (GS(5) or none) and exactly
    rounded dots, on a random right-hand side (interior cells, zero halo).

    wk is a WorkBuffers with the model loaded; its fields pcg_r, pcg_p,
    pcg_p_ and pcg_q are overwritten.  A (negative definite) and M^-1 have
    the same sign, so α, β > 0 and the Ritz values approximate the spectrum
    of M^-1 A (without a preconditioner, of -A).  Five red-black sweeps
    from zero are not a symmetric preconditioner, so for GS(5) the values
    are estimates of the spectrum the reference's CG! iterates with, not
    exact Lanczos bounds, and the run stops if (r, M^-1 r) changes sign.
    Yields (k, smallest, largest, relative residual) every `every` steps
    and at the last step."""
    rng = np.random.default_rng(seed)
    b = np.zeros(wk.shape, order="F")
    b[1:-1, 1:-1, 1:-1] = rng.standard_normal(tuple(m - 2 for m in wk.shape))
    wk.upload("pcg_r", b)
    r0 = np.sqrt(wk.fdot1("pcg_r"))
    wk.fill("pcg_p", 0.0)
    wk.fill("pcg_q", 0.0)
    alphas, betas, out = [], [], []
    rz_old = None
    for k in range(1, iters + 1):
        if smoother is None:
            wk.copy("pcg_p_", "pcg_r")
        else:
            # Preconditioner! starts from its output's contents: myfill! first,
            # as CG! does (src/NonUniform.jl:157-159, 203-205)
            wk.fill("pcg_p_", 0.0)
            wk.precond("pcg_p_", "pcg_r", smoother, dt, is_steady)
        rz = wk.fdot2("pcg_r", "pcg_p_")
        if rz_old is not None:
            betas.append(rz / rz_old)
            wk.bicg1("pcg_p", "pcg_p_", "pcg_q", betas[-1], 0.0)   # p = z + β p
        else:
            wk.copy("pcg_p", "pcg_p_")
        wk.calc_ax("pcg_q", "pcg_p", dt, is_steady)
        pq = wk.fdot2("pcg_p", "pcg_q")
        a = rz / pq
        if not (np.isfinite(a) and a != 0.0 and (not betas or betas[-1] >= 0.0)):
            log("  k %6d  stopped: (r, z) = %.6e, (p, Ap) = %.6e, beta %s"
                % (k, rz, pq, "%.6e" % betas[-1] if betas else "-"))
            if betas:
                betas.pop()
            break
        alphas.append(a)
        wk.triad("pcg_r", "pcg_q", "pcg_r", -a)                    # Triad!: r = -α q + r
        rz_old = rz
        rel = np.sqrt(wk.fdot1("pcg_r")) / r0
        last = k == iters or rel < rtol
        if k % every == 0 or last:
            lo, hi = _ritz(alphas, (betas + [0.0])[:len(alphas)])
            out.append((k, lo, hi, rel))
            log("  k %6d  ritz min %.6e  max %.6e  ratio %.4e  rel res %.3e" % (k, lo, hi, hi / lo, rel))
        if last:
            break
    if alphas and (not out or out[-1][0] != len(alphas)):
        lo, hi = _ritz(alphas, (betas + [0.0])[:len(alphas)])
        out.append((len(alphas), lo, hi, rel))
        log("  k %6d  ritz min %.6e  max %.6e  ratio %.4e  rel res %.3e" % (len(alphas), lo, hi, hi / lo, rel))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--nx", type=int, default=240)
    ap.add_argument("--nz", type=int, default=30)
    ap.add_argument("--power", default="map")
    ap.add_argument("--iters", type=int, default=3000)
    ap.add_argument("--every", type=int, default=250)
    ap.add_argument("--device", action="store_true",
                    help="CG on the device through the library's operators: M^-1 A for GS(5) and A itself")
    a = ap.parse_args(argv)
    if a.device:
        from . import _lib as L
        from .heat3ds import WorkBuffers, build_model, load_model
        m = build_model(a.nx, a.nx, a.nz, power=a.power)
        with WorkBuffers(*m.shape) as wk:
            load_model(wk, m)
            for name, sm in (("M^-1 A, M = GS(5)", L.SMOOTHER_GS), ("A (no preconditioner)", None)):
                print("%dx%dx%d %s:" % (a.nx, a.nx, a.nz, name), flush=True)
                t0 = time.time()
                res = device_pcg_lanczos(wk, a.iters, sm, every=a.every, log=lambda s: print(s, flush=True))
                k, lo, hi, rel = res[-1]
                print("  after %d steps (%.1f s): kappa >= %.4e, kappa * 2^-24 >= %.3e"
                      % (k, time.time() - t0, hi / lo, hi / lo * 2.0 ** -24), flush=True)
        return 0
    from .heat3ds import build_model
    t0 = time.time()
    m = build_model(a.nx, a.nx, a.nz, power=a.power)
    A, d = assemble(m)
    print("%dx%dx%d: %d unknowns, assembled in %.1f s; diagonal spans %.3e .. %.3e"
          % (a.nx, a.nx, a.nz, A.shape[0], time.time() - t0, d.min(), d.max()), flush=True)
    import scipy.sparse as sp
    for name, M in (("-A", A), ("D^-1/2 (-A) D^-1/2", sp.diags(1 / np.sqrt(d)) @ A @ sp.diags(1 / np.sqrt(d)))):
        print("%s:" % name, flush=True)
        res = cg_lanczos(M.tocsr(), a.iters, every=a.every, log=lambda s: print(s, flush=True))
        k, lo, hi = res[-1]
        print("  after %d steps: kappa >= %.4e, kappa * 2^-24 >= %.3e" % (k, hi / lo, hi / lo * 2.0 ** -24),
              flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
