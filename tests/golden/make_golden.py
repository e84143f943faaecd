"""Writes the golden fixtures under tests/golden/ from the C oracle.

  python tests/golden/make_golden.py [--threads N]

Fixtures (all FP64, modelA, BC mode 3, steady, Δt = 1e4 s):
  readme_pins.json        the reference README's published log
                          (README.md:166-183), typed in verbatim
  c1_<solver>_<mode>.npz  40x40x30, ε = 1e-6: residual history, res0, iteration
                          count, θ statistics and the SHA-256 of the full θ
  c2_<solver>_<mode>.npz  240x240x30, PBiCGSTAB+GS ε = 1e-4 and CG+GS ε = 1e-6:
                          residual history, res0, iteration count, θ
                          min/max/L2, the SHA-256 of the full θ array, the
                          centre Z line and every 8th cell of three XY planes
  <mode> is "seq" (the reference's sequential grid sums) or "exact"
  (exactly rounded grid sums, what the device library computes).
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

README_PINS = {
    "source": "README.md:166-183 of kenoogl/H2 (240x240x30, pbicgstab + gs, eps = 1e-4, steady)",
    "grid": [242, 242, 32],
    "res0": "0.002771608776496581",
    "history": {"1": "6.58308938608891E-01", "2": "8.89700596995947E-01", "208": "9.72579214884977E-05"},
    "converged_at": 208,
    "theta_min_K": "300.6",
    "theta_max_K": "349.4",
    "seconds": 183,
}


def theta_digest(th):
    return hashlib.sha256(np.asfortranarray(th).tobytes(order="F")).hexdigest()


def run(nx, nz, solver, tol, mode):
    with O.sum_mode(mode):
        c = O.modelA_case(nx, nx, nz)
        c.cal_rhs()
        t0 = time.time()
        conv, itr, res0, hist = c.solve(solver, tol=tol, smoother=1)
        dt = time.time() - t0
    th = c.arr["theta"]
    s = th[1:-1, 1:-1, 1:-1]
    out = dict(converged=conv, itr=itr, res0=res0, history=hist, theta_min=s.min(), theta_max=s.max(),
               theta_l2=np.linalg.norm(s.ravel()), theta_sha256=theta_digest(th), b_sha256=theta_digest(c.arr["b"]))
    print("%dx%dx%d %s %s: itr %d res0 %.17g θmax %.6f (%.1f s)" % (nx, nx, nz, solver, "exact" if mode else "seq",
                                                                   itr, res0, s.max(), dt), flush=True)
    return out, th


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--skip-c2", action="store_true")
    a = ap.parse_args()
    O.set_threads(a.threads)
    with open(os.path.join(HERE, "readme_pins.json"), "w") as f:
        json.dump(README_PINS, f, indent=1)
    for solver in ("pbicgstab", "cg"):
        for mode, tag in ((O.SEQUENTIAL, "seq"), (O.EXACT, "exact")):
            out, th = run(40, 30, solver, 1e-6, mode)
            out["tol"] = 1e-6
            np.savez_compressed(os.path.join(HERE, "c1_%s_%s.npz" % (solver, tag)), **out)
    if a.skip_c2:
        return
    for solver, tol in (("pbicgstab", 1e-4), ("cg", 1e-6)):
        for mode, tag in ((O.SEQUENTIAL, "seq"), (O.EXACT, "exact")):
            out, th = run(240, 30, solver, tol, mode)
            mid = th.shape[0] // 2
            out["zline_center"] = th[mid, mid, :].copy()
            out["xy_planes_k"] = np.array([5, 13, 28])
            out["xy_planes"] = th[::8, ::8, [5, 13, 28]].copy()
            out["tol"] = tol
            np.savez_compressed(os.path.join(HERE, "c2_%s_%s.npz" % (solver, tag)), **out)


if __name__ == "__main__":
    main()
