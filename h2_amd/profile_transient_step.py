"""Workload for the profiles of the 960x960x120 transient case.

Run under rocprofv3 (kernel trace, or one counter set per run):
  rocprofv3 --kernel-trace --stats -d OUT -o c5 -- python -m h2_amd.profile_transient_step
  rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT -o f -- python -m h2_amd.profile_transient_step
It builds the 960x960x120 case with the seeded power map, forms the
right-hand side of one implicit-Euler step (dt = 1 ms), runs the first 40
iterations of its solve (every kernel of an iteration in its steady-state
proportions; a whole step is about 750 FP64 PBiCGSTAB+GS(5) iterations or
1300 inner iterations of the FP32-refined CG+GS(5) solve), makes three
plain device copies of one H-cell field (the counters' calibration) and
times 10 full red-black sweeps.  --solver cg --precision fp32 profiles the
FP32-refined CG solve instead (its first inner solve: 40 iterations end
the refinement loop).
"""
import argparse
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import h2_amd
from h2_amd import _lib as L
from h2_amd.heat3ds import WorkBuffers, load_model

ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
ap.add_argument("--solver", choices=("pbicgstab", "cg"), default="pbicgstab")
ap.add_argument("--precision", choices=("fp64", "fp32"), default="fp64")
a = ap.parse_args()
solver = L.SOLVER_CG if a.solver == "cg" else L.SOLVER_PBICGSTAB
precision = L.FP32_REFINED if a.precision == "fp32" else L.FP64

m = h2_amd.build_model(960, 960, 120, power="map")
wk = WorkBuffers(*m.shape)
load_model(wk, m)
wk.fill("b", 0.0)
wk.calc_rhs(1e-3, False)
conv, itr, res0 = wk.solve(solver, 1e-3, 1e-8, 40, L.SMOOTHER_GS, False, precision=precision)
print("solver", a.solver, "precision", a.precision, "iterations", itr, "device ms", wk.last_solve_ms(), flush=True)
for _ in range(3):
    wk.copy("pcg_p", "theta")      # calibration: a plain copy of H*8 bytes
print("sweep", wk.time_kernel(0, 10, 1e-3, False), flush=True)
wk.close()
