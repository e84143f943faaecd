"""Workload for the HBM-traffic counters of the red-black sweep kernel.

Run under rocprofv3 twice, once per counter (one pass each):
  rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT -o f -- python -m h2_amd.profile_sweep_traffic 240
  rocprofv3 --pmc WRITE_SIZE --output-format csv -d OUT -o w -- python -m h2_amd.profile_sweep_traffic 240
It builds the modelA case (240x240x30, or 480x480x60 with argument 480),
runs a 3-iteration solve so every Krylov vector holds data, makes three
plain device copies of one H-cell FP64 field (the calibration: H*8 bytes
read and written by the runtime's copy kernel), then times 20 full
red-black sweeps (40 launches of k_rb_sweep<double>).
"""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import h2_amd
from h2_amd.heat3ds import WorkBuffers, load_model
from h2_amd import _lib as L
nx = int(sys.argv[1]) if len(sys.argv) > 1 else 240
nz = 30 if nx == 240 else 60
m = h2_amd.build_model(nx, nx, nz); wk = WorkBuffers(*m.shape); load_model(wk, m)
wk.calc_rhs(10000.0, True)
wk.solve(L.SOLVER_PBICGSTAB, 10000.0, 1e-30, 3, 1, True)
for _ in range(3):
    wk.copy("pcg_p", "theta")      # calibration: a plain copy of H*8 bytes
print("sweep", wk.time_kernel(0, 20, 10000.0, True), flush=True)
