"""Configuration 5 at its full size: the 960x960x120 modelA stack (110.6 M
cells) with the seeded power map scaled by 1 + 0.5·sin(2πt/20 ms),
implicit-Euler steps of Δt = 1 ms from 300 K (the reference's step is
calRHS! then the Krylov solve with is_steady = false, src/heat3ds.jl:122-151).
Two steps, each an FP32-refined CG+GS(5) solve to a true relative residual of
1e-8 (per-colour sweep launches, four FP32 cells per thread), a checkpoint
every 2 steps (so the adjoint recomputes step 1 from θ⁰), the adjoint of
the time-integrated smooth maximum, and one central difference (relative
step 1e-6) on a layer conductivity.  Checks: every step's true FP64
residual is at most 1e-8 relative or at the rounding floor of its
evaluation; the recomputed state is used; the adjoint gradient agrees with
the central difference within 1e-5 relative; device memory is recorded."""
import time

import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="no HIP device")]

import h2_amd
from h2_amd import _lib as L
from h2_amd.transient import TransientRun

FD_GROUP = 8


@pytest.mark.timeout(480)  # about 95 s on one MI355X
def test_c5_two_steps_adjoint_and_central_difference():
    print("case 960x960x120 power map, 2 steps of 1 ms, FP32-refined CG+GS(5), checkpoints every 2", flush=True)
    t0 = time.perf_counter()
    m = h2_amd.build_model(960, 960, 120, power="map")
    print("model built in %.1f s" % (time.perf_counter() - t0), flush=True)
    run = TransientRun(m, dt=1e-3, nsteps=2, checkpoint_every=2, tol=1e-8, solver=L.SOLVER_CG,
                       precision=L.FP32_REFINED, true_residual=True, verbose=True)
    try:
        used0, total = run.wk.mem_info()
        J = run.forward()
        for n, rel, ulps, at_floor in run.stats["true_residual"]:
            print("step %d true relative residual %.3e (%.2f u|s|, floor %s)" % (n, rel, ulps, at_floor), flush=True)
            assert rel <= 1e-8 or at_floor
        g = run.adjoint()
        assert len(run.stats["recompute_iters"]) == 1 and len(run.stats["adjoint_iters"]) == 2
        peak = run.stats["mem_peak"]
        print("J_total %.10f, forward %.1f s, adjoint %.1f s, device memory %.1f GB before, %.1f GB peak"
              % (J, run.stats["forward_s"], run.stats["adjoint_s"], used0 / 1e9, peak / 1e9), flush=True)
        assert used0 <= peak < total
        t1 = time.perf_counter()
        fd = run.central_difference(FD_GROUP, 1e-6, precision=L.FP32_REFINED)
        a = g["dJ_dscale"][FD_GROUP]
        rel = abs(a - fd) / abs(fd)
        print("group %d adjoint %.12e central difference %.12e rel %.2e (%.1f s)"
              % (FD_GROUP, a, fd, rel, time.perf_counter() - t1), flush=True)
        assert rel <= 1e-5
    finally:
        run.close()
