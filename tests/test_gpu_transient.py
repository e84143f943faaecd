"""GPU parity of the transient adjoint (implicit-Euler steps, time-integrated
smooth maximum, θ checkpoints with recompute):
  * against the oracle's direct-solve forward and adjoint recursion on a
    small modelA case (12x10x30, 4 steps of 1 ms);
  * checkpoint spacing changes nothing (recomputed states are bitwise the
    forward's, so the gradients are bitwise equal);
  * against central finite differences (relative step 1e-6, agreement
    within 1e-5 relative) on 4 layer conductivities at 240x240x30 with the
    seeded power map, 20 steps of 1 ms, checkpoints every 5 steps; the
    perturbed runs are computed through their changes from the unperturbed
    states (TransientRun.central_difference);
  * FP32-refined steps (forward and adjoint to 1e-8) against FP64 ones."""
import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="no HIP device")]

import h2_amd
from h2_amd import _lib as L
from h2_amd import sensitivity as S
from h2_amd.transient import TransientRun

FD_GROUPS = (2, 5, 8, 11)


def test_transient_adjoint_vs_oracle(oracle):
    from oracle import adjoint as OA
    m = h2_amd.build_model(12, 10, 30)
    run = TransientRun(m, dt=1e-3, nsteps=4, checkpoint_every=2, tol=1e-13)
    J = run.forward()
    g = run.adjoint()
    prob = OA.TransientProblem(12, 10, 30, run.groups, 1e-3, 4)
    ref = prob.adjoint(np.ones(S.NGROUPS), S.BETA)
    print("J", J, ref["J"])
    assert abs(J - ref["J"]) <= 1e-12 * abs(ref["J"])
    d, dr = g["dJ_dscale"], ref["dJ_dscale"]
    for k in range(S.NGROUPS):
        print("group", k, d[k], dr[k], abs(d[k] - dr[k]) / max(abs(dr[k]), 1e-300))
    assert np.max(np.abs(d - dr)) <= 1e-6 * np.max(np.abs(dr))
    gh, ghr = g["dJ_dhsrc"], ref["dJ_dhsrc"]
    print("hsrc max rel", np.max(np.abs(gh - ghr)) / np.max(np.abs(ghr)))
    assert np.max(np.abs(gh - ghr)) <= 1e-8 * np.max(np.abs(ghr))
    run.close()


def test_checkpoint_spacing_is_exact():
    m = h2_amd.build_model(24, 20, 30)
    out = []
    for every in (1, 3, 7):
        run = TransientRun(m, dt=1e-3, nsteps=7, checkpoint_every=every, tol=1e-10)
        J = run.forward()
        g = run.adjoint()
        out.append((J, g))
        print("every", every, "J", J, "recompute", len(run.stats["recompute_iters"]))
        run.close()
    for J, g in out[1:]:
        assert J == out[0][0]
        assert np.array_equal(g["dJ_dlambda"], out[0][1]["dJ_dlambda"])
        assert np.array_equal(g["dJ_dhsrc"], out[0][1]["dJ_dhsrc"])


def test_transient_adjoint_vs_central_differences():
    m = h2_amd.build_model(240, 240, 30, power="map")
    run = TransientRun(m, dt=1e-3, nsteps=20, checkpoint_every=5, tol=1e-12)
    run.forward()
    g = run.adjoint()
    print("forward s", run.stats["forward_s"], "adjoint s", run.stats["adjoint_s"])
    print("iterations per step", run.stats["forward_iters"])
    thetas = run.states()
    worst = 0.0
    for k in FD_GROUPS:
        fd = run.central_difference(k, 1e-6, thetas)
        a = g["dJ_dscale"][k]
        rel = abs(a - fd) / abs(fd)
        worst = max(worst, rel)
        print("group %2d adjoint % .12e fd % .12e rel %.2e" % (k, a, fd, rel))
    assert worst <= 1e-5
    # the difference runs replaced the checkpointed forward on the device
    with pytest.raises(RuntimeError):
        run.adjoint()
    run.close()


def test_fp32_refined_steps_match_fp64():
    m = h2_amd.build_model(120, 120, 30, power="map")
    res = {}
    for prec in (L.FP64, L.FP32_REFINED):
        run = TransientRun(m, dt=1e-3, nsteps=6, checkpoint_every=3, tol=1e-8, precision=prec)
        J = run.forward()
        g = run.adjoint()
        res[prec] = (J, g["dJ_dscale"], g["dJ_dhsrc"])
        print("precision", prec, "J", J, "forward iters", run.stats["forward_iters"],
              "adjoint iters", run.stats["adjoint_iters"])
        run.close()
    (J64, s64, h64), (J32, s32, h32) = res[L.FP64], res[L.FP32_REFINED]
    assert abs(J32 - J64) <= 1e-9 * abs(J64)
    assert np.max(np.abs(s32 - s64)) <= 1e-6 * np.max(np.abs(s64))
    assert np.max(np.abs(h32 - h64)) <= 1e-6 * np.max(np.abs(h64))


def test_true_residual_log():
    """With true_residual the forward run records, per step, the true
    residual ||b - A theta|| / res0 and its size in units of 2^-53·||s||.
    A converged FP64 step is at 1e-8 of res0 or within a few tens of
    rounding units of it (the recursive residual drifts from the true one
    by rounding); a refined step is below 1e-8 or reports the floor."""
    m = h2_amd.build_model(40, 40, 30, power="map")
    for prec in (L.FP64, L.FP32_REFINED):
        run = TransientRun(m, dt=1e-3, nsteps=3, checkpoint_every=3, tol=1e-8, precision=prec,
                           solver=L.SOLVER_CG, true_residual=True)
        run.forward()
        tr = run.stats["true_residual"]
        print("precision", prec, tr)
        assert [t[0] for t in tr] == [1, 2, 3]
        for n, rel, ulps, at_floor in tr:
            assert np.isfinite(rel) and np.isfinite(ulps) and ulps > 0
            if prec == L.FP64:
                assert rel < 1e-8 or ulps < 64
                assert not at_floor
            else:
                assert rel < 1e-8 or (at_floor and ulps <= 9.0)
        run.close()
