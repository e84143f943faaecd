"""GPU parity of the sensitivity extension (smooth-max objective, adjoint
solve, gradients): against the oracle's CPU adjoint on a small modelA case,
and against central finite differences on the 240x240x30 case with the
seeded power map (relative step 1e-6, 13 conductivity groups and 8 seeded
heat-source cells, agreement within 1e-5 relative)."""
import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="no HIP device")]

import h2_amd
from h2_amd import _lib as L
from h2_amd import sensitivity as S


def test_objective_adjoint_and_gradients_vs_oracle(oracle):
    from oracle import adjoint as OA
    m = h2_amd.build_model(12, 10, 30)
    sens = S.SteadySensitivity(m, tol=1e-13)
    J = sens.forward()
    th = sens.theta()
    Jo, go = OA.smooth_max(th, sens.region, S.BETA)
    assert abs(J - Jo) <= 1e-13 * abs(Jo)
    g = sens.adjoint()
    mu, seed = sens.wk.adjoint_fields()
    assert np.max(np.abs(seed - go)) <= 1e-14
    prob = OA.SteadyProblem(12, 10, 30, sens.groups)
    ref = prob.adjoint(np.ones(S.NGROUPS), S.BETA)
    assert np.max(np.abs(th - ref["theta"])) < 1e-9
    assert np.max(np.abs(mu - ref["mu"])) <= 1e-8 * np.max(np.abs(ref["mu"]))
    # the oracle's operator derivative is an extrapolated difference (~1e-7)
    d, dr = g["dJ_dscale"], ref["dJ_dscale"]
    for k in range(S.NGROUPS):
        print("group", k, d[k], dr[k], abs(d[k] - dr[k]) / max(abs(dr[k]), 1e-300))
    assert np.max(np.abs(d - dr)) <= 1e-6 * np.max(np.abs(dr))
    assert np.max(np.abs(g["dJ_dhsrc"] - ref["dJ_dhsrc"])) <= 1e-8 * np.max(np.abs(ref["dJ_dhsrc"]))
    sens.close()


def test_torch_function_backward_is_the_adjoint():
    torch = pytest.importorskip("torch")
    m = h2_amd.build_model(12, 10, 30)
    sens = S.SteadySensitivity(m, tol=1e-13)
    scales = torch.ones(S.NGROUPS, dtype=torch.float64, requires_grad=True)
    hsrc = torch.from_numpy(np.array(m.hsrc)).requires_grad_(True)
    J = S.SmoothMaxTemperature.apply(scales, hsrc, sens)
    J.backward()
    sens2 = S.SteadySensitivity(m, tol=1e-13)
    sens2.forward()
    g = sens2.adjoint()
    assert np.allclose(scales.grad.numpy(), g["dJ_dscale"], rtol=1e-12, atol=0)
    assert np.allclose(hsrc.grad.numpy(), g["dJ_dhsrc"], rtol=1e-12, atol=0)
    sens.close()
    sens2.close()


def test_c4_adjoint_vs_central_differences():
    m = h2_amd.build_model(240, 240, 30, power="map")
    sens = S.SteadySensitivity(m, tol=1e-12)
    sens.forward()
    th0 = sens.theta()
    g = sens.adjoint()
    # the finite-difference solves: flux-form operator (rounding relative to
    # temperature differences), started from the unperturbed solution
    sens.wk.set_operator_form(L.FORM_FLUX)
    errs = []
    for k in range(S.NGROUPS):
        if not np.any(sens.groups == k):
            continue
        fd = sens.central_difference("scale", k, 1e-6, theta0=th0)
        errs.append(("scale", k, g["dJ_dscale"][k], fd))
    rng = np.random.default_rng(20251203)
    src = np.argwhere(m.hsrc > 0)
    for ci in src[rng.choice(len(src), 8, replace=False)]:
        ci = tuple(int(v) for v in ci)
        fd = sens.central_difference("hsrc", ci, 1e-6, theta0=th0)
        errs.append(("hsrc", ci, g["dJ_dhsrc"][ci], fd))
    for kind, idx, a, fd in errs:
        print(kind, idx, a, fd, abs(a - fd) / abs(fd))
    worst = max(abs(a - fd) / abs(fd) for _, _, a, fd in errs)
    assert worst <= 1e-5, errs
    # the difference solves replaced the forward solution on the device
    with pytest.raises(RuntimeError):
        sens.adjoint()
    sens.close()
