"""The transient adjoint (implicit-Euler steps, time-integrated smooth
maximum), pinned on the CPU with the oracle:
  * ∂b/∂θ_old probed from the oracle's calRHS! is diagonal, equals
    -ρ·cp·V/Δt away from convective faces, and differs from it by -h·A on
    the minus faces' cells and +h·A on the plus faces' cells (the sign
    asymmetry of src/RHS.jl: minus faces subtract h·A·(θ∞ - θ), plus faces
    add it, before calRHS! negates b);
  * the adjoint recursion's gradients against Richardson-extrapolated full
    central differences of J_total (re-running every step).
Small case: 12x10x30 interior cells, 4 steps of Δt = 1 ms."""
import numpy as np
import pytest

import h2_amd
from h2_amd import sensitivity as S

NSTEPS, DT = 4, 1.0e-3


@pytest.fixture(scope="module")
def problem(oracle):
    from oracle import adjoint as OA
    m = h2_amd.build_model(12, 10, 30)
    prob = OA.TransientProblem(12, 10, 30, S.layer_groups(m), DT, NSTEPS)
    return OA, prob, m


def test_rhs_derivative_is_capacity_and_convection(problem, oracle):
    OA, prob, m = problem
    c = prob.case(np.ones(S.NGROUPS))
    D = OA.probe_rhs_derivative(c)
    assert (D - D.multiply(np.eye(D.shape[0], dtype=bool))).nnz == 0
    d = D.diagonal().reshape((m.NX, m.NY, m.NZ), order="F")
    I = (slice(1, -1),) * 3
    V = m.dx * m.dy * m.dZ[None, None, 1:-1]
    cap = m.rho[I] * m.cp[I] * V / DT
    # modelA BC mode 3: x±, y±, z+ convective (h = 5), z- isothermal
    h = 5.0
    expect = -cap.copy()
    areas = {0: m.dy * m.dZ[1:-1][None, :], 1: m.dx * m.dZ[1:-1][None, :], 2: m.dx * m.dy}
    expect[0, :, :] -= h * areas[0]
    expect[-1, :, :] += h * areas[0]
    expect[:, 0, :] -= h * areas[1]
    expect[:, -1, :] += h * areas[1]
    expect[:, :, -1] += h * areas[2]
    assert np.max(np.abs(d - expect) / np.abs(expect)) < 1e-9


def transient_fd(prob, beta, perturb, h):
    """Richardson-extrapolated central differences of J_total, each step's
    J difference without cancellation."""
    def cd(step):
        (sa, ha), (sb, hb) = perturb(step), perturb(-step)
        _, tp, _, _, _, _ = prob.run(sa, beta, ha)
        _, tm, _, _, _, _ = prob.run(sb, beta, hb)
        return sum(prob.dt * S.delta_J(tp[n], tm[n], prob.region, beta) for n in range(1, prob.nsteps + 1)) \
            / (2 * step)
    return (4 * cd(h) - cd(2 * h)) / 3


def test_transient_adjoint_matches_full_central_differences(problem):
    OA, prob, m = problem
    beta = S.BETA
    ones = np.ones(S.NGROUPS)
    groups = [2, 4, 9, 12]
    adj = prob.adjoint(ones, beta, groups=groups)
    scale = np.max(np.abs(adj["dJ_dscale"]))
    for k in groups:
        def perturb(d, k=k):
            sc = ones.copy()
            sc[k] += d
            return sc, None
        fd = transient_fd(prob, beta, perturb, 1e-2)
        a = adj["dJ_dscale"][k]
        assert abs(a - fd) <= 2e-6 * max(abs(fd), 1e-3 * scale), (k, a, fd)
    src = np.argwhere(prob.hsrc > 0)
    for ci in (src[0], src[len(src) // 2]):
        ci = tuple(ci)
        def perturb(d, ci=ci):
            hs = prob.hsrc.copy(order="F")
            hs[ci] += d * prob.hsrc[ci]
            return ones, hs
        fd = transient_fd(prob, beta, perturb, 1e-2) / prob.hsrc[ci]
        assert abs(adj["dJ_dhsrc"][ci] - fd) <= 2e-6 * abs(fd), (ci, adj["dJ_dhsrc"][ci], fd)
