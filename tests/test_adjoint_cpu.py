"""The adjoint method of the sensitivity extension, pinned on the CPU with
the oracle: gradients by the adjoint (a direct solve of Aᵀμ = ∂J/∂θ with a
central difference of the oracle's operator application) against full
central differences of J that re-solve the perturbed problem, for the 13
conductivity-group scales and for single heat-source cells.  Also the
probed operator against CalcAX! and the objective's closed forms."""
import numpy as np
import pytest
from scipy.special import logsumexp

import h2_amd
from h2_amd import sensitivity as S


@pytest.fixture(scope="module")
def problem(oracle):
    from oracle import adjoint as OA
    m = h2_amd.build_model(12, 10, 30)
    prob = OA.SteadyProblem(12, 10, 30, S.layer_groups(m))
    return OA, prob, m


def test_probed_matrix_reproduces_calc_ax(problem, oracle):
    OA, prob, m = problem
    c = prob.case(np.ones(S.NGROUPS))
    A = OA.probe_matrix(c)
    x = np.zeros(c.shape, order="F")
    x[1:-1, 1:-1, 1:-1] = np.random.default_rng(1).normal(size=x[1:-1, 1:-1, 1:-1].shape)
    ax = np.zeros(c.shape, order="F")
    c.calc_ax(ax, x)
    y = A @ x[1:-1, 1:-1, 1:-1].ravel(order="F")
    assert np.max(np.abs(y - ax[1:-1, 1:-1, 1:-1].ravel(order="F"))) <= 1e-13 * np.max(np.abs(y))
    assert A.nnz <= 7 * A.shape[0]


def test_smooth_max_closed_form():
    from oracle import adjoint as OA
    rng = np.random.default_rng(0)
    th = 300 + 50 * rng.random((6, 5, 4))
    reg = (rng.random((6, 5, 4)) < 0.5).astype(np.uint8)
    J, g = OA.smooth_max(th, reg, 10.0)
    assert abs(J - logsumexp(10.0 * th[reg.astype(bool)]) / 10.0) < 1e-12
    assert abs(g.sum() - 1.0) < 1e-14 and np.all(g[reg == 0] == 0)
    assert th[reg.astype(bool)].max() <= J <= th[reg.astype(bool)].max() + np.log(reg.sum()) / 10.0


def test_groups_cover_interior(problem):
    OA, prob, m = problem
    g = S.layer_groups(m)
    assert np.all(g[1:-1, 1:-1, 1:-1] >= 0) and np.all(g[0] == -1)
    assert set(np.unique(g[1:-1, 1:-1, 1:-1])) <= set(range(S.NGROUPS))
    assert np.all(g[m.ID == 1] == 12)


def richardson_fd(prob, beta, perturb, h):
    """Central differences of J at steps h and 2h, extrapolated (O(h⁴)).
    Large steps keep the rounding of the FP64 solves (~1e-10 K in θ, from
    residuals of temperatures near 300 K) far below the J differences."""
    def cd(step):
        a, b = perturb(step), perturb(-step)
        _, thp, _, _, _ = prob.objective(a[0], beta, a[1])
        _, thm, _, _, _ = prob.objective(b[0], beta, b[1])
        return S.delta_J(thp, thm, prob.region, beta) / (2 * step)
    return (4 * cd(h) - cd(2 * h)) / 3


def test_adjoint_matches_full_central_differences(problem):
    """The adjoint (operator derivative by extrapolated differences at
    relative steps 1e-3, 2e-3) against Richardson-extrapolated central
    differences of J at relative steps 1e-2 and 2e-2."""
    OA, prob, m = problem
    beta = 10.0
    ones = np.ones(S.NGROUPS)
    adj = OA.SteadyProblem.adjoint(prob, ones, beta)
    scale = np.max(np.abs(adj["dJ_dscale"]))
    for k in [k for k in range(S.NGROUPS) if np.any(prob.groups == k)]:
        def perturb(d, k=k):
            sc = ones.copy()
            sc[k] += d
            return sc, None
        fd = richardson_fd(prob, beta, perturb, 1e-2)
        a = adj["dJ_dscale"][k]
        assert abs(a - fd) <= 2e-6 * max(abs(fd), 1e-3 * scale), (k, a, fd)
    src = np.argwhere(prob.hsrc > 0)
    for ci in (src[0], src[len(src) // 2]):
        ci = tuple(ci)
        def perturb(d, ci=ci):
            hs = prob.hsrc.copy(order="F")
            hs[ci] += d * prob.hsrc[ci]
            return ones, hs
        fd = richardson_fd(prob, beta, perturb, 1e-2) / prob.hsrc[ci]
        assert abs(adj["dJ_dhsrc"][ci] - fd) <= 2e-6 * abs(fd), (ci, adj["dJ_dhsrc"][ci], fd)


def _theta_with_case_halo(c, rng):
    """A random interior θ (250-350 K) with the halo the oracle case sets."""
    th = c.arr["theta"].copy(order="F")
    th[1:-1, 1:-1, 1:-1] = 250.0 + 100.0 * rng.random(th[1:-1, 1:-1, 1:-1].shape)
    return th


def test_operator_change_matches_oracle_calc_ax(problem):
    """(A(λp) - A(λ0))·θ of the flux form (operator_change, the change the
    transient central difference applies) against the difference of the
    oracle's CalcAX! of the scaled and unscaled cases on the same θ, for
    every group the case holds: the same within rounding (the convective
    and ap0 terms cancel)."""
    OA, prob, m = problem
    for k in [k for k in range(S.NGROUPS) if np.any(prob.groups == k)]:
        _check_operator_change(prob, m, k, 1.5 if k % 2 else 0.6)


def _check_operator_change(prob, m, k, s):
    ones = np.ones(S.NGROUPS)
    sc = ones.copy()
    sc[k] = s
    c0, cp = prob.case(ones), prob.case(sc)
    th = _theta_with_case_halo(c0, np.random.default_rng(k))
    assert np.array_equal(c0.arr["theta"][0], cp.arr["theta"][0])
    a0 = np.zeros(th.shape, order="F")
    ap = np.zeros(th.shape, order="F")
    c0.calc_ax(a0, th)
    cp.calc_ax(ap, th)
    groups = S.layer_groups(m)
    lam0 = np.array(m.lam, order="F")
    assert np.array_equal(lam0[1:-1, 1:-1, 1:-1], prob.lam0[1:-1, 1:-1, 1:-1])
    d = S.operator_change(m, S.scaled_lambda(m, groups, sc), lam0, th)
    want = (ap - a0)[1:-1, 1:-1, 1:-1]
    got = d[1:-1, 1:-1, 1:-1]
    assert np.max(np.abs(want)) > 0, k
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(ap)), k


@pytest.mark.parametrize("k", [0, 5, 8, 11])
def test_conductance_change_box_equals_full(problem, k):
    """The changes formed over the changed box only are bitwise those over
    the whole interior (entries and their application to θ)."""
    OA, prob, m = problem
    sc = np.ones(S.NGROUPS)
    sc[k] = 1.0 + 1e-6
    groups = S.layer_groups(m)
    lam0 = np.array(m.lam, order="F")
    lam = S.scaled_lambda(m, groups, sc)
    full = S.conductance_change(m, lam, lam0, box=False)
    boxed = S.conductance_change(m, lam, lam0)
    th = _theta_with_case_halo(prob.case(np.ones(S.NGROUPS)), np.random.default_rng(7))
    assert np.array_equal(S.apply_conductance_change(m, full, th), S.apply_conductance_change(m, boxed, th))
    assert len(full) == len(boxed) == 6
    for (nb_f, sel_f, v_f), (nb_b, sel_b, v_b) in zip(full, boxed):
        assert nb_f == nb_b
        assert all(np.array_equal(a, b) for a, b in zip(sel_f, sel_b))
        assert np.array_equal(v_f, v_b)
    assert S.conductance_change(m, lam0, lam0) == []
