"""The condition estimates of h2_amd.spectrum: Lanczos through CG's α, β
against dense eigenvalues, on a small modelA operator (CPU)."""
import numpy as np
import pytest

from h2_amd import spectrum
from h2_amd.heat3ds import build_model


class _HostFields:
    """The WorkBuffers operations device_pcg_lanczos uses, on the host: A
    the assembled operator, the preconditioner a Jacobi solve (symmetric,
    so Lanczos bounds are exact)."""

    def __init__(self, m, A, d):
        self.NX, self.NY, self.NZ = m.NX, m.NY, m.NZ
        self.shape = (m.NX + 2, m.NY + 2, m.NZ + 2)
        self.A, self.d, self.f = A, d, {}

    def _i(self, n):
        return self.f[n][1:-1, 1:-1, 1:-1].ravel(order="F")

    def _set(self, n, v):
        a = np.zeros(self.shape, order="F")
        a[1:-1, 1:-1, 1:-1] = v.reshape((self.NX, self.NY, self.NZ), order="F")
        self.f[n] = a

    def upload(self, n, a):
        self.f[n] = np.asfortranarray(np.array(a, dtype=np.float64))

    def fdot1(self, n):
        v = self._i(n)
        return float(v @ v)

    def fdot2(self, a, b):
        return float(self._i(a) @ self._i(b))

    def fill(self, n, v):
        self.f[n] = np.full(self.shape, float(v), order="F")

    def copy(self, a, b):
        self.f[a] = self.f[b].copy()

    def precond(self, z, r, smoother, dt, is_steady):
        self._set(z, self._i(r) / (-self.d))

    def bicg1(self, p, r, q, beta, omg):
        self._set(p, self._i(r) + beta * (self._i(p) - omg * self._i(q)))

    def calc_ax(self, q, p, dt, is_steady):
        self._set(q, self.A @ self._i(p))

    def triad(self, z, x, y, a):
        # Triad! (src/NonUniform.jl:627-639): z = a x + y
        self._set(z, a * self._i(x) + self._i(y))


@pytest.fixture(scope="module")
def small():
    m = build_model(8, 6, 30, power="map")
    negA, d = spectrum.assemble(m)
    return m, negA, d


def test_lanczos_extremes_of_jacobi_preconditioned_operator(small):
    m, negA, d = small
    wk = _HostFields(m, -negA, d)
    out = spectrum.device_pcg_lanczos(wk, 3000, 1, every=10 ** 6, log=lambda s: None)
    k, lo, hi, rel = out[-1]
    D = np.diag(1.0 / np.sqrt(d))
    ev = np.linalg.eigvalsh(D @ negA.toarray() @ D)
    assert rel < 1e-12
    assert lo == pytest.approx(ev[0], rel=1e-8)
    assert hi == pytest.approx(ev[-1], rel=1e-8)


def test_lanczos_extremes_of_the_operator_itself(small):
    m, negA, d = small
    wk = _HostFields(m, -negA, d)
    out = spectrum.device_pcg_lanczos(wk, 6000, None, every=10 ** 6, log=lambda s: None)
    k, lo, hi, rel = out[-1]
    ev = np.linalg.eigvalsh(negA.toarray())
    assert lo == pytest.approx(ev[0], rel=1e-8)
    assert hi == pytest.approx(ev[-1], rel=1e-8)
    # the assembled matrix is -A: symmetric positive definite
    assert ev[0] > 0
