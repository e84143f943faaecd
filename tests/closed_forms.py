"""Small boundary-condition cases whose steady or one-step solutions have a
closed form, shared by the oracle's pins (tests/test_oracle.py) and the
device's (tests/test_gpu_closed_forms.py).

Each case gives the grid, the six faces (x-, x+, y-, y+, z-, z+), the
interior λ, ρ, cp and source Q, the initial θ, and the expected interior θ
as a function of the grid metrics.  The sign and area of the heat-flux
right-hand side (src/RHS.jl:83-134: minus faces add q·A, plus faces
subtract it, before calRHS! negates b), the a_p_0 = ρ·cp·V/Δt term of the
transient operator (src/NonUniform.jl:86, 498) and the copy of the interior
λ into an isothermal halo (src/boundary_conditions.jl:156-215) each change
the answer of at least one case by O(1).
"""
from dataclasses import dataclass, field
from typing import Callable, List

import numpy as np

from h2_amd.heat3ds import adiabatic_bc, heat_flux_bc, isothermal_bc


@dataclass
class Case:
    NX: int
    NY: int
    NZ: int
    faces: List
    lam: object                    # scalar or (NX, NY, NZ) interior array
    expected: Callable = None      # (dx, dy, ZC) -> interior θ (broadcastable)
    tol: float = 0.0               # max |θ - expected| allowed (K)
    steady: bool = True
    dt: float = 10000.0
    theta0: float = 300.0
    rho: float = 1.0
    cp: float = 1.0
    Q: float = 0.0
    start_is_solution: bool = False  # then the residual of θ0 is checked instead of a solve
    energy: Callable = None        # optional (θ, dx, dy, ZC) -> max relative energy-balance error
    extra: dict = field(default_factory=dict)


def flux_top_column(q):
    """Flux q along +z leaves through the z+ face, isothermal z- at T0,
    adiabatic sides: θ_k = T0 - q·R_k, R_k the resistance (per unit area)
    from the bottom face to the centre of cell k, whose first segment is the
    doubled half-cell conductance of an isothermal z face."""
    NZ, T0, lam = 30, 300.0, 149.0

    def expected(dx, dy, ZC):
        seg = np.empty(NZ)
        seg[0] = (ZC[1] - ZC[0]) / (2 * lam)
        seg[1:] = np.diff(ZC[1:NZ + 1]) / lam
        return (T0 - q * np.cumsum(seg))[None, None, :]

    def energy(th, dx, dy, ZC):
        # heat through the bottom face = q·A, cell by cell
        flow = 2 * lam / (ZC[1] - ZC[0]) * (T0 - th[:, :, 0])
        return np.max(np.abs(flow - q)) / abs(q)

    rise = abs(q) * 4.0e-6  # Σ resistances over the Zcase2 stack at λ = 149 (6e-4 m / 149)
    return Case(3, 4, NZ, [adiabatic_bc(), adiabatic_bc(), adiabatic_bc(), adiabatic_bc(), isothermal_bc(T0),
                           heat_flux_bc(q)], lam, expected, 1e-7 * rise, energy=energy)


def flux_side_line(flux_face):
    """Flux q along the axis through an X or Y face, the opposite face
    isothermal at T1, the rest adiabatic: θ_i = T1 + q·d_i/λ, d_i the signed
    distance (along +axis) from the cell centre to the isothermal face."""
    N, T1, lam, q = 12, 320.0, 38.0, 5.0e4
    faces = [adiabatic_bc() for _ in range(6)]
    axis = 0 if flux_face[0] == "x" else 1
    minus = flux_face[1] == "-"
    f = 2 * axis + (0 if minus else 1)
    faces[f] = heat_flux_bc(q)
    faces[f ^ 1] = isothermal_bc(T1)
    NX, NY = (N, 3) if axis == 0 else (3, N)

    def expected(dx, dy, ZC):
        h = dx if axis == 0 else dy
        centre = (np.arange(N) + 0.5) * h
        dist = (N * h - centre) if minus else -centre
        line = T1 + q * dist / lam
        return line[:, None, None] if axis == 0 else line[None, :, None]

    return Case(NX, NY, 30, faces, lam, expected, 1e-7 * q * 1.2e-3 / lam)  # 1e-7 of the rise q·L/λ


def transient_uniform_source(Q):
    """One implicit-Euler step, every face adiabatic, uniform ρ, cp, Q from a
    uniform θ0: θ1 = θ0 + QΔt/(ρcp) in every cell; for Q = 0 the start θ0 is
    the solution (its residual is the rounding of the stencil sums, against
    a_p_0·θ0 if a_p_0 were dropped on either side)."""
    th0, dt, rho, cp = 300.0, 2.5e-3, 2330.0, 700.0
    rise = Q * dt / (rho * cp)
    return Case(5, 4, 30, [adiabatic_bc() for _ in range(6)], 149.0,
                lambda dx, dy, ZC: np.full((1, 1, 1), th0 + rise), 1e-12 * th0, steady=False, dt=dt,
                theta0=th0, rho=rho, cp=cp, Q=Q, start_is_solution=(Q == 0.0))


def isothermal_two_materials(axis):
    """T0 on the minus face and T1 on the plus face of one axis, other faces
    adiabatic, λ1 in the first half of the line and λ2 in the second: series
    resistances Δ/(2λ) per half cell, Δ/(2λ_edge) at each isothermal face
    with λ_edge the interior λ copied into the halo."""
    N, T0, T1, l1, l2 = 10, 300.0, 345.0, 149.0, 0.9
    faces = [adiabatic_bc() for _ in range(6)]
    faces[2 * axis] = isothermal_bc(T0)
    faces[2 * axis + 1] = isothermal_bc(T1)
    lam_line = np.where(np.arange(N) < N // 2, l1, l2)
    NX, NY, NZ = (N, 2, 30) if axis == 0 else (2, N, 30)
    lam = np.empty((NX, NY, NZ))
    lam[:] = lam_line[:, None, None] if axis == 0 else lam_line[None, :, None]

    def expected(dx, dy, ZC):
        h = dx if axis == 0 else dy
        half = h / (2 * lam_line)
        R = np.concatenate([[half[0]], half[:-1] + half[1:]]).cumsum()
        line = T0 + (T1 - T0) * R / (R[-1] + half[-1])
        return line[:, None, None] if axis == 0 else line[None, :, None]

    return Case(NX, NY, NZ, faces, lam, expected, 1e-9 * (T1 - T0))


CASES = {
    "flux_top_out": lambda: flux_top_column(2.0e4),
    "flux_top_in": lambda: flux_top_column(-7.5e3),
    "flux_x-": lambda: flux_side_line("x-"),
    "flux_x+": lambda: flux_side_line("x+"),
    "flux_y-": lambda: flux_side_line("y-"),
    "flux_y+": lambda: flux_side_line("y+"),
    "transient_no_source": lambda: transient_uniform_source(0.0),
    "transient_source": lambda: transient_uniform_source(3.0e10),
    "isothermal_x_two_materials": lambda: isothermal_two_materials(0),
    "isothermal_y_two_materials": lambda: isothermal_two_materials(1),
}


def check(case, th, metrics, extra):
    """Assert the solution (or, for start_is_solution, the residual) of a
    case against its closed form."""
    dx, dy, ZC = metrics
    if case.start_is_solution:
        res, bnorm = extra
        assert bnorm > 0.0
        assert res <= 1e-10 * bnorm
        return
    exp = case.expected(dx, dy, ZC)
    err = float(np.max(np.abs(th - exp)))
    assert err <= case.tol, "max |θ - closed form| = %.3e K > %.3e" % (err, case.tol)
    if case.energy is not None:
        assert case.energy(th, dx, dy, ZC) <= 1e-7
