"""Run outputs of the solve path in the reference's text formats, and the
z-grid file format.

* log.txt: the run header of `conditions` (src/heat3ds.jl:67-84), the
  residual lines "%10d %24.14E" and the per-step θ statistics line of
  `main` (src/heat3ds.jl:145-148).
* convergence CSV: `export_convergence_csv` (src/convergence_history.jl:119-144)
  from residual lines read back as `parse_residuals_from_log!` does
  (src/parse_log_residuals.jl:14-44); `get_convergence_info`
  (src/convergence_history.jl:234-248).
* Z-line CSV: `plot_line_z_nu`'s sample (src/plotter.jl:343-369, cell
  lookup `find_i`/`find_j`/`chk_idx` :3-21) written by `export_zline_csv`
  (src/plotter.jl:401-417).
* grid file: `read_grid_file` (src/Zcoord.jl:19-62), a count line, then
  "index coord" lines with 1-based indices; the z metrics of `genZ!`
  (src/Zcoord.jl:133-166) from any face list.

One intended difference: the reference prints its residual lines to
stdout only, so its log.txt holds none and its convergence CSV comes out
empty; here `write_log` puts them into log.txt, which makes the CSV hold
the history.
"""
import math
import re
import time
import warnings

import numpy as np

ITR_MAX_PRINTED = 8000   # `ItrMax` (src/common.jl:10), printed in the header; the solvers cap at 20000
_RES_LINE = re.compile(r"^\s+(\d+)\s+([0-9]+\.[0-9]+E[+-]\d+)")


def residual_lines(history):
    """The solvers' per-iteration lines, "%10d %24.14E"."""
    return ["%10d %24.14E" % (k + 1, r) for k, r in enumerate(history)]


def conditions(SZ, dh, solver, smoother, eps):
    """Header lines of log.txt (`conditions`, src/heat3ds.jl:67-84)."""
    out = ["Problem : IC on NonUniform grid (Opt. 13 layers)",
           "Grid  : %d %d %d" % tuple(SZ),
           "Pitch : %6.4e %6.4e %6.4e" % tuple(dh)]
    if solver in ("pbicgstab", "cg"):
        out.append("Solver: %s without preconditioner" % solver if not smoother
                   else "Solver: %s with smoother %s" % (solver, smoother))
    else:
        out.append("Solver: %s" % solver)
    out.append("ItrMax : %e" % ITR_MAX_PRINTED)
    out.append("ε      : %e" % eps)
    return out


def theta_stats_line(step, t, theta):
    """`main`'s per-step line over the interior (src/heat3ds.jl:145-148)."""
    s = theta[1:-1, 1:-1, 1:-1]
    return "%d %f : θmin=%e  θmax=%e  L2 norm of θ=%e" % (step, t, float(s.min()), float(s.max()),
                                                         float(np.linalg.norm(s.ravel())))


def write_log(filename, header, steps):
    """log.txt: header lines, then per step its residual lines and its
    θ statistics line; steps = [(history, stats_line), ...]."""
    with open(filename, "w") as f:
        for line in header:
            f.write(line + "\n")
        for history, stats in steps:
            for line in residual_lines(history):
                f.write(line + "\n")
            f.write(stats + "\n")


class ConvergenceData:
    """Residual history of one solve (`ConvergenceData`,
    src/convergence_history.jl:16-35)."""

    def __init__(self, solver, smoother=""):
        self.solver_name = solver
        self.smoother_name = smoother
        self.iterations = []
        self.residuals = []
        self.start_time = time.time()

    def add_residual(self, iteration, residual):
        self.iterations.append(int(iteration))
        self.residuals.append(float(residual))


def parse_residuals_from_log(conv, filename):
    """Residual lines of a log file into conv; True when any were found."""
    n = 0
    with open(filename) as f:
        for line in f:
            m = _RES_LINE.match(line)
            if m:
                conv.add_residual(int(m.group(1)), float(m.group(2)))
                n += 1
    return n > 0


def export_convergence_csv(conv, filename):
    """Writes the history as `export_convergence_csv` does; False (nothing
    written) for an empty history."""
    if not conv.residuals:
        return False
    with open(filename, "w") as f:
        f.write("# Convergence History Data\n")
        f.write("# Solver: %s\n" % conv.solver_name)
        if conv.smoother_name:
            f.write("# Smoother: %s\n" % conv.smoother_name)
        f.write("# Generated: %s\n" % time.strftime("%Y-%m-%d %H:%M:%S", time.gmtime(conv.start_time)))
        f.write("Iteration,Residual\n")
        for it, r in zip(conv.iterations, conv.residuals):
            f.write("%d,%.14E\n" % (it, r))
    return True


def get_convergence_info(conv):
    if not conv.residuals:
        return {}
    r0, r1 = conv.residuals[0], conv.residuals[-1]
    return {"solver": conv.solver_name, "smoother": conv.smoother_name, "iterations": len(conv.residuals),
            "initial_residual": r0, "final_residual": r1, "convergence_rate": r1 / r0,
            "reduction_factor": math.log10(r0 / r1)}


def _chk_idx(i, n):
    return min(max(i, 1), n)


def find_ij(x, x0, dx, n):
    """1-based cell index of coordinate x (`find_i`/`find_j`)."""
    return _chk_idx(int(math.floor((x - x0) / dx + 1.5)), n)


def line_z(theta, dh, Z, xc, yc, ox=(0.0, 0.0, 0.0)):
    """(z [mm], θ) along z at (xc, yc) as `plot_line_z_nu` samples it: the
    interior planes k = 2..MZ-1 (1-based) against Z[k]·1000."""
    MX, MY, MZ = theta.shape
    i = find_ij(xc, ox[0], dh[0], MX)
    j = find_ij(yc, ox[1], dh[1], MY)
    ks = range(2, MZ)
    return [Z[k - 1] * 1000.0 for k in ks], [float(theta[i - 1, j - 1, k - 1]) for k in ks]


def export_zline_csv(z, d, filename):
    """`export_zline_csv`: writes filename + ".csv"."""
    with open(filename + ".csv", "w") as f:
        f.write("Z [mm], Temperature [K]\n")
        for a, b in zip(z, d):
            f.write("%f, %f\n" % (a, b))


def read_grid_file(filename="grid.txt"):
    """(coord, numNodes) of a grid file; a line without exactly two fields
    is an error, an index other than its line number a warning."""
    with open(filename) as f:
        n = int(f.readline().strip())
        coord = np.zeros(n)
        for i in range(1, n + 1):
            line = f.readline()
            parts = line.split()
            if len(parts) != 2:
                raise ValueError("grid file format error (line %d): %r" % (i + 1, line))
            if int(parts[0]) != i:
                warnings.warn("grid index %s differs from the expected %d" % (parts[0], i))
            coord[i - 1] = float(parts[1])
    return coord, n


def write_grid_file(filename, coord):
    with open(filename, "w") as f:
        f.write("%d\n" % len(coord))
        for i, z in enumerate(coord, 1):
            f.write("%d %.17g\n" % (i, z))


def genZ_from_faces(z_face):
    """(Z, ZC, ΔZ) of `genZ!` for the nk+1 faces z_face: ΔZ with guard
    cells equal to their neighbours, ZC = face midpoints with the guard
    entries on the outer faces, Z the faces extended linearly by one on
    each side."""
    zf = np.asarray(z_face, dtype=np.float64)
    nk = len(zf) - 1
    dz = np.diff(zf)
    dZ = np.zeros(nk + 2)
    dZ[1:nk + 1] = dz
    dZ[0], dZ[nk + 1] = dZ[1], dZ[nk]
    ZC = np.zeros(nk + 2)
    ZC[0], ZC[nk + 1] = zf[0], zf[nk]
    ZC[1:nk + 1] = (zf[1:] + zf[:-1]) * 0.5
    Z = np.zeros(nk + 3)
    Z[1:nk + 2] = zf
    Z[0] = 2 * zf[0] - zf[1]
    Z[nk + 2] = 2 * zf[nk] - zf[nk - 1]
    return Z, ZC, dZ
