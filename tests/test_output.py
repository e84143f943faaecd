"""Run outputs and the z-grid file (no GPU): residual-line format against
the README's log lines, log parse round trip, convergence CSV layout,
Z-line sampling and CSV, grid-file round trip and errors, and the genZ!
metrics from a face list against the model builder's Zcase2 metrics."""
import warnings

import numpy as np
import pytest

import h2_amd
from h2_amd import output as O


def test_residual_lines_match_readme_format():
    # README.md:175-178, the 240x240x30 run's iterations 1, 2 and 208
    lines = O.residual_lines([6.58308938608891E-01, 8.89700596995947E-01])
    assert lines == ["         1     6.58308938608891E-01", "         2     8.89700596995947E-01"]
    assert O.residual_lines([0.0] * 207 + [9.72579214884977E-05])[-1] == "       208     9.72579214884977E-05"


def test_log_round_trip_and_convergence_csv(tmp_path):
    hist = [0.65, 0.12, 3.5e-3, 9.9e-5]
    th = np.full((5, 4, 6), 300.0, order="F")
    th[1:-1, 1:-1, 1:-1] = np.arange(2 * 3 * 4).reshape(3, 2, 4, order="F").transpose(0, 1, 2).reshape(
        3, 2, 4) + 300.5
    log = tmp_path / "log.txt"
    header = O.conditions((5, 4, 6), (1e-5, 1e-5, 1.0), "pbicgstab", "gs", 1e-4)
    assert header[1] == "Grid  : 5 4 6" and header[3] == "Solver: pbicgstab with smoother gs"
    assert header[4] == "ItrMax : 8.000000e+03" and header[5] == "ε      : 1.000000e-04"
    stats = O.theta_stats_line(1, 10000.0, th)
    assert stats.startswith("1 10000.000000 : θmin=3.005000e+02  θmax=3.235000e+02")
    O.write_log(str(log), header, [(hist, stats)])
    conv = O.ConvergenceData("pbicgstab", "gs")
    assert O.parse_residuals_from_log(conv, str(log))
    assert conv.iterations == [1, 2, 3, 4]
    assert np.allclose(conv.residuals, hist, rtol=1e-14, atol=0)
    csv = tmp_path / "c.csv"
    assert O.export_convergence_csv(conv, str(csv))
    text = csv.read_text().splitlines()
    assert text[0] == "# Convergence History Data" and text[1] == "# Solver: pbicgstab"
    assert text[2] == "# Smoother: gs" and text[4] == "Iteration,Residual"
    assert text[5] == "1,6.50000000000000E-01" and len(text) == 9
    info = O.get_convergence_info(conv)
    assert info["iterations"] == 4 and abs(info["reduction_factor"] - np.log10(0.65 / 9.9e-5)) < 1e-12
    assert not O.export_convergence_csv(O.ConvergenceData("cg"), str(tmp_path / "empty.csv"))


def test_zline_sampling_and_csv(tmp_path):
    m = h2_amd.build_model(40, 40, 30)
    th = np.asfortranarray(np.random.default_rng(0).normal(300, 5, m.shape))
    Z, ZC, dZ = h2_amd.genZ(30)
    z, d = O.line_z(th, (m.dx, m.dy, 1.0), Z, 0.6e-3, 0.6e-3)
    i = int(np.floor(0.6e-3 / m.dx + 1.5))
    assert d == [float(th[i - 1, i - 1, k - 1]) for k in range(2, 32)]
    assert z == [Z[k - 1] * 1000 for k in range(2, 32)]
    # outside the domain the index is clamped (chk_idx)
    assert O.find_ij(-1.0, 0.0, m.dx, 42) == 1 and O.find_ij(1.0, 0.0, m.dx, 42) == 42
    O.export_zline_csv(z, d, str(tmp_path / "temp3Z_ctr"))
    lines = (tmp_path / "temp3Z_ctr.csv").read_text().splitlines()
    assert lines[0] == "Z [mm], Temperature [K]" and len(lines) == 31
    assert lines[1] == "%f, %f" % (z[0], d[0])


def test_grid_file_round_trip_and_errors(tmp_path):
    Z, ZC, dZ = h2_amd.genZ(30)
    faces = Z[1:32]
    p = tmp_path / "grid.txt"
    O.write_grid_file(str(p), faces)
    coord, n = O.read_grid_file(str(p))
    assert n == 31 and np.array_equal(coord, faces)
    bad = tmp_path / "bad.txt"
    bad.write_text("2\n1 0.0\n2 1.0 extra\n")
    with pytest.raises(ValueError):
        O.read_grid_file(str(bad))
    odd = tmp_path / "odd.txt"
    odd.write_text("2\n1 0.0\n5 1.0e-3\n")
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        coord, n = O.read_grid_file(str(odd))
    assert n == 2 and coord[1] == 1.0e-3 and len(w) == 1


@pytest.mark.parametrize("nk", [30, 60, 120])
def test_genz_from_faces_matches_model_builder(nk):
    Z, ZC, dZ = h2_amd.genZ(nk)
    Z2, ZC2, dZ2 = O.genZ_from_faces(Z[1:nk + 2])
    assert np.array_equal(Z, Z2) and np.array_equal(ZC, ZC2) and np.array_equal(dZ, dZ2)
