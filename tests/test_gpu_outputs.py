"""q3d's text outputs on the GPU path: log.txt with the header, one
residual line per iteration and the θ statistics line; the convergence CSV
with the same history; the two Z-line CSVs."""
import os

import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="no HIP device")]

import h2_amd
from h2_amd import output as O


def test_q3d_writes_reference_outputs(tmp_path):
    res = h2_amd.q3d(40, 40, 30, "pbicgstab", "gs", epsilon=1e-6, is_steady=True, verbose=False,
                     output_dir=str(tmp_path))
    log = (tmp_path / "log.txt").read_text().splitlines()
    assert log[0].startswith("Problem : IC on NonUniform grid")
    assert log[1] == "Grid  : 42 42 32"
    res_lines = [l for l in log if O._RES_LINE.match(l)]
    assert len(res_lines) == res.itr
    assert log[-1].startswith("1 10000.000000 : θmin=")
    csv = (tmp_path / "convergence_pbicgstab_40x40x30_gs.csv").read_text().splitlines()
    rows = [l for l in csv if not l.startswith("#")][1:]
    assert len(rows) == res.itr
    assert float(rows[-1].split(",")[1]) == pytest.approx(res.history[-1], rel=1e-13)
    for tag in ("ctr", "tsv"):
        lines = (tmp_path / ("temp3Z_%s.csv" % tag)).read_text().splitlines()
        assert len(lines) == 31
        t = np.array([float(l.split(",")[1]) for l in lines[1:]])
        assert np.all((t > 299.0) & (t < 400.0))
