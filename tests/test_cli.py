"""Command line and output record (the reference's cli.py and run() record
schema, pipeline.py:243-274): every top-level key of the reference's stored
PDF records is present, the grids are the reference's, and the JSON is
strict.  The device solve is mocked here; tests/test_gpu_cli.py runs it."""
import json
import os
from unittest.mock import patch

import numpy as np

from sgl_analytic_amd import api, cli

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_scalars.json")


def fake_run(zs, config="full", cosmo=None, xi_out=None, closure="gauss", edge="r200"):
    from sgl_analytic_amd.solver import Shape
    cp = api.check_cosmology(cosmo)
    diag = np.array([1.0, 0.01, 0.02, 1.5, 0.07, 1.0, 0.01, 0.0])
    info = api.reference_schema(cp, Shape(), diag, config, closure, edge, 0.25, 0.05)
    info.update(zs=float(zs), cfg=config)
    x = np.asarray(xi_out)
    return np.exp(-0.5 * (x / 0.1) ** 2) / (0.1 * np.sqrt(2 * np.pi)), info


def test_record_has_the_reference_key_set(tmp_path):
    keys = set(json.load(open(GOLDEN))["fixture_keys"])
    out_file = tmp_path / "pdf.json"
    with patch("sgl_analytic_amd.cli.run", side_effect=fake_run):
        cli.main(["--zs", "1", "--growth-mode", "auto", "--radiation", "none",
                  "--out", str(out_file)])
    rec = json.load(open(out_file))
    assert keys <= set(rec), sorted(keys - set(rec))
    assert len(rec["xi"]) == len(rec["P_s"]) == 1601 and rec["xi"][0] == -1.0
    assert len(rec["xi_tail"]) == len(rec["P_s_tail"]) == 1600 and rec["xi_tail"][0] == 1.005
    assert rec["dark_energy"]["growth_mode"] == "legacy" and rec["dark_energy"]["radiation"] is None
    assert rec["dark_energy"]["Omega_r"] == 0.0
    assert abs(rec["clipped"]["mean"]) < 1e-12 and abs(rec["diagnostics"]["grid_mass"] - 1) < 1e-9
    assert rec["moment_interval"] == [rec["grid"]["xi_min"], rec["grid"]["xi_max"]]


def test_cli_rejects_bad_input(tmp_path):
    import pytest
    with patch("sgl_analytic_amd.cli.run", side_effect=fake_run):
        with pytest.raises(SystemExit):
            cli.main(["--zs", "1", "--Om", "-1", "--out", str(tmp_path / "x.json")])
        with pytest.raises(SystemExit):
            cli.main(["--zs", "1", "--nproc", "0", "--out", str(tmp_path / "x.json")])
