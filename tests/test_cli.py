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
    cp = api.check_cosmology(cosmo)
    grid = dict(nk_k=10, nk_g=12, k_k_max=1.0, k_g_max=2.0, n_kap=5, n_gam=6, kap_min=-0.1,
                kap_max=0.99, gam_max=0.8, xi_min=-0.2, xi_max=1.3, sigma=0.03)
    info = api.record_2d(cp, config, api.on(config), closure, edge,
                         dict(kbar=0.05, N_expl=1e6, N_fil=10.0, n_bins=100, n_points=1000),
                         dict(sigma=0.03, mean=0.0, var=0.01, skew=1.5, sigmaDL=0.03, grid=grid),
                         dict(build=0.2, lam=0.0, invert=0.05))
    info.update(zs=float(zs))
    x = np.asarray(xi_out)
    return np.exp(-0.5 * (x / 0.1) ** 2) / (0.1 * np.sqrt(2 * np.pi)), info


def test_record_has_the_reference_key_set(tmp_path):
    keys = set(json.load(open(GOLDEN))["fixture_keys"])
    out_file = tmp_path / "pdf.json"
    with patch("sgl_analytic_amd.cli.run", side_effect=fake_run):
        cli.main(["--zs", "1", "--growth-mode", "auto", "--radiation", "none", "--config", "+fil",
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
    # subhalos pair with the r200 edge only: a usage error, not a silent other model
    with pytest.raises(SystemExit):
        cli.main(["--zs", "1", "--edge", "rvir", "--out", str(tmp_path / "x.json")])
