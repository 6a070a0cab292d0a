"""The command line on the device (needs a GPU): the record of a real solve
has the reference's key set and a unit-mass body."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_scalars.json")


def test_cli_record_on_device(tmp_path):
    from sgl_analytic_amd import cli
    print("cli record zs=1", flush=True)
    out_file = tmp_path / "pdf_zs1.json"
    cli.main(["--zs", "1", "--config", "+fil", "--out", str(out_file)])
    rec = json.load(open(out_file))
    keys = set(json.load(open(GOLDEN))["fixture_keys"])
    assert keys <= set(rec), sorted(keys - set(rec))
    x = np.array(rec["xi"] + rec["xi_tail"])
    p = np.array(rec["P_s"] + rec["P_s_tail"])
    print("grid mass", rec["diagnostics"]["grid_mass"], "clipped", rec["clipped"])
    assert abs(np.trapezoid(p, x) - 1.0) < 2e-3
    assert rec["n_bins"] > 1000 and rec["N_fil"] > 0 and rec["sigma_kappa"] > 0
    assert rec["xi_closure"] == "gauss" and rec["seconds"] > 0
