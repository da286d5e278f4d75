"""Command line front end (python -m gap_amd.cli)."""
import json

import numpy as np
import pytest

from gap_amd import capi, cli, gridgen


def test_cli_requires_exactly_one_input():
    with pytest.raises(SystemExit):
        cli.main([])
    with pytest.raises(SystemExit):
        cli.main(["in.json", "--synthetic", "200"])


@pytest.mark.gpu
def test_cli_synthetic_200_matches_capi(device, tmp_path):
    out = tmp_path / "r.json"
    assert cli.main(["--synthetic", "200", "-o", str(out), "--repeat", "2"]) == 0
    doc = json.loads(out.read_text())
    net = gridgen.small_mesh(200, seed=0, n_extra=40)
    ref = capi.Plan([net]).solve(1e-8, 100, 100e6)
    node = doc["networks"][0]["node"]
    assert doc["networks"][0]["iterations"] == int(ref["iterations"][0])
    u = np.array([n["u_pu"] for n in node])
    assert np.abs(u - np.abs(ref["v"][0])).max() < 1e-12
    plan = capi.Plan([net])
    plan.solve(1e-8, 100, 100e6)
    fl = plan.branch_flows()
    line = doc["networks"][0]["line"]
    assert [row["id"] for row in line] == [int(b) for b in net.br_id]
    for f in capi.FLOW_FIELDS:
        assert np.array_equal(np.array([row[f] for row in line]), fl[f])
