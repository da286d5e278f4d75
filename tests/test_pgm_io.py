"""The C++ PGM JSON reader / writer and the gap_hip command line
(gap_amd/csrc/pgm_io.cpp, gap_hip_main.cpp).

CPU: the C++ conversion gives rows bit-identical to the Python restatement of
the reference harness (gap_amd/pgm.py) on every PGM input fixture; malformed
input, missing components and unknown backends fail with status 2.
GPU: end-to-end runs on the PGM fixtures against power-grid-model's own node
and line results and against the Python API on the same rows."""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import load_golden
from gap_amd import capi
from gap_amd.network import Network
from gap_amd.pgm import read_pgm_input

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "gap_amd" / "lib" / "gap_hip"
INPUTS = sorted((ROOT / "tests" / "golden" / "pgm_inputs").glob("*.json"))
PGM_LINES = json.loads((ROOT / "tests" / "golden" / "pgm_lines.json").read_text())


def run(*args, check=False):
    if not EXE.exists():
        pytest.skip("gap_hip not built (python -m gap_amd.build)")
    return subprocess.run([str(EXE), *map(str, args)], capture_output=True, text=True, timeout=120, check=check)


@pytest.mark.parametrize("path", INPUTS, ids=[p.stem for p in INPUTS])
def test_cpp_rows_bit_identical_to_harness_conversion(path):
    r = run("--dump-rows", "-i", path)
    assert r.returncode == 0, r.stderr
    rows = json.loads(r.stdout)
    bus, br, _ = read_pgm_input(path, 1e6)
    assert rows["bus"] == [[float(x) for x in row] for row in bus]
    assert rows["branch"] == [[float(x) for x in row[:6]] for row in br]


def test_base_power_option_changes_conversion():
    path = INPUTS[0]
    r = run("--dump-rows", "--base-power", "100e6", "-i", path)
    bus, br, _ = read_pgm_input(path, 100e6)
    rows = json.loads(r.stdout)
    assert rows["branch"] == [[float(x) for x in row[:6]] for row in br]


@pytest.mark.parametrize("text,msg", [
    ("{", "JSON"),
    ('{"data": {"node": []}}', "No nodes found"),
    ('{"data": {"node": [{"id": 1, "u_rated": 10}], "line": [{"id": 2, "from_node": 1, "to_node": 1, "r1": 1, "x1": 1}]}}',
     "No sources found"),
    ('{"node": [{"id": 1, "u_rated": 10}], "line": [{"id": 2, "from_node": 1, "to_node": 7, "r1": 1, "x1": 1}],'
     ' "source": [{"id": 3, "node": 1, "u_ref": 1.0}]}', "unknown node 7"),
    ('{"node": [1, 2] ] }', "JSON"),
    ('{"node": "x"}', "not an array"),
])
def test_bad_input_fails_with_status_2(tmp_path, text, msg):
    f = tmp_path / "bad.json"
    f.write_text(text)
    r = run("--dump-rows", "-i", f)
    assert r.returncode == 2
    assert msg in r.stderr


def test_json_escapes_exponents_and_extra_members(tmp_path):
    doc = {"version": "1.0", "type": "input", "note": "café \"q\" \\ 😀 \n",
           "data": {"node": [{"id": 2, "u_rated": 1.05e4, "name": "büs"}, {"id": 1, "u_rated": 10500.0}],
                    "line": [{"id": 5, "from_node": 1, "to_node": 2, "r1": 2.5E-1, "x1": 0.4, "c1": 1e-7,
                              "extra": [True, False, None, {"k": [1, 2, 3]}]}],
                    "source": [{"id": 9, "node": 1, "u_ref": 1.02}],
                    "sym_load": [{"id": 7, "node": 2, "status": 1, "p_specified": -1.5e5, "q_specified": 2e4}]}}
    f = tmp_path / "x.json"
    f.write_text(json.dumps(doc, ensure_ascii=True))
    r = run("--dump-rows", "-i", f)
    assert r.returncode == 0, r.stderr
    bus, br, _ = read_pgm_input(f, 1e6)
    rows = json.loads(r.stdout)
    assert rows["bus"] == [[float(x) for x in row] for row in bus]
    assert rows["branch"] == [[float(x) for x in row[:6]] for row in br]


def test_unknown_backend_and_usage():
    assert run("-b", "cpu", "-i", INPUTS[0]).returncode == 2
    assert run().returncode == 2
    assert run("-h").returncode == 0


@pytest.mark.gpu
@pytest.mark.parametrize("path", INPUTS, ids=[p.stem for p in INPUTS])
def test_gap_hip_end_to_end_on_pgm_fixture(device, tmp_path, path):
    case = "pgm_" + path.stem
    out = tmp_path / "out.json"
    r = run("-i", path, "-o", out)
    assert r.returncode == 0, r.stderr
    summary = json.loads(r.stdout)
    net, g = load_golden(case)
    assert summary["iterations"] == int(g["iterations"])
    doc = json.loads(out.read_text())
    nodes = sorted(doc["data"]["node"], key=lambda n: n["id"])
    u = np.array([n["u_pu"] for n in nodes])
    ang = np.array([n["u_angle"] for n in nodes])
    assert np.abs(u - g["pgm_u_pu"]).max() < 5e-6          # the reference's bar against PGM
    assert np.abs(ang - np.deg2rad(g["pgm_u_angle"])).max() < 5e-6
    # same rows through the Python API: the same solution
    bus, br, _ = read_pgm_input(path, 1e6)
    plan = capi.Plan([Network.from_rows(bus, br)])
    v = plan.solve(1e-8, 50, 1e6)["v"][0]
    assert np.abs(u - np.abs(v)).max() < 1e-12
    fl = plan.branch_flows()
    lines = {row["id"]: row for row in doc["data"]["line"]}
    for k, bid in enumerate(net.br_id):
        assert lines[int(bid)]["p_from"] == pytest.approx(fl["p_from"][k], rel=1e-12, abs=1e-9)
    ref = PGM_LINES[case]
    scale = max(abs(x) for x in ref["p_from"])
    for j, lid in enumerate(ref["id"]):
        for f in ("p_from", "q_from", "p_to", "q_to", "s_from", "i_from"):
            assert abs(lines[lid][f] - ref[f][j]) < 1e-6 * max(scale, abs(ref[f][j])), (lid, f)
    # each node's injection is what leaves it through its lines
    inj = {int(b): [0.0, 0.0] for b in net.bus_id}
    for k, bid in enumerate(net.br_id):
        row = lines[int(bid)]
        inj[int(net.br_from[k])][0] += row["p_from"]
        inj[int(net.br_to[k])][0] += row["p_to"]
        inj[int(net.br_from[k])][1] += row["q_from"]
        inj[int(net.br_to[k])][1] += row["q_to"]
    for n in nodes:
        assert n["p"] == pytest.approx(inj[n["id"]][0], rel=1e-12, abs=1e-6)
        assert n["q"] == pytest.approx(inj[n["id"]][1], rel=1e-12, abs=1e-6)
