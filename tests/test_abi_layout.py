"""The plugin boundary's types laid out as the reference's: tests/abi_probe.cpp
compiled against the reference headers and against include/gap must print the
same sizes, alignments, field offsets, enumerator values and vtable slots.
Needs the reference tree (skipped where it is absent); builds only into the
test's temporary directory."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
REF_INC = Path("/root/reference/include")


@pytest.mark.skipif(not (REF_INC / "gap" / "core" / "types.h").exists(), reason="reference headers not present")
def test_boundary_layout_matches_reference_headers(tmp_path):
    outs = {}
    for tag, inc in (("reference", REF_INC), ("ours", ROOT / "include")):
        exe = tmp_path / f"probe_{tag}"
        subprocess.run(["g++", "-std=c++20", "-O0", "-Wno-invalid-offsetof", "-I", str(inc),
                        str(ROOT / "tests" / "abi_probe.cpp"), "-o", str(exe)], check=True)
        outs[tag] = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines()
    assert len(outs["reference"]) > 100
    diff = [(a, b) for a, b in zip(outs["reference"], outs["ours"]) if a != b]
    assert not diff and len(outs["reference"]) == len(outs["ours"]), diff
