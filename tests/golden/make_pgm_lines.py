"""Writes tests/golden/pgm_lines.json: the `line` results (p, q, i, s at both
ends) that power-grid-model's own output fixtures hold for the PGM cases of
the golden set (reference tests/pgm_validation/test_data/<case>/output.json),
keyed by line id; and tests/golden/pgm_inputs/<case>.json, the PGM input
datasets of those cases (the same data, compact).  Data only; run in a
container that has the reference tree:

    python tests/golden/make_pgm_lines.py
"""
import json
from pathlib import Path

REF = Path("/root/reference/tests/pgm_validation/test_data")
OUT = Path(__file__).resolve().parent / "pgm_lines.json"
FIELDS = ("p_from", "q_from", "i_from", "s_from", "p_to", "q_to", "i_to", "s_to", "loading")


def main():
    from make_golden import PGM_CASES
    data = {}
    for case in PGM_CASES:
        out = json.loads((REF / case / "output.json").read_text())
        lines = out["line"] if "line" in out else out["data"]["line"]
        data[f"pgm_{case}"] = {"id": [int(r["id"]) for r in lines],
                               **{f: [float(r[f]) for r in lines] for f in FIELDS}}
    OUT.write_text(json.dumps(data, indent=0))
    (OUT.parent / "pgm_inputs").mkdir(exist_ok=True)
    for case in PGM_CASES:
        doc = json.loads((REF / case / "input.json").read_text())
        (OUT.parent / "pgm_inputs" / f"{case}.json").write_text(json.dumps(doc, separators=(",", ":")))
    print("wrote", OUT, sum(len(v["id"]) for v in data.values()), "lines")


if __name__ == "__main__":
    main()
