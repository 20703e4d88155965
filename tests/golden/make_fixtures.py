"""Writes the compact golden fixtures under tests/golden/ from the reference's Arb data.

The reference commits arbitrary-precision (Arb, python-flint 0.9.0) ground truth as JSON lines
under ``reference/data/*.jsonl`` (format documented in ``reference/README.md:1-16``: each value
as hi + lo + lo2 with an error ball). This script keeps the inputs and the double-precision
``hi`` part of every value the path's tests need, and drops the rest (double-double digits,
error balls, the decimal strings), so the fixtures stay small. It reads the reference's data
only; no reference code is imported or run.

Usage: python tests/golden/make_fixtures.py [/path/to/reference]
"""

from __future__ import annotations

import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent


def _load(root: Path, name: str) -> list[dict]:
    with open(root / "reference" / "data" / name) as f:
        return [json.loads(line) for line in f if line.strip()]


def _hi(outputs: dict) -> dict:
    return {k: v["hi"] for k, v in outputs.items()}


def _write(name: str, records: list[dict]) -> None:
    path = HERE / name
    with open(path, "w") as f:
        for r in records:
            f.write(json.dumps(r, separators=(",", ":")) + "\n")
    print(f"{path.name}: {len(records)} records, {path.stat().st_size} bytes")


def main() -> None:
    root = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    for name in ("mie.jsonl", "mie_derivatives.jsonl", "layered.jsonl",
                 "layered_derivatives.jsonl", "mie_greek.jsonl"):
        recs = [
            {"case": r["case"], "label": r["label"], "inputs": r["inputs"],
             "outputs": _hi(r["outputs"])}
            for r in _load(root, name)
        ]
        _write(name, recs)
    # Gauss-Legendre rules up to 1000 points (the sizes the path uses).
    gl = [
        {"case": r["case"], "inputs": r["inputs"], "outputs": _hi(r["outputs"])}
        for r in _load(root, "gauss_legendre.jsonl")
        if r["inputs"]["n"] <= 1000
    ]
    _write("gauss_legendre.jsonl", gl)
    # Wigner d for the four pairs the Greek projection uses, degrees up to 4000.
    pairs = {(0, 0), (2, 2), (2, -2), (0, 2)}
    wd = [
        {"case": r["case"], "inputs": r["inputs"], "outputs": _hi(r["outputs"])}
        for r in _load(root, "wigner_d.jsonl")
        if (r["inputs"]["m"], r["inputs"]["n"]) in pairs and r["inputs"]["l_max"] <= 4000
    ]
    _write("wigner_d.jsonl", wd)


if __name__ == "__main__":
    main()
