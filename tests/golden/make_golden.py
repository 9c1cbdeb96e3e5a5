"""Writes the golden fixtures tests/golden/<deck>.json from the reference itself.

Runs oracle/_ref/ref_trace (the reference driver + serial model, built from the
reference's sources by ``make -C oracle ref``) or, for the large decks,
oracle/_ref/ref_trace_omp (the same driver with the reference omp model) on a
deck under tests/golden/decks/ and keeps, for every step, the reference's time,
dt and the five field_summary sums with 17 significant digits, and x_vol ..
x_press: the exact sums of the same per-cell terms of the reference's state
(double-double accumulation in the trace harness, oracle/ref_trace.cpp), which
do not depend on a summation order.

The reference prints a second, identical summary at the final step
(driver/hydro.cpp:93 and :111); the fixture keeps one record per step.

usage: python tests/golden/make_golden.py [deck ...]
"""
import json
import subprocess
import sys
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
SERIAL = ["clover_default", "clover_sodx", "clover_sody", "clover_sodxy", "small_bm_64x48", "small_geom_mix",
          "uniform_32x24", "clover_bm_short"]
OMP = ["clover_bm16_very_short", "clover_bm2_short", "clover_bm4_short", "clover_bm16_short", "clover_bm16_300",
       "clover_bm256_short"]


def trace(deck: str) -> dict:
    exe = REPO / "oracle" / "_ref" / ("ref_trace" if deck in SERIAL else "ref_trace_omp")
    with tempfile.TemporaryDirectory() as td:
        out = Path(td) / "t.json"
        subprocess.run([str(exe), str(HERE / "decks" / f"{deck}.in"), str(out)], cwd=td, check=True,
                       stdout=subprocess.DEVNULL)
        data = json.loads(out.read_text())
    recs, seen = [], set()
    for r in data["summaries"]:
        if r["step"] in seen:
            continue
        seen.add(r["step"])
        recs.append(r)
    return {"deck": f"{deck}.in", "model": "serial" if deck in SERIAL else "omp",
            "x_cells": data["x_cells"], "y_cells": data["y_cells"], "summaries": recs}


def main(decks):
    for d in decks or SERIAL + OMP:
        g = trace(d)
        (HERE / f"{d}.json").write_text(json.dumps(g, indent=0) + "\n")
        print(d, len(g["summaries"]), "records")


if __name__ == "__main__":
    main(sys.argv[1:])
