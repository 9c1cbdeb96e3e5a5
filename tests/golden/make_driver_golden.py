"""Writes tests/golden/driver_output.json.gz: what the reference's own binary prints.

Runs oracle/_ref/serial-cloverleaf (the reference driver and serial model, built
from the reference's sources by ``make -C oracle ref``) and keeps its stdout, its
clover.out and, with --dump, every field dump file it writes
(driver/shared.cpp:96-189), and the visit files (driver/visit.cpp:38-196: clover.visit
and the .vtk grids), for:
  tiny   decks/tiny_dump_6x5.in with --dump d (6 x 5 cells, 2 steps)
  qa     decks/clover_default.in with --profile (test problem 1, 10 x 2 cells)
  visit  decks/visit_12x10.in (12 x 10 cells, 5 steps, visit_frequency 2)
The host driver's output is compared with these (tests/test_gpu_driver.py).

usage: python tests/golden/make_driver_golden.py
"""
import gzip
import json
import subprocess
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
EXE = REPO / "oracle" / "_ref" / "serial-cloverleaf"
RUNS = {"tiny": ("tiny_dump_6x5", ["--dump", "d"]), "qa": ("clover_default", ["--profile"]), "visit": ("visit_12x10", [])}


def run(deck, extra):
    with tempfile.TemporaryDirectory() as td:
        (Path(td) / "deck.in").write_text((HERE / "decks" / f"{deck}.in").read_text())
        args = ["--in", "deck.in", "--out", "clover.out", *extra]
        p = subprocess.run([str(EXE), *args], cwd=td, capture_output=True, text=True, check=True)
        dumps = {}
        if (Path(td) / "d").is_dir():
            dumps = {f.name: f.read_text() for f in sorted((Path(td) / "d").iterdir())}
        visit = {f.name: f.read_text() for f in sorted(Path(td).iterdir()) if f.name == "clover.visit" or f.suffix == ".vtk"}
        return {"deck": f"{deck}.in", "args": args, "stdout": p.stdout, "clover_out": (Path(td) / "clover.out").read_text(),
                "dumps": dumps, "visit": visit}


def main():
    out = {name: run(deck, extra) for name, (deck, extra) in RUNS.items()}
    with gzip.open(HERE / "driver_output.json.gz", "wt") as f:
        json.dump(out, f)
    for k, v in out.items():
        print(k, len(v["stdout"].splitlines()), "stdout lines,", len(v["dumps"]), "dumps,", len(v["visit"]), "visit files")


if __name__ == "__main__":
    main()
