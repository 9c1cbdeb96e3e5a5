"""Writes tests/golden/batch_seed355.json: the reference's own results for the 64
problems of the seeded batch workload.

Each problem's deck comes from the library's generator (clv_batch_deck, seed 355);
oracle/_ref/ref_trace (the reference driver + serial model, built from the
reference's sources by ``make -C oracle ref``) runs it to end_step 100 and
records, with 17 significant digits, the dt of every step and the five
field_summary sums.  The fixture keeps per problem: the deck's sha256 (pins the
generator), the grid side, dt of steps 1..100, the final time, and the summaries
of steps 1, 50 and 100.

usage: python tests/golden/make_batch_golden.py [--jobs N]
"""
import argparse
import hashlib
import json
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
SEED = 355
N = 64
KEEP = (1, 50, 100)


def run_one(td: Path, i: int, text: str) -> dict:
    deck = td / f"b{i:02d}.in"
    out = td / f"b{i:02d}.json"
    deck.write_text(text)
    subprocess.run([str(REPO / "oracle" / "_ref" / "ref_trace"), str(deck), str(out)], cwd=td, check=True,
                   stdout=subprocess.DEVNULL)
    data = json.loads(out.read_text())
    by_step = {}
    for r in data["summaries"]:
        by_step.setdefault(r["step"], r)
    steps = sorted(s for s in by_step if s > 0)
    return {
        "index": i,
        "deck_sha256": hashlib.sha256(text.encode()).hexdigest(),
        "side": data["x_cells"],
        "dt": [by_step[s]["dt"] for s in steps],
        "time": by_step[steps[-1]]["time"],
        "summaries": [{k: by_step[s][k] for k in ("step", "time", "vol", "mass", "ie", "ke", "press")} for s in KEEP],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    a = ap.parse_args()
    sys.path.insert(0, str(REPO))
    import cloverleaf_amd as clv
    texts = [clv.batch_deck(SEED, i) for i in range(N)]
    with tempfile.TemporaryDirectory() as td:
        # largest first so the pool drains evenly
        order = sorted(range(N), key=lambda i: -clv.parse_deck(texts[i]).x_cells)
        with ThreadPoolExecutor(a.jobs) as ex:
            res = dict(zip(order, ex.map(lambda i: run_one(Path(td), i, texts[i]), order)))
    out = {"seed": SEED, "end_step": 100, "model": "serial", "problems": [res[i] for i in range(N)]}
    (HERE / f"batch_seed{SEED}.json").write_text(json.dumps(out, separators=(",", ":")) + "\n")
    print("wrote", HERE / f"batch_seed{SEED}.json")


if __name__ == "__main__":
    main()
