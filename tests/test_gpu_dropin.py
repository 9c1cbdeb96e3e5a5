"""The drop-in: the reference's own, unchanged driver (driver/*.cpp) linked with the
MI355X plugin model (cloverleaf_amd/model_mi355x) prints what the reference
prints.

The binary is built by cloverleaf_amd/model_mi355x/Makefile from the reference
sources where the reference tree is present (this file skips where it was not
built).  tests/golden/clover_bm_short.reference_lines.txt holds the reference's
own printed dt lines and field-summary table (the reference omp binary built
here; the serial binary prints identical lines).
"""
import os
import re
import subprocess
from pathlib import Path

import pytest

import clover_testlib as T

REPO = Path(__file__).resolve().parent.parent
DROPIN = REPO / "cloverleaf_amd" / "model_mi355x" / "_dropin" / "mi355x-cloverleaf"

pytestmark = pytest.mark.skipif(not DROPIN.exists(), reason="drop-in binary not built (needs the reference tree)")


def _printed_lines(stdout: str, clover_out: str):
    steps = [l for l in stdout.splitlines() if re.match(r"^ Step \d+ time", l)]
    table = [l for l in clover_out.splitlines() if l.startswith(" step:")]
    return steps + table


def _run(tmp_path, deck, env_extra=None, args=()):
    env = dict(os.environ)
    env.update(env_extra or {})
    return subprocess.run([str(DROPIN), "--file", str(T.DECKS / f"{deck}.in"), "--out", str(tmp_path / "clover.out"), *args],
                          cwd=tmp_path, env=env, capture_output=True, text=True, timeout=600)


@pytest.mark.gpu
@pytest.mark.parametrize("order,chains", [("tree", "chains"), ("serial", "chains"), ("tree", "off")])
def test_dropin_prints_the_reference_output(gpu, tmp_path, order, chains):
    """clover_bm_short (960^2, 87 steps, test problem 2): every dt line and every
    field-summary line equal to the reference's, and the QA check passes; with the
    chains of entry points as one launch each (the default) and one entry point at a time."""
    env = {"CLOVER_MI355X_SUMMARY": order}
    if chains == "off":
        env["CLOVER_MI355X_CHAINS"] = "off"
    p = _run(tmp_path, "clover_bm_short", env)
    assert p.returncode == 0, p.stderr[-2000:]
    got = _printed_lines(p.stdout, (tmp_path / "clover.out").read_text())
    want = (T.GOLDEN / "clover_bm_short.reference_lines.txt").read_text().splitlines()
    assert got == want
    assert "This test is considered PASSED" in p.stdout
    assert "MI355X" in p.stdout


@pytest.mark.gpu
def test_dropin_profiler_and_known_answer(gpu, tmp_path):
    """clover_bm16_very_short (3840^2, 8 steps, test problem 168) with --profile: the
    reference's per-phase profile table is produced and the QA check passes."""
    p = _run(tmp_path, "clover_bm16_very_short", args=("--profile",))
    assert p.returncode == 0, p.stderr[-2000:]
    assert "This test is considered PASSED" in p.stdout
    out = (tmp_path / "clover.out").read_text()
    assert "Profiler Output" in out and "Cell Advection" in out  # driver/hydro.cpp:177-194


def test_dropin_without_gpu_fails_cleanly(tmp_path):
    """Without a HIP device the model says so and exits non-zero (no CPU fallback)."""
    import cloverleaf_amd as clv
    if clv.lib().clv_device_count() > 0:
        pytest.skip("a GPU is visible")
    p = _run(tmp_path, "clover_default")
    assert p.returncode != 0
    assert "no HIP device" in p.stderr


@pytest.mark.gpu
def test_dropin_dumps_equal_reference(gpu, tmp_path):
    """--dump (driver/shared.cpp:96-189): the field dumps the reference's driver writes
    after every start-up and step phase, with this model, byte-identical to the
    reference serial model's own (tests/golden/driver_output.json.gz, case tiny: 6 x 5
    cells, 2 steps).  Every array of the chunk is dumped, work arrays included."""
    import gzip
    import json
    g = json.load(gzip.open(T.GOLDEN / "driver_output.json.gz", "rt"))["tiny"]
    (tmp_path / "deck.in").write_text((T.DECKS / g["deck"]).read_text())
    p = subprocess.run([str(DROPIN), *g["args"]], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    files = sorted(f.name for f in (tmp_path / "d").iterdir())
    assert files == sorted(g["dumps"])
    for name in files:
        got = (tmp_path / "d" / name).read_text()
        if got != g["dumps"][name]:
            a, b = got.splitlines(), g["dumps"][name].splitlines()
            bad = next((k for k in range(min(len(a), len(b))) if a[k] != b[k]), min(len(a), len(b)))
            raise AssertionError(f"{name}: line {bad}: {a[bad] if bad < len(a) else None!r} != {b[bad] if bad < len(b) else None!r}")



@pytest.mark.gpu
def test_dropin_visit_files_equal_reference(gpu, tmp_path):
    """visit_frequency: the reference driver's visit output (driver/visit.cpp:38-196)
    with this model -- clover.visit and every VTK grid -- byte-identical to the
    reference serial model's (tests/golden/driver_output.json.gz, case visit)."""
    import gzip
    import json
    g = json.load(gzip.open(T.GOLDEN / "driver_output.json.gz", "rt"))["visit"]
    (tmp_path / "deck.in").write_text((T.DECKS / g["deck"]).read_text())
    p = subprocess.run([str(DROPIN), *g["args"]], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    files = sorted(f.name for f in tmp_path.iterdir() if f.name == "clover.visit" or f.suffix == ".vtk")
    assert files == sorted(g["visit"])
    for name in files:
        got, want = (tmp_path / name).read_text().splitlines(), g["visit"][name].splitlines()
        bad = [(k, a, b) for k, (a, b) in enumerate(zip(got, want)) if a != b]
        assert len(got) == len(want) and not bad, (name, bad[:3])


def _printed_value_ok(printed: float, exact: float) -> bool:
    """printed is exact rounded to the 7 significant digits of the reference's format
    (std::scientific, driver/report.cpp), a last-digit flip allowed only within 1e-11
    of a rounding boundary."""
    if exact == 0.0:
        return printed == 0.0
    import math
    half = 0.5 * 10.0 ** (math.floor(math.log10(abs(exact))) - 6)
    return abs(printed - exact) <= half * (1 + 1e-9) + 1e-11 * abs(exact)


@pytest.mark.gpu
@pytest.mark.parametrize("deck", ["clover_bm16_300", "clover_bm256_short"])
def test_dropin_large_decks_vs_reference_trace(gpu, tmp_path, deck):
    """The drop-in on InputDecks/clover_bm16_300.in (3840^2, 300 steps) and the headline
    deck clover_bm256_short.in (15360^2, 87 steps) against the reference omp model's
    full-precision trace of the same deck (tests/golden/<deck>.json): every printed dt
    and time equal to the reference's value in the driver's own format
    (driver/timestep.cpp:98-103), and every printed field summary
    (driver/report.cpp:48-71) equal to the exact sums of the reference's state rounded
    to the printed digits."""
    gold = T.golden(deck)["summaries"]
    by_step = {g["step"]: g for g in gold}
    (tmp_path / "deck.in").write_text((T.DECKS / f"{deck}.in").read_text())
    print(f"# drop-in {deck}", flush=True)
    p = subprocess.run([str(DROPIN), "--file", "deck.in", "--out", "clover.out"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=600)
    assert p.returncode == 0, p.stderr[-2000:]
    steps = [l for l in p.stdout.splitlines() if re.match(r"^ Step \d+ time", l)]
    assert len(steps) == len(gold) - 1
    for line in steps:
        m = re.match(r"^ Step (\d+) time (\S+) control \S+ timestep  (\S+) ", line)
        assert m, line
        n = int(m.group(1))
        assert m.group(3) == f"{by_step[n]['dt']:g}", (line, by_step[n]["dt"])
        assert m.group(2) == f"{by_step[n - 1]['time']:g}", (line, by_step[n - 1]["time"])
    rows = [l.split() for l in (tmp_path / "clover.out").read_text().splitlines() if l.startswith(" step:")]
    assert rows and rows[-1][1] == str(len(gold) - 1)
    bad = []
    for r in rows:
        g = by_step[int(r[1])]
        vol, mass, ie, ke, press = g["x_vol"], g["x_mass"], g["x_ie"], g["x_ke"], g["x_press"]
        want = [vol, mass, mass / vol, press / vol, ie, ke, ie + ke]
        got = [float(x) for x in r[2:9]]
        bad += [(r[1], k, a, b) for k, (a, b) in enumerate(zip(got, want)) if not _printed_value_ok(a, b)]
    assert not bad, bad[:5]
