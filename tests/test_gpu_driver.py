"""The host driver (cloverleaf_amd/bin/clover_mi355x) prints what the reference's
driver prints.

tests/golden/driver_output.json.gz holds the reference binary's own stdout,
clover.out and field dumps (written by tests/golden/make_driver_golden.py from
oracle/_ref/serial-cloverleaf, the reference built from its sources).  The host
driver runs the same decks with the same arguments; after masking what differs by
nature -- the timings (wall clock, grind times, profiler times), the device list
and the Model block (model-specific by the reference's design), the QA percentage
(a function of the summation order) -- every line is equal, and every field dump
(driver/shared.cpp:96-189, after every start-up and step phase) is byte-identical.
"""
import gzip
import json
import re
import subprocess

import pytest

import clover_testlib as T
import cloverleaf_amd as clv

pytestmark = pytest.mark.gpu

EXE = clv.PKG / "bin" / "clover_mi355x"
GOLD = json.load(gzip.open(T.GOLDEN / "driver_output.json.gz", "rt"))
TIMED = re.compile(r"^(\s*(?:Wall clock|Average time per cell|Step time per cell   |First step overhead))\s+\S+$")
PROFILE_ROW = re.compile(r"^( [A-Za-z ]+:)\S+ \S+$")


def _mask(line: str) -> str:
    m = TIMED.match(line)
    if m:
        return m.group(1) + " T"
    m = PROFILE_ROW.match(line)
    if m:
        return m.group(1) + "T P"
    return re.sub(r"(Test problem \d+ is within )\S+(% of the expected solution)", r"\1Q\2", line)


def normalise_stdout(text: str):
    out, in_devices, in_kernels = [], False, False
    for line in text.splitlines():
        if line == "Devices:":
            in_devices = True
            out.append(line)
            continue
        if in_devices and re.match(r"^ \d+: ", line):
            continue
        in_devices = False
        if line.startswith(" Kernel profile (HIP events"):
            in_kernels = True
            continue
        if in_kernels and line.startswith("  "):
            continue
        in_kernels = False
        if re.match(r"^ - (Name|Execution|Device):", line):
            continue
        out.append(_mask(line))
    return out


def normalise_out(text: str):
    return [_mask(l) for l in text.splitlines()]


def _run(tmp_path, case, extra=()):
    g = GOLD[case]
    (tmp_path / "deck.in").write_text((T.DECKS / g["deck"]).read_text())
    p = subprocess.run([str(EXE), *g["args"], *extra], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    return g, p


@pytest.mark.parametrize("case", ["tiny", "qa"])
def test_driver_output_equals_reference(gpu, tmp_path, case):
    g, p = _run(tmp_path, case)
    got, want = normalise_stdout(p.stdout), normalise_stdout(g["stdout"])
    assert got == want, "\n".join(f"{a!r} | {b!r}" for a, b in zip(got, want) if a != b)[:3000]
    got, want = normalise_out((tmp_path / "clover.out").read_text()), normalise_out(g["clover_out"])
    assert got == want, "\n".join(f"{a!r} | {b!r}" for a, b in zip(got, want) if a != b)[:3000]


def test_driver_dumps_equal_reference(gpu, tmp_path):
    """--dump: the same files with the same text as the reference's dumps."""
    g, _ = _run(tmp_path, "tiny")
    files = sorted(f.name for f in (tmp_path / "d").iterdir())
    assert files == sorted(g["dumps"])
    for name in files:
        got = (tmp_path / "d" / name).read_text()
        if got != g["dumps"][name]:
            a, b = got.splitlines(), g["dumps"][name].splitlines()
            bad = next(k for k in range(min(len(a), len(b))) if a[k] != b[k]) if a != b else None
            raise AssertionError(f"{name}: line {bad}: {a[bad] if bad is not None else ''!r} != {b[bad] if bad is not None else ''!r}")


@pytest.mark.parametrize("plan", ["fused", "reference"])
@pytest.mark.parametrize("geometry", ["arrays", "mesh"])
def test_driver_plans_print_the_same(gpu, tmp_path, plan, geometry):
    """Both plans and both geometry modes print the reference's lines on the 64x48
    deck (the fused plan runs there; masked as above)."""
    deck = "small_bm_64x48.in"
    (tmp_path / "deck.in").write_text((T.DECKS / deck).read_text())
    ref = subprocess.run([str(EXE), "--in", "deck.in", "--plan", "reference"], cwd=tmp_path, capture_output=True, text=True,
                         timeout=300)
    out_ref = (tmp_path / "clover.out").read_text()
    p = subprocess.run([str(EXE), "--in", "deck.in", "--plan", plan, "--geometry", geometry], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    assert ref.returncode == 0 and p.returncode == 0, p.stderr[-2000:]
    assert normalise_out((tmp_path / "clover.out").read_text()) == normalise_out(out_ref)
    strip = lambda t: [l for l in normalise_stdout(t) if not l.startswith(" Args:")]
    assert strip(p.stdout) == strip(ref.stdout)


def test_driver_small_deck_defaults_to_a_plan_that_runs(gpu, tmp_path):
    """Test problem 1 (10 x 2 cells) with the default plan: the fused plan needs 4 cells
    a side, so the driver runs the reference plan and the QA check passes."""
    (tmp_path / "deck.in").write_text((T.DECKS / "clover_default.in").read_text())
    p = subprocess.run([str(EXE), "--in", "deck.in"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    assert "This test is considered PASSED" in p.stdout
    assert " - Outcome: PASSED" in p.stdout


@pytest.mark.parametrize("plan", ["fused", "reference"])
def test_driver_visit_files_equal_reference(gpu, tmp_path, plan):
    """visit_frequency (driver/visit.cpp:38-196): clover.visit and every VTK grid, written
    after start-up, every second step and at the end, byte-identical to the
    reference's (case visit: 12 x 10 cells, 5 steps)."""
    g, _ = _run(tmp_path, "visit", ("--plan", plan))
    files = sorted(f.name for f in tmp_path.iterdir() if f.name == "clover.visit" or f.suffix == ".vtk")
    assert files == sorted(g["visit"])
    for name in files:
        got, want = (tmp_path / name).read_text().splitlines(), g["visit"][name].splitlines()
        bad = [(k, a, b) for k, (a, b) in enumerate(zip(got, want)) if a != b]
        assert len(got) == len(want) and not bad, (name, bad[:3])
