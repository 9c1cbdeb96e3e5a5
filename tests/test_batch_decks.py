"""The seeded batch workload's deck generator (clv_batch_deck), on the CPU.

The generator has no counterpart in the reference; it is pinned by determinism,
by the ranges and distribution it promises, by the reference's own parser
semantics (the library's clv_deck_parse, itself checked against
driver/read_input.cpp), and by the sha256 of every deck recorded in
tests/golden/batch_seed355.json next to the reference's results for it.
"""
import hashlib
import json
import re

import numpy as np
import pytest

import cloverleaf_amd as clv

GOLDEN = clv.PKG.parent / "tests" / "golden" / "batch_seed355.json"


def _params(text):
    m = re.search(r"state 2 density=(\S+) energy=(\S+) geometry=rectangle xmin=(\S+) xmax=(\S+) ymin=(\S+) ymax=(\S+)", text)
    return [float(x) for x in m.groups()]


def test_deterministic_and_seeded():
    a = [clv.batch_deck(355, i) for i in range(8)]
    assert a == [clv.batch_deck(355, i) for i in range(8)]
    assert len(set(a)) == 8
    assert a != [clv.batch_deck(356, i) for i in range(8)]


def test_ranges_and_fixed_fields():
    for i in range(200):
        t = clv.batch_deck(355, i)
        cfg = clv.parse_deck(t)
        assert cfg.x_cells == cfg.y_cells and (cfg.x_cells - 256) % 128 == 0 and 256 <= cfg.x_cells <= 2048
        rho, e, x0, x1, y0, y1 = _params(t)
        assert 0.5 <= rho < 2.0 and 1.5 <= e < 3.5
        assert 0.0 <= x0 <= x1 < 10.0 and 0.0 <= y0 <= y1 < 10.0
        assert (cfg.xmin, cfg.ymin, cfg.xmax, cfg.ymax) == (0.0, 0.0, 10.0, 10.0)
        assert (cfg.dtinit, cfg.dtmax, cfg.dtrise, cfg.end_step) == (0.04, 0.04, 1.5, 100)
        assert cfg.number_of_states == 2
        s1, s2 = cfg.states[0], cfg.states[1]
        assert (s1.density, s1.energy) == (0.2, 1.0)
        assert (s2.density, s2.energy, s2.geometry) == (rho, e, 1)


def test_distribution():
    sides = np.array([clv.parse_deck(clv.batch_deck(7, i)).x_cells for i in range(3000)])
    counts = np.bincount((sides - 256) // 128, minlength=15)
    assert len(counts) == 15 and counts.min() > 150  # uniform over the 15 sides: ~200 each
    mean_cells = float(np.mean(sides.astype(np.float64) ** 2))
    expect = np.mean((256.0 + 128.0 * np.arange(15)) ** 2)  # 16384 * E[(2 + k)^2] = 1.633e6
    assert abs(mean_cells / expect - 1.0) < 0.05


def test_text_round_trips_17_digits():
    t = clv.batch_deck(355, 3)
    for v in _params(t):
        assert float(repr(v)) == v
    assert "%.17g" % _params(t)[0] in t


@pytest.mark.skipif(not GOLDEN.exists(), reason="batch golden not generated")
def test_golden_decks_match_generator():
    g = json.loads(GOLDEN.read_text())
    assert g["seed"] == 355 and len(g["problems"]) == 64
    for i, prob in enumerate(g["problems"]):
        t = clv.batch_deck(355, i)
        assert hashlib.sha256(t.encode()).hexdigest() == prob["deck_sha256"], i
        assert prob["side"] == clv.parse_deck(t).x_cells
        assert len(prob["dt"]) == 100 and all(0 < d <= 0.04 for d in prob["dt"])
        assert [r["step"] for r in prob["summaries"]] == [1, 50, 100]
        # the reference's volume sum is the domain area to round-off (dx = 10 / side is
        # inexact for most sides); mass is conserved to the small drift the reference's
        # reflective walls allow once the flow reaches them (up to ~1e-6 over 100 steps)
        v0, m100 = prob["summaries"][0]["vol"], prob["summaries"][2]["mass"]
        assert abs(v0 - 100.0) < 1e-10 * 100.0 and abs(m100 / prob["summaries"][0]["mass"] - 1.0) < 1e-5
