"""The C ABI without a GPU: the library loads, exports exactly what
include/clover_mi355x.h declares, and its host-side deck reader agrees with the
test-side restatement of driver/read_input.cpp."""
import re
import subprocess
from pathlib import Path

import pytest

import clover_testlib as T
import cloverleaf_amd as clv

REPO = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (REPO / "include" / "clover_mi355x.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    inline = set(re.findall(r"static inline [^(]*\b(clv_[A-Za-z0-9_]+)\s*\(", text))
    return sorted(set(re.findall(r"\b(clv_[A-Za-z0-9_]+)\s*\(", text)) - inline)


def test_library_exports_every_declared_symbol():
    so = clv.LIB_PATH
    assert so.exists(), "build the library first (make -C cloverleaf_amd/csrc)"
    out = subprocess.run(["nm", "-D", "--defined-only", str(so)], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (clv_[A-Za-z0-9_]+)", out))
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing
    bound = {name for name, _, _ in clv.SIGNATURES}
    assert set(declared_symbols()) == bound, set(declared_symbols()) ^ bound


def test_library_loads_and_reports_version():
    L = clv.lib()
    assert b"gfx950" in L.clv_version()


@pytest.mark.parametrize("deck", sorted(p.stem for p in T.DECKS.glob("*.in")))
def test_deck_parse_matches_reference_semantics(deck):
    text = (T.DECKS / f"{deck}.in").read_text()
    cfg = clv.parse_deck(text)
    d = T.parse_deck(text)
    for k in ("xmin", "ymin", "xmax", "ymax", "x_cells", "y_cells", "dtinit", "dtmax", "dtmin", "dtrise", "end_time",
              "end_step", "summary_frequency", "test_problem", "dtc_safe", "dtu_safe", "dtv_safe", "dtdiv_safe"):
        assert getattr(cfg, k) == getattr(d, k), k
    assert cfg.number_of_states == len(d.states)
    for i, s in enumerate(d.states):
        for k in ("density", "energy", "xvel", "yvel", "geometry", "xmin", "ymin", "xmax", "ymax", "radius"):
            assert getattr(cfg.states[i], k) == getattr(s, k), (i, k)


def test_deck_errors():
    with pytest.raises(clv.CloverError, match="No states"):
        clv.parse_deck("*clover\n x_cells=10\n*endclover\n")
    with pytest.raises(clv.CloverError, match="defined twice"):
        clv.parse_deck("*clover\n state 1 density=1 energy=1\n state 1 density=2 energy=1\n*endclover\n")


def test_chunk_create_without_gpu_fails_loudly():
    if clv.lib().clv_device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(clv.CloverError, match="no HIP device"):
        clv.Chunk(16, 16)
