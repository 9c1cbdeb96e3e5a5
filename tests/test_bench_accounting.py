"""bench.py's byte accounting for the fused step agrees with itself: each fused
kernel's compulsory arrays (ALG_TERMS) are its march-probe shape (PATTERN, reads +
writes), less the volume and face areas in mesh mode (GEOM_ARRAYS), and one fused
step's compulsory bytes (B_FUSED_PER_CELL_STEP) are the timestep, the lagrange pass
and one advection pass of each sweep."""
import importlib.util
from pathlib import Path

REPO = Path(__file__).resolve().parent.parent
_spec = importlib.util.spec_from_file_location("bench_under_test", REPO / "bench.py")
bench = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(bench)

FUSED = ["f_timestep", "f_lagrange", "f_cm_x1", "f_cm_y1", "f_cm_x2", "f_cm_y2"]


def _arrays(name, mesh):
    (k, extent), = bench.ALG_TERMS[name]
    assert extent == "C"
    return k - (bench.GEOM_ARRAYS[name] if mesh else 0)


def test_probe_shape_is_the_compulsory_arrays():
    for mesh in (False, True):
        for name in FUSED:
            reads, writes = bench.PATTERN[mesh][name]
            assert reads + writes == _arrays(name, mesh), (name, mesh)


def test_step_bytes_are_the_sum_of_its_kernels():
    nx, ny = 1000, 700
    for mesh in (False, True):
        per_cell = {n: bench.alg_bytes(n, nx, ny, mesh) / (nx * ny) for n in FUSED}
        assert per_cell["f_cm_x1"] == per_cell["f_cm_y1"] and per_cell["f_cm_x2"] == per_cell["f_cm_y2"]
        step = per_cell["f_timestep"] + per_cell["f_lagrange"] + per_cell["f_cm_x1"] + per_cell["f_cm_y2"]
        assert step == bench.B_FUSED_PER_CELL_STEP[mesh], mesh


def test_per_entry_point_forms_store_more():
    """The per-entry-point forms store what the fused step keeps in registers: q (and p,
    c) in the timestep pass, p' and c' in the lagrange pass (which reads q rather than
    recomputing it), the mass fluxes in the advection passes."""
    assert _arrays("f_timestep_q", False) == _arrays("f_timestep", False) + 1
    assert _arrays("f_timestep_pc", False) == _arrays("f_timestep", False) + 3
    assert _arrays("f_lagrange_pc", False) == _arrays("f_lagrange", False) + 3
    for s in ("x1", "y1", "x2", "y2"):
        assert _arrays(f"f_cm_{s}_mf", False) == _arrays(f"f_cm_{s}", False) + 1
