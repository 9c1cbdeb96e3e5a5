"""cloverleaf_amd -- MI355X-native CloverLeaf hydro kernels (gfx950, FP64).

Python face of the C ABI in include/clover_mi355x.h (libclover_mi355x.so, built
by ``make -C cloverleaf_amd/csrc``).  The HIP library is the only compute path:
importing the package does not need a GPU, but every compute call goes through
the library and raises if it is missing or fails; there is no CPU fallback.

    from cloverleaf_amd import Chunk, parse_deck
    cfg = parse_deck(open("clover_bm_short.in").read())
    ch = Chunk(cfg.x_cells, cfg.y_cells)
    st, s0 = ch.start(cfg)
    while not st.complete:
        ch.step(cfg, st)
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libclover_mi355x.so"

# field ids (include/clover_mi355x.h; 0..14 = driver/definitions.h:62-78)
FIELDS = ["density0", "density1", "energy0", "energy1", "pressure", "viscosity", "soundspeed",
          "xvel0", "xvel1", "yvel0", "yvel1", "vol_flux_x", "vol_flux_y", "mass_flux_x", "mass_flux_y",
          "work1", "work2", "work3", "work4", "work5", "work6", "work7", "volume", "xarea", "yarea"]
FID = {n: i for i, n in enumerate(FIELDS)}
MESH = ["cellx", "celly", "vertexx", "vertexy", "celldx", "celldy", "vertexdx", "vertexdy"]
MID = {n: i for i, n in enumerate(MESH)}
PLAN_REFERENCE, PLAN_FUSED = 0, 1
SUMMARY_TREE, SUMMARY_SERIAL = 0, 1
GEOMETRY_ARRAYS, GEOMETRY_MESH = 0, 1
ENTRY_REFERENCE, ENTRY_FUSED = 0, 1
MAX_STATES = 16
ERR_SMALL_DT = -4


class CloverError(RuntimeError):
    pass


class SmallTimestep(CloverError):
    """driver/timestep.cpp:104-106 -- dt fell below dtmin."""


class _State(C.Structure):
    _fields_ = [("density", C.c_double), ("energy", C.c_double), ("xvel", C.c_double), ("yvel", C.c_double),
                ("geometry", C.c_int), ("xmin", C.c_double), ("ymin", C.c_double), ("xmax", C.c_double),
                ("ymax", C.c_double), ("radius", C.c_double)]


class Config(C.Structure):
    """clv_config: the deck (driver/definitions.h:97-155)."""
    _fields_ = [("xmin", C.c_double), ("ymin", C.c_double), ("xmax", C.c_double), ("ymax", C.c_double),
                ("x_cells", C.c_int), ("y_cells", C.c_int),
                ("dtinit", C.c_double), ("dtmin", C.c_double), ("dtmax", C.c_double), ("dtrise", C.c_double),
                ("dtc_safe", C.c_double), ("dtu_safe", C.c_double), ("dtv_safe", C.c_double),
                ("dtdiv_safe", C.c_double), ("end_time", C.c_double), ("end_step", C.c_int),
                ("summary_frequency", C.c_int), ("visit_frequency", C.c_int), ("tiles_per_chunk", C.c_int),
                ("test_problem", C.c_int), ("profiler_on", C.c_int), ("number_of_states", C.c_int),
                ("states", _State * MAX_STATES)]


class HydroState(C.Structure):
    """clv_hydro_state (driver/definitions.h:265-274)."""
    _fields_ = [("step", C.c_int), ("advect_x", C.c_int), ("time", C.c_double), ("dt", C.c_double),
                ("dtold", C.c_double), ("jdt", C.c_int), ("kdt", C.c_int), ("dt_control", C.c_int),
                ("complete", C.c_int)]


class BatchFamily(C.Structure):
    """clv_batch_family: rows per tile, persistent blocks per CU (0: one per tile), member order."""
    _fields_ = [("seg", C.c_int), ("blocks_per_cu", C.c_int), ("order", C.c_int)]


BATCH_FAMILIES = ["timestep", "lagrange", "cm_x", "cm_y"]
ORDER_GIVEN, ORDER_LARGEST_FIRST = 0, 1


class BatchSchedule(C.Structure):
    _fields_ = [("fam", BatchFamily * len(BATCH_FAMILIES))]

    def as_dict(self) -> dict:
        return {n: dict(seg=f.seg, blocks_per_cu=f.blocks_per_cu, order=f.order)
                for n, f in zip(BATCH_FAMILIES, self.fam)}


class ChunkView(C.Structure):
    _fields_ = [("x_cells", C.c_int), ("y_cells", C.c_int), ("pitch", C.c_int64),
                ("field", C.c_void_p * 25), ("mesh", C.c_void_p * 8)]


_LIB = None

# (name, restype, argtypes) of every exported entry point
_P = C.c_void_p
_D = C.c_double
_I = C.c_int
_PD = C.POINTER(C.c_double)
_PI = C.POINTER(C.c_int)
SIGNATURES = [
    ("clv_last_error", C.c_char_p, []),
    ("clv_version", C.c_char_p, []),
    ("clv_device_count", _I, []),
    ("clv_chunk_create", _I, [_I, _I, C.POINTER(_P)]),
    ("clv_chunk_wrap", _I, [C.POINTER(ChunkView), C.POINTER(_P)]),
    ("clv_chunk_destroy", _I, [_P]),
    ("clv_chunk_get_view", _I, [_P, C.POINTER(ChunkView)]),
    ("clv_chunk_geometry_from_mesh", _I, [_P]),
    ("clv_chunk_set_geometry_mode", _I, [_P, _I]),
    ("clv_chunk_set_entry_plan", _I, [_P, _I]),
    ("clv_chunk_set_stream", _I, [_P, _P]),
    ("clv_chunk_synchronize", _I, [_P]),
    ("clv_chunk_bytes", C.c_size_t, [_P]),
    ("clv_field_upload", _I, [_P, _I, _PD, C.c_int64]),
    ("clv_field_download", _I, [_P, _I, _PD, C.c_int64]),
    ("clv_mesh_upload", _I, [_P, _I, _PD]),
    ("clv_mesh_download", _I, [_P, _I, _PD]),
    ("clv_build_field", _I, [_P]),
    ("clv_initialise_chunk", _I, [_P, C.POINTER(Config)]),
    ("clv_generate_chunk", _I, [_P, C.POINTER(Config)]),
    ("clv_ideal_gas", _I, [_P, _I]),
    ("clv_viscosity", _I, [_P]),
    ("clv_calc_dt", _I, [_P, _D, _D, _D, _D, _D, _PD, _PI, _PI, _PI, _PI]),
    ("clv_timestep_kernels", _I, [_P, _D, _D, _D, _D, _D, _I, _PD, _PI, _PI, _PI, _PI]),
    ("clv_lagrange_kernels", _I, [_P, _D]),
    ("clv_advection_pass", _I, [_P, _I, _I]),
    ("clv_advection_pass_reset", _I, [_P, _I, _I]),
    ("clv_PdV", _I, [_P, _I, _D]),
    ("clv_PdV_driver", _I, [_P, _I, _D]),
    ("clv_revert", _I, [_P]),
    ("clv_accelerate", _I, [_P, _D]),
    ("clv_flux_calc", _I, [_P, _D]),
    ("clv_advec_cell", _I, [_P, _I, _I]),
    ("clv_advec_mom", _I, [_P, _I, _I, _I]),
    ("clv_reset_field", _I, [_P]),
    ("clv_update_halo", _I, [_P, _PI, _I]),
    ("clv_initialise_tile", _I, [_P, C.POINTER(Config), _I, _I]),
    ("clv_update_halo_faces", _I, [_P, _PI, _I, _PI]),
    ("clv_update_tile_halo", _I, [_P, _P, _I, _PI, _I]),
    ("clv_pack_message", _I, [_P, _P, _I, _I, _I, _P, C.c_size_t, C.c_int64]),
    ("clv_unpack_message", _I, [_P, _P, _I, _I, _I, _P, C.c_size_t, C.c_int64]),
    ("clv_message_length", C.c_size_t, [_PI, _I, _I]),
    ("clv_pack_fields", _I, [_P, _I, _PI, _I, _I, C.c_int64, _P, C.c_size_t]),
    ("clv_unpack_fields", _I, [_P, _I, _PI, _I, _I, C.c_int64, _P, C.c_size_t]),
    ("clv_field_summary", _I, [_P, _PD]),
    ("clv_chunk_set_summary_order", _I, [_P, _I]),
    ("clv_field_summary_noalloc", _I, [C.POINTER(ChunkView), _P, _PD]),
    ("clv_deck_parse", _I, [C.c_char_p, C.POINTER(Config)]),
    ("clv_hydro_start", _I, [_P, C.POINTER(Config), C.POINTER(HydroState), _PD]),
    ("clv_timestep", _I, [_P, C.POINTER(Config), C.POINTER(HydroState), _I]),
    ("clv_hydro_step", _I, [_P, C.POINTER(Config), C.POINTER(HydroState), _I]),
    ("clv_hydro_summary", _I, [_P, _PD]),
    ("clv_hydro_phase", _I, [_P, C.POINTER(Config), C.POINTER(HydroState), _I]),
    ("clv_profile_enable", _I, [_P, _I]),
    ("clv_profile_read", _I, [_P, C.POINTER(C.c_char_p), C.POINTER(C.c_int64), _PD, _I]),
    ("clv_profile_reset", _I, [_P]),
    ("clv_bandwidth_probe", _I, [C.c_size_t, _I, _PD]),
    ("clv_march_probe", _I, [_I, _I, _I, _I, _I, _I, _PD]),
    ("clv_device_alloc", _I, [C.c_size_t, C.POINTER(_P)]),
    ("clv_device_free", _I, [_P]),
    ("clv_copy_to_host", _I, [_P, _P, C.c_size_t]),
    ("clv_copy_to_device", _I, [_P, _P, C.c_size_t]),
    ("clv_batch_deck", _I, [C.c_uint64, _I, C.c_char_p, C.c_size_t]),
    ("clv_batch_create", _I, [C.POINTER(_P), C.POINTER(Config), C.POINTER(HydroState), _I, _I, C.POINTER(_P)]),
    ("clv_batch_destroy", _I, [_P]),
    ("clv_batch_get_schedule", _I, [_P, C.POINTER(BatchSchedule)]),
    ("clv_batch_set_schedule", _I, [_P, C.POINTER(BatchSchedule)]),
    ("clv_batch_autotune", _I, [_P, _I, C.POINTER(BatchSchedule), _PD]),
    ("clv_batch_run", _I, [_P, _I, _I, _PI]),
    ("clv_batch_states", _I, [_P, C.POINTER(HydroState)]),
    ("clv_batch_dt_history", _I, [_P, _I, _PD, _I, _PI]),
    ("clv_adj_ideal_gas", _I, [_P, _P, _I]),
    ("clv_adj_viscosity", _I, [_P, _P]),
    ("clv_adj_PdV", _I, [_P, _P, _I, _D]),
    ("clv_adj_revert", _I, [_P, _P]),
    ("clv_adj_accelerate", _I, [_P, _P, _D]),
    ("clv_adj_flux_calc", _I, [_P, _P, _D]),
    ("clv_adj_advec_cell", _I, [_P, _P, _I, _I]),
    ("clv_adj_advec_mom", _I, [_P, _P, _I, _I]),
    ("clv_adj_reset_field", _I, [_P, _P]),
    ("clv_adj_update_halo", _I, [_P, _PI, _I]),
    ("clv_adj_kinetic_energy", _I, [_P, _P, _D]),
    ("clv_adjoint_create", _I, [_P, _D, _I, _I, C.POINTER(_P)]),
    ("clv_adjoint_destroy", _I, [_P]),
    ("clv_adjoint_bytes", C.c_size_t, [_P]),
    ("clv_adjoint_set_initial", _I, [_P, _P, _P, C.c_int64]),
    ("clv_adjoint_forward", _I, [_P, _PD]),
    ("clv_adjoint_backward", _I, [_P, _D, _P, _P, C.c_int64]),
    ("clv_adjoint_chunks", _I, [_P, C.POINTER(_P), C.POINTER(_P)]),
]


def bandwidth_probe(nbytes: int = 4 << 30, iters: int = 20) -> float:
    """Copy bandwidth (GB/s, read + write) of the current GPU's HBM."""
    out = C.c_double()
    _check(lib().clv_bandwidth_probe(int(nbytes), int(iters), C.byref(out)), "clv_bandwidth_probe")
    return out.value


def march_probe(nx: int, ny: int, reads: int, writes: int, seg: int = 16, iters: int = 10) -> float:
    """ms per launch of the fused kernels' march over an nx x ny grid reading `reads`
    and writing `writes` arrays, no arithmetic: the ceiling of a fused kernel with
    those arrays."""
    out = C.c_double()
    _check(lib().clv_march_probe(int(nx), int(ny), int(reads), int(writes), int(seg), int(iters), C.byref(out)), "clv_march_probe")
    return out.value


def lib() -> C.CDLL:
    """Load libclover_mi355x.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise CloverError(f"{LIB_PATH} is missing: build it with `make -C {PKG / 'csrc'}` "
                              "(or __graft_entry__.build())")
        L = C.CDLL(str(LIB_PATH))
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def _check(rc: int, what: str):
    if rc != 0:
        msg = lib().clv_last_error().decode()
        if rc == ERR_SMALL_DT:
            raise SmallTimestep(f"{what}: {msg}")
        raise CloverError(f"{what} failed ({rc}): {msg}")


def parse_deck(text: str) -> Config:
    """Deck text -> Config, with driver/read_input.cpp semantics (done by the library)."""
    cfg = Config()
    _check(lib().clv_deck_parse(text.encode(), C.byref(cfg)), "clv_deck_parse")
    return cfg


FACE_LEFT, FACE_RIGHT, FACE_BOTTOM, FACE_TOP = 0, 1, 2, 3
CELL_DATA, VERTEX_DATA, X_FACE_DATA, Y_FACE_DATA = 1, 2, 3, 4


def message_length(fields, depth: int, chunk_cells: int) -> int:
    """Doubles in one face message of `fields` (clv_message_length)."""
    return int(lib().clv_message_length(_mask(fields), depth, chunk_cells))


class DeviceBuffer:
    """n doubles of device memory (clv_device_alloc), e.g. an exchange message."""

    def __init__(self, n: int):
        self.L = lib()
        self.n = int(n)
        p = C.c_void_p()
        _check(self.L.clv_device_alloc(self.n * 8, C.byref(p)), "clv_device_alloc")
        self.ptr = p.value

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == self.n
        _check(self.L.clv_copy_to_device(self.ptr, a.ctypes.data, self.n * 8), "clv_copy_to_device")

    def download(self) -> np.ndarray:
        out = np.empty(self.n)
        _check(self.L.clv_copy_to_host(out.ctypes.data, self.ptr, self.n * 8), "clv_copy_to_host")
        return out

    def close(self):
        if getattr(self, "ptr", None):
            self.L.clv_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _mask(names) -> "C.Array":
    m = (C.c_int * 15)()
    for n in names:
        m[FID[n] if isinstance(n, str) else int(n)] = 1
    return m


class Chunk:
    """One chunk resident on the current GPU (one arena allocation)."""

    def __init__(self, x_cells: int, y_cells: int, geometry: int = GEOMETRY_ARRAYS):
        """geometry: GEOMETRY_ARRAYS (default: the kernels read volume, xarea and yarea)
        or GEOMETRY_MESH (computed from the 1D spacings while the arrays hold what
        initialise_chunk wrote)."""
        self.L = lib()
        self.nx, self.ny = int(x_cells), int(y_cells)
        h = C.c_void_p()
        _check(self.L.clv_chunk_create(self.nx, self.ny, C.byref(h)), "clv_chunk_create")
        self.h = h
        self.set_geometry_mode(geometry)

    def set_geometry_mode(self, mode: int):
        _check(self.L.clv_chunk_set_geometry_mode(self.h, int(mode)), "set_geometry_mode")

    def set_entry_plan(self, plan: int):
        """ENTRY_REFERENCE (one launch per reference loop) or ENTRY_FUSED (advec_cell,
        advec_mom and the PdV predictor chain one launch each)."""
        _check(self.L.clv_chunk_set_entry_plan(self.h, int(plan)), "set_entry_plan")

    def close(self):
        if getattr(self, "h", None):
            self.L.clv_chunk_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- data movement (host arrays in the oracle layout: (ny+5, nx+5)) ----
    def download(self, name: str) -> np.ndarray:
        a = np.empty((self.ny + 5, self.nx + 5), dtype=np.float64)
        _check(self.L.clv_field_download(self.h, FID[name], a.ctypes.data_as(_PD), self.nx + 5), "download")
        return a

    def upload(self, name: str, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == (self.ny + 5, self.nx + 5), a.shape
        _check(self.L.clv_field_upload(self.h, FID[name], a.ctypes.data_as(_PD), self.nx + 5), "upload")

    def mesh_length(self, name: str) -> int:
        """Cell arrays n + 4, vertex arrays n + 5 (driver/definitions.cpp:28-35)."""
        n = self.nx if name in ("cellx", "vertexx", "celldx", "vertexdx") else self.ny
        return n + (5 if name.startswith("vertex") else 4)

    def download_mesh(self, name: str) -> np.ndarray:
        n = self.mesh_length(name)
        a = np.empty(n, dtype=np.float64)
        _check(self.L.clv_mesh_download(self.h, MID[name], a.ctypes.data_as(_PD)), "download_mesh")
        return a

    def upload_mesh(self, name: str, a: np.ndarray):
        n = self.mesh_length(name)
        a = np.ascontiguousarray(np.asarray(a, dtype=np.float64)[:n])
        assert a.shape == (n,), a.shape
        _check(self.L.clv_mesh_upload(self.h, MID[name], a.ctypes.data_as(_PD)), "upload_mesh")

    def synchronize(self):
        _check(self.L.clv_chunk_synchronize(self.h), "synchronize")

    def view(self) -> ChunkView:
        v = ChunkView()
        _check(self.L.clv_chunk_get_view(self.h, C.byref(v)), "view")
        return v

    # ---- per-kernel entry points ----
    def ideal_gas(self, predict=False):
        _check(self.L.clv_ideal_gas(self.h, int(bool(predict))), "ideal_gas")

    def viscosity(self):
        _check(self.L.clv_viscosity(self.h), "viscosity")

    def calc_dt(self, dtmin=1e-7, dtc_safe=0.7, dtu_safe=0.5, dtv_safe=0.5, dtdiv_safe=0.7):
        dt = C.c_double()
        ctl, jl, kl, small = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _check(self.L.clv_calc_dt(self.h, dtmin, dtc_safe, dtu_safe, dtv_safe, dtdiv_safe, C.byref(dt), C.byref(ctl),
                                  C.byref(jl), C.byref(kl), C.byref(small)), "calc_dt")
        return dt.value, ctl.value, jl.value, kl.value, small.value

    def timestep_kernels(self, pressure_halo_depth=1, dtmin=1e-7, dtc_safe=0.7, dtu_safe=0.5, dtv_safe=0.5, dtdiv_safe=0.7):
        """ideal_gas, viscosity and calc_dt in one pass, with the pressure and viscosity
        halos (clv_timestep_kernels)."""
        dt = C.c_double()
        ctl, jl, kl, small = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _check(self.L.clv_timestep_kernels(self.h, dtmin, dtc_safe, dtu_safe, dtv_safe, dtdiv_safe, int(pressure_halo_depth),
                                           C.byref(dt), C.byref(ctl), C.byref(jl), C.byref(kl), C.byref(small)),
               "timestep_kernels")
        return dt.value, ctl.value, jl.value, kl.value, small.value

    def lagrange_kernels(self, dt):
        """PdV(predict) chain, accelerate, PdV(correct) and flux_calc in one pass
        (clv_lagrange_kernels)."""
        _check(self.L.clv_lagrange_kernels(self.h, float(dt)), "lagrange_kernels")

    def advection_pass(self, sweep_number, direction):
        """advec_cell, the halo between, advec_mom for both velocities in one pass
        (clv_advection_pass)."""
        _check(self.L.clv_advection_pass(self.h, sweep_number, direction), "advection_pass")

    def advection_pass_reset(self, sweep_number, direction):
        """advection_pass followed by reset_field, the copy done by the pass's own stores
        (clv_advection_pass_reset)."""
        _check(self.L.clv_advection_pass_reset(self.h, sweep_number, direction), "advection_pass_reset")

    def PdV(self, predict, dt):
        _check(self.L.clv_PdV(self.h, int(bool(predict)), float(dt)), "PdV")

    def PdV_driver(self, predict, dt):
        _check(self.L.clv_PdV_driver(self.h, int(bool(predict)), float(dt)), "PdV_driver")

    def revert(self):
        _check(self.L.clv_revert(self.h), "revert")

    def accelerate(self, dt):
        _check(self.L.clv_accelerate(self.h, float(dt)), "accelerate")

    def flux_calc(self, dt):
        _check(self.L.clv_flux_calc(self.h, float(dt)), "flux_calc")

    def advec_cell(self, sweep_number, direction):
        _check(self.L.clv_advec_cell(self.h, sweep_number, direction), "advec_cell")

    def advec_mom(self, which_vel, direction, sweep_number):
        _check(self.L.clv_advec_mom(self.h, which_vel, direction, sweep_number), "advec_mom")

    def reset_field(self):
        _check(self.L.clv_reset_field(self.h), "reset_field")

    def update_halo(self, fields, depth):
        _check(self.L.clv_update_halo(self.h, _mask(fields), depth), "update_halo")

    # ---- decomposed domains (faces: FACE_LEFT, FACE_RIGHT, FACE_BOTTOM, FACE_TOP) ----
    def initialise_tile(self, cfg: Config, left: int, bottom: int):
        """initialise_chunk for a tile whose first cell is (left, bottom) of the grid."""
        _check(self.L.clv_initialise_tile(self.h, C.byref(cfg), int(left), int(bottom)), "initialise_tile")

    def update_halo_faces(self, fields, depth, external):
        """The reflective halo on the faces with external[face] true only."""
        ext = (C.c_int * 4)(*[1 if e else 0 for e in external])
        _check(self.L.clv_update_halo_faces(self.h, _mask(fields), depth, ext), "update_halo_faces")

    def update_tile_halo(self, neighbour: "Chunk", face: int, fields, depth):
        """This tile's ghosts on `face` from the neighbouring tile's interior."""
        _check(self.L.clv_update_tile_halo(self.h, neighbour.h, int(face), _mask(fields), depth), "update_tile_halo")

    def pack_message(self, field_ptr: int, field_type: int, face: int, depth: int, buf: "DeviceBuffer", offset: int, unpack=False):
        """One array's exchange message (field_ptr: a device array of this chunk's layout)."""
        fn = self.L.clv_unpack_message if unpack else self.L.clv_pack_message
        _check(fn(self.h, field_ptr, field_type, face, depth, buf.ptr, buf.n, int(offset)), "pack_message")

    def pack_fields(self, face: int, fields, depth: int, chunk_cells: int, tile_offset: int, buf: "DeviceBuffer", unpack=False):
        """Every requested field of this tile into (or out of) one face message."""
        fn = self.L.clv_unpack_fields if unpack else self.L.clv_pack_fields
        _check(fn(self.h, int(face), _mask(fields), depth, int(chunk_cells), int(tile_offset), buf.ptr, buf.n), "pack_fields")

    def geometry_from_mesh(self) -> bool:
        """True while volume, xarea and yarea hold what initialise_chunk wrote (the fused
        plan then derives them from the mesh instead of reading them)."""
        return self.L.clv_chunk_geometry_from_mesh(self.h) == 1

    def field_ptr(self, name: str) -> int:
        """Device address of element (0, 0) of a 2D field."""
        v = ChunkView()
        _check(self.L.clv_chunk_get_view(self.h, C.byref(v)), "clv_chunk_get_view")
        return int(v.field[FID[name]])

    def field_summary(self):
        out = (C.c_double * 5)()
        _check(self.L.clv_field_summary(self.h, out), "field_summary")
        return list(out)

    def field_summary_noalloc(self, stream: int = 0):
        """clv_field_summary over this chunk's arrays through the allocation-free entry
        point (a view of the arrays, the library's fixed scratch array)."""
        out = (C.c_double * 5)()
        v = ChunkView()
        _check(self.L.clv_chunk_get_view(self.h, C.byref(v)), "clv_chunk_get_view")
        _check(self.L.clv_field_summary_noalloc(C.byref(v), C.c_void_p(stream or None), out), "clv_field_summary_noalloc")
        return list(out)

    def set_summary_order(self, order: int):
        """SUMMARY_TREE (default) or SUMMARY_SERIAL (bit-identical to the serial model)."""
        _check(self.L.clv_chunk_set_summary_order(self.h, order), "set_summary_order")

    def build_field(self):
        _check(self.L.clv_build_field(self.h), "build_field")

    def initialise_chunk(self, cfg: Config):
        _check(self.L.clv_initialise_chunk(self.h, C.byref(cfg)), "initialise_chunk")

    def generate_chunk(self, cfg: Config):
        _check(self.L.clv_generate_chunk(self.h, C.byref(cfg)), "generate_chunk")

    # ---- whole-step forms ----
    def start(self, cfg: Config):
        st = HydroState()
        out = (C.c_double * 5)()
        _check(self.L.clv_hydro_start(self.h, C.byref(cfg), C.byref(st), out), "hydro_start")
        return st, list(out)

    def step(self, cfg: Config, st: HydroState, plan: int = PLAN_FUSED):
        _check(self.L.clv_hydro_step(self.h, C.byref(cfg), C.byref(st), plan), "hydro_step")
        return st

    def summary(self):
        out = (C.c_double * 5)()
        _check(self.L.clv_hydro_summary(self.h, out), "hydro_summary")
        return list(out)

    # ---- measurement ----
    def profile(self, on=True):
        _check(self.L.clv_profile_enable(self.h, int(bool(on))), "profile_enable")

    def profile_reset(self):
        _check(self.L.clv_profile_reset(self.h), "profile_reset")

    def profile_read(self) -> dict:
        n = 128
        names = (C.c_char_p * n)()
        launches = (C.c_int64 * n)()
        ms = (C.c_double * n)()
        k = self.L.clv_profile_read(self.h, names, launches, ms, n)
        return {names[i].decode(): (int(launches[i]), float(ms[i])) for i in range(min(k, n))}


def run(cfg: Config, plan: int = PLAN_FUSED, trace: bool = False, summary_order: int = 0,
        geometry: int = GEOMETRY_ARRAYS):
    """Run a deck to completion (driver/hydro.cpp:53-111); returns the summary records."""
    recs = []
    with Chunk(cfg.x_cells, cfg.y_cells, geometry) as ch:
        ch.set_summary_order(summary_order)
        st, s = ch.start(cfg)
        recs.append(dict(step=0, time=st.time, dt=st.dt, vol=s[0], mass=s[1], ie=s[2], ke=s[3], press=s[4]))
        while True:
            ch.step(cfg, st, plan)
            if trace or st.complete or (cfg.summary_frequency and st.step % cfg.summary_frequency == 0):
                s = ch.summary()
                recs.append(dict(step=st.step, time=st.time, dt=st.dt, vol=s[0], mass=s[1], ie=s[2], ke=s[3],
                                 press=s[4]))
            if st.complete:
                return recs


def batch_deck(seed: int, index: int) -> str:
    """Deck text of problem `index` of the seeded batch workload (clv_batch_deck)."""
    buf = C.create_string_buffer(4096)
    _check(lib().clv_batch_deck(C.c_uint64(seed), int(index), buf, len(buf)), "clv_batch_deck")
    return buf.value.decode()


class Batch:
    """Many independent problems advanced together (clv_batch_*).

    Each member is a started Chunk; the batch owns nothing but its schedule,
    views and per-member driver state.  Members must not be stepped outside the
    batch while it exists (reading them is fine between runs).
    """

    def __init__(self, chunks, cfgs, states, hist_len: int = 0):
        self.L = lib()
        n = len(chunks)
        self.chunks = list(chunks)
        self.n = n
        hs = (_P * n)(*[c.h.value for c in chunks])
        cf = (Config * n)(*cfgs)
        st = (HydroState * n)(*states)
        h = C.c_void_p()
        _check(self.L.clv_batch_create(hs, cf, st, n, int(hist_len), C.byref(h)), "clv_batch_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.clv_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def schedule(self) -> BatchSchedule:
        s = BatchSchedule()
        _check(self.L.clv_batch_get_schedule(self.h, C.byref(s)), "clv_batch_get_schedule")
        return s

    def set_schedule(self, s: BatchSchedule):
        _check(self.L.clv_batch_set_schedule(self.h, C.byref(s)), "clv_batch_set_schedule")

    def autotune(self, reps: int = 3):
        s = BatchSchedule()
        ms = (C.c_double * len(BATCH_FAMILIES))()
        _check(self.L.clv_batch_autotune(self.h, int(reps), C.byref(s), ms), "clv_batch_autotune")
        return s, list(ms)

    def run(self, max_steps: int, use_graph: bool = True) -> int:
        done = C.c_int()
        _check(self.L.clv_batch_run(self.h, int(max_steps), int(bool(use_graph)), C.byref(done)), "clv_batch_run")
        return done.value

    def states(self):
        out = (HydroState * self.n)()
        _check(self.L.clv_batch_states(self.h, out), "clv_batch_states")
        return list(out)

    def dt_history(self, member: int, max_steps: int = 1 << 20) -> np.ndarray:
        a = np.empty(max_steps, dtype=np.float64)
        k = C.c_int()
        _check(self.L.clv_batch_dt_history(self.h, int(member), a.ctypes.data_as(_PD), int(max_steps), C.byref(k)),
               "clv_batch_dt_history")
        return a[:k.value].copy()


def adjoint_kernel(name: str, primal: Chunk, adjoint: Chunk, *args):
    """Run the adjoint of one reference kernel (clv_adj_<name>): primal holds the
    forward kernel's input state, adjoint the output adjoints (in) / input
    adjoints (out)."""
    L = lib()
    if name == "update_halo":
        fields, depth = args
        _check(L.clv_adj_update_halo(adjoint.h, _mask(fields), int(depth)), "clv_adj_update_halo")
        return
    _check(getattr(L, f"clv_adj_{name}")(primal.h, adjoint.h, *args), f"clv_adj_{name}")


class Adjoint:
    """Checkpointed gradient of the final kinetic energy of a fixed-dt run with
    respect to the initial density0 / energy0 interior fields (clv_adjoint_*)."""

    def __init__(self, initial: Chunk, dt: float, steps: int, checkpoint_every: int = 20):
        self.L = lib()
        self.nx, self.ny = initial.nx, initial.ny
        h = C.c_void_p()
        _check(self.L.clv_adjoint_create(initial.h, float(dt), int(steps), int(checkpoint_every), C.byref(h)), "clv_adjoint_create")
        self.h = h
        # bumped by every set_initial / forward: a caller that holds on to a forward's
        # checkpoints (the autograd binding) can tell whether they are still the ones stored
        self.generation = 0

    def close(self):
        if getattr(self, "h", None):
            self.L.clv_adjoint_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def bytes(self) -> int:
        return int(self.L.clv_adjoint_bytes(self.h))

    def set_initial(self, density0, energy0, pitch=None):
        """density0 / energy0: (ny, nx) float64 numpy arrays, or device pointers (ints) with a row pitch."""
        self.generation += 1
        if isinstance(density0, np.ndarray):
            d = np.ascontiguousarray(density0, dtype=np.float64)
            e = np.ascontiguousarray(energy0, dtype=np.float64)
            assert d.shape == e.shape == (self.ny, self.nx)
            _check(self.L.clv_adjoint_set_initial(self.h, d.ctypes.data, e.ctypes.data, self.nx), "clv_adjoint_set_initial")
        else:
            _check(self.L.clv_adjoint_set_initial(self.h, int(density0), int(energy0), int(pitch)), "clv_adjoint_set_initial")

    def forward(self) -> list:
        out = (C.c_double * 5)()
        self.generation += 1
        _check(self.L.clv_adjoint_forward(self.h, out), "clv_adjoint_forward")
        return list(out)

    def backward(self, ke_bar: float = 1.0, grad_density0=None, grad_energy0=None, pitch=None):
        """Host numpy gradients by default; or device pointers (ints) with a row pitch.
        The backward pass consumes the forward run: another backward needs another forward."""
        self.generation += 1
        if grad_density0 is None and grad_energy0 is None:
            gd = np.empty((self.ny, self.nx))
            ge = np.empty((self.ny, self.nx))
            _check(self.L.clv_adjoint_backward(self.h, float(ke_bar), gd.ctypes.data, ge.ctypes.data, self.nx), "clv_adjoint_backward")
            return gd, ge
        _check(self.L.clv_adjoint_backward(self.h, float(ke_bar), int(grad_density0 or 0), int(grad_energy0 or 0), int(pitch)),
               "clv_adjoint_backward")
        return None


# ---- the adjoint workload -------------------------------------------------------------------
# The two-state problem of InputDecks/clover_bm_short.in on a 1920^2 grid, its
# initial density scaled by (1 + 1e-3 s(x, y)) with s a seeded sum of low sine
# modes (max |s| = 1), advanced 200 steps with a fixed dt; the gradient of the
# final kinetic energy with respect to the initial density and energy fields.
# dt = 0.003: a fixed 0.004 drives this problem unstable by step 5 at 1920^2
# (its adaptive dt runs 0.0019 ... 0.0031); 0.003 stays stable for 200 steps.
ADJOINT_WORKLOAD = dict(nx=1920, ny=1920, steps=200, dt=0.003, checkpoint_every=20, amplitude=1e-3, seed=4)


def smooth_field(nx: int, ny: int, seed: int, modes: int = 4) -> np.ndarray:
    """A seeded sum of low sine modes on the interior cells, (ny, nx), max |value| = 1."""
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    p = np.zeros((ny, nx))
    for _ in range(modes):
        kx, ky = rng.integers(1, 5, 2)
        p += rng.uniform(-1, 1) * np.outer(np.sin(np.pi * ky * y), np.sin(np.pi * kx * x))
    return p / np.abs(p).max()


def bm_deck(nx: int, ny: int) -> str:
    """InputDecks/clover_bm_short.in's two-state problem on an nx x ny grid."""
    return (f"*clover\n state 1 density=0.2 energy=1.0\n"
            f" state 2 density=1.0 energy=2.5 geometry=rectangle xmin=0.0 xmax=5.0 ymin=0.0 ymax=2.0\n"
            f" x_cells={nx}\n y_cells={ny}\n xmin=0.0\n ymin=0.0\n xmax=10.0\n ymax=10.0\n"
            f" initial_timestep=0.04\n timestep_rise=1.5\n max_timestep=0.04\n end_step=87\n*endclover\n")


def adjoint_initial(density0: np.ndarray, energy0: np.ndarray, amplitude: float, seed: int):
    """The workload's initial interior fields from the unperturbed ones."""
    ny, nx = density0.shape
    return density0 * (1.0 + amplitude * smooth_field(nx, ny, seed)), energy0.copy()
