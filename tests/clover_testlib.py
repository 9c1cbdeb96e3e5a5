"""Test-side helpers: the deck reader, the CPU oracle (oracle/liboracle.so) and the
reference serial kernels (oracle/_ref/librefkern.so) behind ctypes.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, and only to check the GPU library's results against or to time a CPU run beside it.

The deck reader restates driver/read_input.cpp:37-277: defaults (:39-69),
``keyword=value`` / ``keyword value`` tokens split on spaces and '=' (:135),
states (:193-254) and the 1/100-cell shrink of every non-background state
(:262-276, with the reference's (float) cast of the cell count).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent.parent
ORACLE_DIR = REPO / "oracle"
GOLDEN = REPO / "tests" / "golden"
DECKS = GOLDEN / "decks"

# 2D field ids (oracle/clover_oracle.h; 0..14 = driver/definitions.h:62-78)
FIELDS = ["density0", "density1", "energy0", "energy1", "pressure", "viscosity", "soundspeed",
          "xvel0", "xvel1", "yvel0", "yvel1", "vol_flux_x", "vol_flux_y", "mass_flux_x", "mass_flux_y",
          "work1", "work2", "work3", "work4", "work5", "work6", "work7", "volume", "xarea", "yarea"]
FID = {n: i for i, n in enumerate(FIELDS)}
MESH = ["cellx", "celly", "vertexx", "vertexy", "celldx", "celldy", "vertexdx", "vertexdy"]
MID = {n: i for i, n in enumerate(MESH)}
# storage extent (x, y) of each 2D field relative to (nx, ny) (driver/definitions.cpp:4-41)
CELL, NODE, XFACE, YFACE = (4, 4), (5, 5), (5, 4), (4, 5)
EXTENT = {**{n: CELL for n in ["density0", "density1", "energy0", "energy1", "pressure", "viscosity",
                               "soundspeed", "volume"]},
          **{n: NODE for n in ["xvel0", "xvel1", "yvel0", "yvel1"] + [f"work{i}" for i in range(1, 8)]},
          "vol_flux_x": XFACE, "mass_flux_x": XFACE, "xarea": XFACE,
          "vol_flux_y": YFACE, "mass_flux_y": YFACE, "yarea": YFACE}
G_SMALL, G_BIG = 1.0e-16, 1.0e21


# --------------------------------------------------------------------------------------------
# deck reader
@dataclass
class State:
    density: float = 0.0
    energy: float = 0.0
    xvel: float = 0.0
    yvel: float = 0.0
    geometry: int = 0
    xmin: float = 0.0
    ymin: float = 0.0
    xmax: float = 0.0
    ymax: float = 0.0
    radius: float = 0.0
    defined: bool = False


@dataclass
class Deck:
    xmin: float = 0.0
    ymin: float = 0.0
    xmax: float = 0.0
    ymax: float = 0.0
    x_cells: int = 10
    y_cells: int = 10
    end_time: float = 10.0
    end_step: int = 640000
    visit_frequency: int = 0
    summary_frequency: int = 10
    tiles_per_chunk: int = 1
    dtinit: float = 0.1
    dtmax: float = 1.0
    dtmin: float = 0.0000001
    dtrise: float = 1.5
    dtc_safe: float = 0.7
    dtu_safe: float = 0.5
    dtv_safe: float = 0.5
    dtdiv_safe: float = 0.7
    profiler_on: bool = False
    test_problem: int = 0
    states: list = field(default_factory=list)


def _f32(x: int) -> float:
    return float(np.float32(x))


def parse_deck(text: str) -> Deck:
    d = Deck()
    lines = text.splitlines()
    nstates = 0
    for line in lines:
        w = line.split()
        if w and w[0] == "state":
            nstates = max(nstates, int(w[1]))
    if nstates < 1:
        raise ValueError("read_input: No states defined.")
    d.states = [State() for _ in range(nstates)]
    keys_d = {"initial_timestep": "dtinit", "max_timestep": "dtmax", "timestep_rise": "dtrise",
              "end_time": "end_time", "xmin": "xmin", "xmax": "xmax", "ymin": "ymin", "ymax": "ymax"}
    keys_i = {"end_step": "end_step", "x_cells": "x_cells", "y_cells": "y_cells",
              "visit_frequency": "visit_frequency", "summary_frequency": "summary_frequency",
              "tiles_per_chunk": "tiles_per_chunk", "test_problem": "test_problem"}
    for line in lines:
        if not line:
            continue
        words = [t for t in line.replace("=", " ").split(" ") if t]
        if not words:
            break  # reference: strtok yields nothing -> words[0] empty -> stop (read_input.cpp:140)
        k = words[0]
        if k in keys_d:
            setattr(d, keys_d[k], float(words[1]))
        elif k in keys_i:
            setattr(d, keys_i[k], int(words[1]))
        elif k == "tiles_per_problem":
            d.tiles_per_chunk = int(words[1])
        elif k == "profiler_on":
            d.profiler_on = True
        elif k == "state":
            s = d.states[int(words[1]) - 1]
            if s.defined:
                raise ValueError("read_input: State defined twice.")
            s.defined = True
            it = iter(words[2:])
            for w in it:
                if w in ("xvel", "yvel", "xmin", "ymin", "xmax", "ymax", "radius", "density", "energy"):
                    setattr(s, w, float(next(it)))
                elif w == "geometry":
                    g = next(it)
                    s.geometry = {"rectangle": 1, "circle": 2, "point": 3}.get(g, s.geometry)
    dx = (d.xmax - d.xmin) / _f32(d.x_cells)
    dy = (d.ymax - d.ymin) / _f32(d.y_cells)
    for s in d.states[1:]:
        s.xmin += dx / 100.0
        s.ymin += dy / 100.0
        s.xmax -= dx / 100.0
        s.ymax -= dy / 100.0
    return d


def load_deck(name_or_path) -> Deck:
    p = Path(name_or_path)
    if not p.exists():
        p = DECKS / (str(name_or_path) + ("" if str(name_or_path).endswith(".in") else ".in"))
    return parse_deck(p.read_text())


# --------------------------------------------------------------------------------------------
# oracle (ctypes)
class _OrcChunk(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("px", C.c_int),
                ("f", C.POINTER(C.c_double) * 25), ("m", C.POINTER(C.c_double) * 8)]


class _OrcState(C.Structure):
    _fields_ = [("density", C.c_double), ("energy", C.c_double), ("xvel", C.c_double), ("yvel", C.c_double),
                ("geometry", C.c_int), ("xmin", C.c_double), ("ymin", C.c_double), ("xmax", C.c_double),
                ("ymax", C.c_double), ("radius", C.c_double)]


class _OrcConfig(C.Structure):
    _fields_ = [("xmin", C.c_double), ("ymin", C.c_double), ("xmax", C.c_double), ("ymax", C.c_double),
                ("x_cells", C.c_int), ("y_cells", C.c_int),
                ("dtinit", C.c_double), ("dtmax", C.c_double), ("dtmin", C.c_double), ("dtrise", C.c_double),
                ("dtc_safe", C.c_double), ("dtu_safe", C.c_double), ("dtv_safe", C.c_double),
                ("dtdiv_safe", C.c_double), ("end_time", C.c_double), ("end_step", C.c_int),
                ("summary_frequency", C.c_int), ("number_of_states", C.c_int),
                ("states", C.POINTER(_OrcState))]


class _OrcGlobals(C.Structure):
    _fields_ = [("step", C.c_int), ("advect_x", C.c_int), ("time", C.c_double), ("dt", C.c_double),
                ("dtold", C.c_double)]


_ORACLE = None


def oracle_lib():
    global _ORACLE
    if _ORACLE is None:
        so = ORACLE_DIR / "liboracle.so"
        if not so.exists():
            subprocess.run(["make", "-s", "-C", str(ORACLE_DIR), "all"], check=True)
        _ORACLE = C.CDLL(str(so))
        L = _ORACLE
        P = C.POINTER
        L.orc_chunk_alloc.argtypes = [P(_OrcChunk), C.c_int, C.c_int]
        L.orc_chunk_free.argtypes = [P(_OrcChunk)]
        L.orc_start.argtypes = [P(_OrcChunk), P(_OrcConfig), P(_OrcGlobals), P(C.c_double)]
        L.orc_hydro_step.argtypes = [P(_OrcChunk), P(_OrcConfig), P(_OrcGlobals)]
        L.orc_hydro_step.restype = C.c_int
        L.orc_timestep.argtypes = [P(_OrcChunk), P(_OrcConfig), P(_OrcGlobals)]
        L.orc_timestep.restype = C.c_int
        L.orc_field_summary.argtypes = [P(_OrcChunk), P(C.c_double)]
        L.orc_field_summary_kernel.argtypes = [P(_OrcChunk), P(C.c_double)]
        L.orc_ideal_gas.argtypes = [P(_OrcChunk), C.c_int]
        L.orc_viscosity.argtypes = [P(_OrcChunk)]
        L.orc_calc_dt.argtypes = [P(_OrcChunk)] + [C.c_double] * 5 + [P(C.c_double)] + [P(C.c_int)] * 4
        L.orc_PdV.argtypes = [P(_OrcChunk), C.c_int, C.c_double]
        L.orc_PdV_driver.argtypes = [P(_OrcChunk), C.c_int, C.c_double]
        L.orc_revert.argtypes = [P(_OrcChunk)]
        L.orc_accelerate.argtypes = [P(_OrcChunk), C.c_double]
        L.orc_flux_calc.argtypes = [P(_OrcChunk), C.c_double]
        L.orc_advec_cell.argtypes = [P(_OrcChunk), C.c_int, C.c_int]
        L.orc_advec_mom.argtypes = [P(_OrcChunk), C.c_int, C.c_int, C.c_int]
        L.orc_advection.argtypes = [P(_OrcChunk), C.c_int]
        L.orc_reset_field.argtypes = [P(_OrcChunk)]
        L.orc_update_halo.argtypes = [P(_OrcChunk), P(C.c_int), C.c_int]
        L.orc_build_field.argtypes = [P(_OrcChunk)]
        L.orc_initialise_chunk.argtypes = [P(_OrcChunk), P(_OrcConfig)]
        L.orc_generate_chunk.argtypes = [P(_OrcChunk), P(_OrcConfig)]
    return _ORACLE


def mask(names) -> "C.Array":
    m = (C.c_int * 15)()
    for n in names:
        m[FID[n]] = 1
    return m


class Oracle:
    """One chunk of the CPU oracle; arrays are exposed as (ny+5, nx+5) numpy views."""

    def __init__(self, deck: Deck):
        self.L = oracle_lib()
        self.deck = deck
        self.nx, self.ny = deck.x_cells, deck.y_cells
        self.c = _OrcChunk()
        if self.L.orc_chunk_alloc(C.byref(self.c), self.nx, self.ny) != 0:
            raise MemoryError("oracle allocation failed")
        self._states = (_OrcState * len(deck.states))()
        for i, s in enumerate(deck.states):
            self._states[i] = _OrcState(s.density, s.energy, s.xvel, s.yvel, s.geometry, s.xmin, s.ymin,
                                        s.xmax, s.ymax, s.radius)
        self.cfg = _OrcConfig(deck.xmin, deck.ymin, deck.xmax, deck.ymax, deck.x_cells, deck.y_cells,
                              deck.dtinit, deck.dtmax, deck.dtmin, deck.dtrise, deck.dtc_safe, deck.dtu_safe,
                              deck.dtv_safe, deck.dtdiv_safe, deck.end_time, deck.end_step,
                              deck.summary_frequency, len(deck.states), self._states)
        self.g = _OrcGlobals()

    def __del__(self):
        try:
            self.L.orc_chunk_free(C.byref(self.c))
        except Exception:
            pass

    def f(self, name) -> np.ndarray:
        n = (self.ny + 5) * (self.nx + 5)
        return np.ctypeslib.as_array(self.c.f[FID[name]], shape=(n,)).reshape(self.ny + 5, self.nx + 5)

    def m(self, name) -> np.ndarray:
        n = self.nx + 5 if name in ("cellx", "vertexx", "celldx", "vertexdx") else self.ny + 5
        return np.ctypeslib.as_array(self.c.m[MID[name]], shape=(n,))

    def start(self):
        out = (C.c_double * 5)()
        self.L.orc_start(C.byref(self.c), C.byref(self.cfg), C.byref(self.g), out)
        return list(out)

    def step(self) -> int:
        return self.L.orc_hydro_step(C.byref(self.c), C.byref(self.cfg), C.byref(self.g))

    def summary(self):
        out = (C.c_double * 5)()
        self.L.orc_field_summary(C.byref(self.c), out)
        return list(out)

    def run(self, trace=True):
        """driver/hydro.cpp:53-111: returns the list of summary records (every step when trace)."""
        recs = []
        s = self.start()
        recs.append(dict(step=0, time=self.g.time, dt=self.g.dt, vol=s[0], mass=s[1], ie=s[2], ke=s[3], press=s[4]))
        while True:
            if self.step():
                raise RuntimeError("timestep: small timestep")
            done = self.g.time + G_SMALL > self.deck.end_time or self.g.step >= self.deck.end_step
            if trace or done or (self.deck.summary_frequency and self.g.step % self.deck.summary_frequency == 0):
                s = self.summary()
                recs.append(dict(step=self.g.step, time=self.g.time, dt=self.g.dt, vol=s[0], mass=s[1], ie=s[2],
                                 ke=s[3], press=s[4]))
            if done:
                return recs


# --------------------------------------------------------------------------------------------
# the reference serial kernels (present only where oracle/_ref was built from /root/reference)
_REFK = None


def refkern_lib():
    global _REFK
    if _REFK is None:
        so = ORACLE_DIR / "_ref" / "librefkern.so"
        if not so.exists():
            return None
        _REFK = C.CDLL(str(so))
        _REFK.refk_run.argtypes = [C.c_char_p, C.c_int, C.c_int, C.POINTER(C.POINTER(C.c_double)),
                                   C.POINTER(C.POINTER(C.c_double)), C.POINTER(C.c_int),
                                   C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int)]
        _REFK.refk_run.restype = C.c_int
        _REFK.refk_tile_halo.restype = C.c_int
        _REFK.refk_pack.restype = C.c_int
        _REFK.refk_chunk_decompose.restype = C.c_int
        _REFK.refk_tile_decompose.restype = C.c_int
    return _REFK


def refk_run(kernel: str, nx: int, ny: int, f2d: dict, m1d: dict, ia=(), da=()):
    """Run one reference serial kernel on copies of f2d/m1d (arrays shaped (ny+5, nx+5) / 1D).
    Returns (f2d_out, dout[4], iout[4])."""
    L = refkern_lib()
    f_out = {k: np.ascontiguousarray(v, dtype=np.float64).copy() for k, v in f2d.items()}
    m_in = {k: np.ascontiguousarray(v, dtype=np.float64).copy() for k, v in m1d.items()}
    fp = (C.POINTER(C.c_double) * 25)(*[f_out[n].ctypes.data_as(C.POINTER(C.c_double)) for n in FIELDS])
    mp = (C.POINTER(C.c_double) * 8)(*[m_in[n].ctypes.data_as(C.POINTER(C.c_double)) for n in MESH])
    iav = (C.c_int * max(1, len(ia)))(*ia)
    dav = (C.c_double * max(1, len(da)))(*da)
    dout = (C.c_double * 4)()
    iout = (C.c_int * 4)()
    rc = L.refk_run(kernel.encode(), nx, ny, fp, mp, iav, dav, dout, iout)
    if rc != 0:
        raise ValueError(f"unknown reference kernel {kernel}")
    return f_out, list(dout), list(iout)


def refk_tile_halo(face: int, nx: int, ny: int, f2d: dict, m1d: dict, nnx: int, nny: int, nf2d: dict, nm1d: dict,
                   fields, depth: int) -> dict:
    """The reference update_tile_halo_{l,r,b,t}_kernel (face 0 left, 1 right, 2 bottom,
    3 top) on a copy of tile f2d from neighbour nf2d; returns the tile's arrays."""
    L = refkern_lib()
    f_out = {k: np.ascontiguousarray(v, dtype=np.float64).copy() for k, v in f2d.items()}
    keep = [f_out, {k: np.ascontiguousarray(v, dtype=np.float64).copy() for k, v in m1d.items()},
            {k: np.ascontiguousarray(v, dtype=np.float64).copy() for k, v in nf2d.items()},
            {k: np.ascontiguousarray(v, dtype=np.float64).copy() for k, v in nm1d.items()}]
    P = C.POINTER(C.c_double)
    fp = (P * 25)(*[keep[0][n].ctypes.data_as(P) for n in FIELDS])
    mp = (P * 8)(*[keep[1][n].ctypes.data_as(P) for n in MESH])
    nfp = (P * 25)(*[keep[2][n].ctypes.data_as(P) for n in FIELDS])
    nmp = (P * 8)(*[keep[3][n].ctypes.data_as(P) for n in MESH])
    rc = L.refk_tile_halo(face, nx, ny, fp, mp, nnx, nny, nfp, nmp, mask(fields), depth)
    if rc != 0:
        raise ValueError("bad face")
    return f_out


def refk_pack(face: int, unpack: int, nx: int, ny: int, field: np.ndarray, field_type: int, depth: int, buf: np.ndarray,
              offset: int):
    """The reference clover_pack_message_* / clover_unpack_message_* on copies of one
    (ny+5, nx+5) array and a message buffer; returns (array, buffer)."""
    L = refkern_lib()
    a = np.ascontiguousarray(field, dtype=np.float64).copy()
    b = np.ascontiguousarray(buf, dtype=np.float64).copy()
    P = C.POINTER(C.c_double)
    rc = L.refk_pack(face, unpack, nx, ny, a.ctypes.data_as(P), field_type, depth, b.ctypes.data_as(P), len(b), offset)
    if rc != 0:
        raise ValueError("bad face")
    return a, b


def random_chunk_arrays(nx, ny, seed, positive=("density0", "density1", "energy0", "energy1", "volume",
                                                 "xarea", "yarea", "pressure", "soundspeed")):
    """Random but physically shaped inputs for per-kernel checks: positive densities,
    energies, volumes and areas; signed velocities and fluxes of mixed sign so every
    upwind / limiter / viscosity branch is taken; a nonuniform mesh spacing."""
    rng = np.random.default_rng(seed)
    shape = (ny + 5, nx + 5)
    f = {}
    for n in FIELDS:
        if n in positive:
            f[n] = rng.uniform(0.5, 2.0, shape)
        elif n in ("viscosity",):
            f[n] = np.where(rng.random(shape) < 0.5, 0.0, rng.uniform(0.0, 0.3, shape))
        else:
            f[n] = rng.uniform(-0.2, 0.2, shape)
    # make roughly half the cells smooth so "diffuw*diffdw > 0" branches occur
    for n in ("density1", "energy1", "xvel1", "yvel1"):
        ramp = np.add.outer(np.linspace(0, 1, ny + 5), np.linspace(0, 1, nx + 5))
        f[n] = np.where(rng.random(shape) < 0.5, 1.0 + ramp, f[n])
    m = {}
    for n in MESH:
        ln = nx + 5 if n in ("cellx", "vertexx", "celldx", "vertexdx") else ny + 5
        if n.startswith("cell") and n.endswith(("dx", "dy")) or n.startswith("vertexd"):
            m[n] = rng.uniform(0.05, 0.15, ln)
        else:
            m[n] = np.cumsum(rng.uniform(0.05, 0.15, ln))
    return f, m


def golden(name: str):
    import json
    return json.loads((GOLDEN / f"{name}.json").read_text())


# --------------------------------------------------------------------------------------------
# the oracle's tangent-linear model (orc_tl_*): forward-mode dual numbers
class _OrcDual(C.Structure):
    _fields_ = [("v", C.c_double), ("d", C.c_double)]


class _OrcTLChunk(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("px", C.c_int),
                ("f", C.POINTER(_OrcDual) * 25), ("m", C.POINTER(C.c_double) * 8)]


_TL_SIGS = {
    "ideal_gas": [C.c_int], "viscosity": [], "PdV": [C.c_int, C.c_double], "revert": [],
    "accelerate": [C.c_double], "flux_calc": [C.c_double], "advec_cell": [C.c_int, C.c_int],
    "advec_mom": [C.c_int, C.c_int, C.c_int], "reset_field": [], "PdV_driver": [C.c_int, C.c_double],
    "advection": [C.c_int], "step_fixed": [C.c_double, C.c_int],
}


def _tl_lib():
    L = oracle_lib()
    if not getattr(L, "_tl_ready", False):
        P = C.POINTER
        L.orc_tl_alloc.argtypes = [P(_OrcTLChunk), C.c_int, C.c_int]
        L.orc_tl_free.argtypes = [P(_OrcTLChunk)]
        L.orc_tl_load.argtypes = [P(_OrcTLChunk), P(_OrcChunk)]
        L.orc_tl_update_halo.argtypes = [P(_OrcTLChunk), P(C.c_int), C.c_int]
        L.orc_tl_field_summary.argtypes = [P(_OrcTLChunk), P(_OrcDual)]
        L.orc_tl_field_summary_kernel.argtypes = [P(_OrcTLChunk), P(_OrcDual)]
        L.orc_step_fixed.argtypes = [P(_OrcChunk), C.c_double, C.c_int]
        L.orc_tl_set_kink_eps.argtypes = [C.c_double]
        L.orc_tl_kinks.restype = C.c_longlong
        for k, args in _TL_SIGS.items():
            getattr(L, f"orc_tl_{k}").argtypes = [P(_OrcTLChunk)] + args
        L._tl_ready = True
    return L


class TLOracle:
    """A chunk of the oracle's tangent-linear model.  Array `name` is exposed as a
    (ny+5, nx+5, 2) view: [..., 0] the value, [..., 1] the directional derivative."""

    def __init__(self, nx, ny):
        self.L = _tl_lib()
        self.nx, self.ny = nx, ny
        self.c = _OrcTLChunk()
        if self.L.orc_tl_alloc(C.byref(self.c), nx, ny) != 0:
            raise MemoryError("tangent-linear oracle allocation failed")

    @classmethod
    def from_oracle(cls, o: "Oracle") -> "TLOracle":
        t = cls(o.nx, o.ny)
        t.L.orc_tl_load(C.byref(t.c), C.byref(o.c))
        return t

    def __del__(self):
        try:
            self.L.orc_tl_free(C.byref(self.c))
        except Exception:
            pass

    def a(self, name) -> np.ndarray:
        n = (self.ny + 5) * (self.nx + 5)
        return np.ctypeslib.as_array(C.cast(self.c.f[FID[name]], C.POINTER(C.c_double)),
                                     shape=(2 * n,)).reshape(self.ny + 5, self.nx + 5, 2)

    def val(self, name) -> np.ndarray:
        return self.a(name)[..., 0]

    def der(self, name) -> np.ndarray:
        return self.a(name)[..., 1]

    def m(self, name) -> np.ndarray:
        n = self.nx + 5 if name in ("cellx", "vertexx", "celldx", "vertexdx") else self.ny + 5
        return np.ctypeslib.as_array(self.c.m[MID[name]], shape=(n,))

    def run(self, kernel, *args):
        getattr(self.L, f"orc_tl_{kernel}")(C.byref(self.c), *args)

    def update_halo(self, names, depth):
        self.L.orc_tl_update_halo(C.byref(self.c), mask(names), depth)

    def summary(self, with_ideal_gas=True):
        out = (_OrcDual * 5)()
        fn = self.L.orc_tl_field_summary if with_ideal_gas else self.L.orc_tl_field_summary_kernel
        fn(C.byref(self.c), out)
        return [(x.v, x.d) for x in out]


def oracle_on_arrays(nx, ny, f2d, m1d) -> "Oracle":
    """An oracle chunk holding the given arrays (for per-kernel checks)."""
    o = Oracle(Deck(x_cells=nx, y_cells=ny, states=[State(density=1, energy=1)]))
    for n in FIELDS:
        o.f(n)[...] = f2d[n]
    for n in MESH:
        o.m(n)[...] = m1d[n]
    return o


def oracle_step_fixed(o: "Oracle", dt: float, advect_x: int):
    _tl_lib().orc_step_fixed(C.byref(o.c), dt, advect_x)


class DecomposedOracle:
    """The oracle on a decomposed domain: one oracle chunk per piece, the reference
    plan's step (driver/hydro.cpp:57-89) kernel by kernel on every piece, and
    update_halo as the tile copies (every piece's bottom/top copies, then
    left/right, driver/update_tile_halo.cpp:26-91) followed by the reflective
    halo on each piece's external faces (src/serial/update_halo.cpp:690-718).
    Each piece's kernels clamp at the piece's own edge as the reference's do
    (e.g. advec_cell's upwind index, src/serial/advec_cell.cpp:77,169), which is
    why a decomposed run may differ slightly from the undecomposed one."""

    def __init__(self, deck: "Deck", pieces, arrays):
        """arrays[k] = (fields, mesh) of piece k: (ny+5, nx+5) / 1D numpy arrays."""
        self.deck, self.pieces = deck, pieces
        self.o = [oracle_on_arrays(p.nx, p.ny, f, m) for p, (f, m) in zip(pieces, arrays)]
        self.L = self.o[0].L
        self.dt = self.dtold = deck.dtinit
        self.advect_x = 1

    def update_halo(self, names, depth):
        mk = mask(names)
        for faces in ((3, 2), (0, 1)):
            for p, o in zip(self.pieces, self.o):
                for fc in faces:
                    if p.neighbours[fc] >= 0:
                        self.L.orc_tile_halo(C.byref(o.c), C.byref(self.o[p.neighbours[fc]].c), fc, mk, depth)
        for p, o in zip(self.pieces, self.o):
            ext = (C.c_int * 4)(*[1 if n < 0 else 0 for n in p.neighbours])
            self.L.orc_update_halo_ext(C.byref(o.c), mk, depth, ext)

    def _each(self, fn):
        for o in self.o:
            fn(C.byref(o.c))

    def step(self):
        L, d = self.L, self.deck
        self._each(lambda c: L.orc_ideal_gas(c, 0))
        self.update_halo(["pressure", "energy0", "density0", "xvel0", "yvel0"], 1)
        self._each(lambda c: L.orc_viscosity(c))
        self.update_halo(["viscosity"], 1)
        dt = G_BIG
        for o in self.o:
            v = C.c_double()
            ints = [C.c_int() for _ in range(4)]
            L.orc_calc_dt(C.byref(o.c), d.dtmin, d.dtc_safe, d.dtu_safe, d.dtv_safe, d.dtdiv_safe, C.byref(v),
                          *[C.byref(x) for x in ints])
            if v.value <= dt:
                dt = v.value
        b = self.dtold * d.dtrise
        m = b if b < dt else dt
        dt = d.dtmax if d.dtmax < m else m
        self.dt = self.dtold = dt
        self._each(lambda c: L.orc_PdV(c, 1, dt))
        self._each(lambda c: L.orc_ideal_gas(c, 1))
        self.update_halo(["pressure"], 1)
        self._each(lambda c: L.orc_revert(c))
        self._each(lambda c: L.orc_accelerate(c, dt))
        self._each(lambda c: L.orc_PdV(c, 0, dt))
        self._each(lambda c: L.orc_flux_calc(c, dt))
        d1, d2 = (1, 2) if self.advect_x else (2, 1)
        cells = ["density1", "energy1", "xvel1", "yvel1", "mass_flux_x", "mass_flux_y"]
        self.update_halo(["energy1", "density1", "vol_flux_x", "vol_flux_y"], 2)
        self._each(lambda c: L.orc_advec_cell(c, 1, d1))
        self.update_halo(cells, 2)
        self._each(lambda c: (L.orc_advec_mom(c, 1, 1, d1), L.orc_advec_mom(c, 2, 1, d1)))
        self._each(lambda c: L.orc_advec_cell(c, 2, d2))
        self.update_halo(cells, 2)
        self._each(lambda c: (L.orc_advec_mom(c, 1, 2, d2), L.orc_advec_mom(c, 2, 2, d2)))
        self._each(lambda c: L.orc_reset_field(c))
        self.advect_x = 1 - self.advect_x
        return dt
