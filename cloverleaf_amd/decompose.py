"""Decomposed domains on one GPU: tiles of a chunk, chunks of a mesh.

The reference splits a chunk into tiles (clover_tile_decompose,
driver/comms.cpp:158-253) whose ghosts are filled by copies from the
neighbouring tiles (driver/update_tile_halo.cpp:26-91), and a mesh into chunks,
one per MPI rank (clover_decompose, driver/comms.cpp:64-156), whose ghosts
travel as packed messages (clover_exchange, src/serial/comms_kernel.cpp:54-190).
Here every piece is a ``cloverleaf_amd.Chunk`` on the one GPU and every halo
operation is a library kernel:

- tiles: direct device copies (``clv_update_tile_halo``), the bottom/top copies
  of every tile first, then the left/right copies, the reference's order;
- chunks: each piece packs its face messages into device memory in the
  reference's layout (``clv_pack_fields``) and the neighbour unpacks them
  (``clv_unpack_fields``), left/right first, then bottom/top, as
  clover_exchange does.  Across GPUs the same messages would travel over
  xGMI / RCCL instead of being read in place;
- then the reflective halo on each piece's external faces only
  (``clv_update_halo_faces``, src/serial/update_halo.cpp:690-718).

The step is driver/hydro.cpp:57-89 with each kernel run per piece (the reference
plan), so each piece's kernels treat the piece's edge as the reference's do: in
particular advec_cell clamps its upwind index at the piece's own last face
(src/serial/advec_cell.cpp:77,169).  Where that clamp engages on a face shared
with another piece, the decomposed run differs from the undecomposed one (as the
reference's own tiled or multi-rank runs do), by round-off-sized amounts on
smooth flow; everywhere else the two are bit-identical.  ``summary`` adds the
pieces' field_summary sums.  Every piece must be on the null stream (the
default), which orders the exchanges between pieces.

Two defects of the reference's tile decomposition are not reproduced:
driver/comms.cpp:212-215 computes 1-based neighbour numbers and then indexes the
0-based tile vector with them, and driver/definitions.h:58 gives tile_bottom and
tile_top the same value, so the two neighbours overwrite each other.  With
tiles_per_chunk > 1 the reference binary crashes.  Here neighbours are 0-based
piece indices, with bottom and top distinct.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import (FACE_BOTTOM, FACE_LEFT, FACE_RIGHT, FACE_TOP, Chunk, Config, DeviceBuffer, HydroState, SmallTimestep,
               message_length)

XDIR, YDIR = 1, 2
G_BIG, G_SMALL = 1.0e21, 1.0e-16
ALL_HALO = ["density0", "density1", "energy0", "energy1", "pressure", "viscosity", "soundspeed", "xvel0", "xvel1",
            "yvel0", "yvel1", "vol_flux_x", "vol_flux_y", "mass_flux_x", "mass_flux_y"]


@dataclass
class Piece:
    """Cells [left, right] x [bottom, top] of the mesh (1-based, inclusive) and the
    index of the neighbouring piece across each face (left, right, bottom, top), -1
    where the face is external."""
    left: int
    right: int
    bottom: int
    top: int
    neighbours: tuple

    @property
    def nx(self) -> int:
        return self.right - self.left + 1

    @property
    def ny(self) -> int:
        return self.top - self.bottom + 1


def _split(nx: int, ny: int, px: int, py: int, left: int, bottom: int) -> list:
    """px x py pieces, the first nx % px columns and ny % py rows one cell wider
    (driver/comms.cpp:106-139 / :192-251), numbered x fastest."""
    dx, dy, mx, my = nx // px, ny // py, nx % px, ny % py
    pieces = []
    add_y_prev = 0
    for ty in range(1, py + 1):
        add_x_prev = 0
        add_y = 1 if ty <= my else 0
        for tx in range(1, px + 1):
            add_x = 1 if tx <= mx else 0
            l = left + (tx - 1) * dx + add_x_prev
            b = bottom + (ty - 1) * dy + add_y_prev
            me = px * (ty - 1) + (tx - 1)
            nb = (me - 1 if tx > 1 else -1, me + 1 if tx < px else -1, me - px if ty > 1 else -1, me + px if ty < py else -1)
            pieces.append(Piece(l, l + dx - 1 + add_x, b, b + dy - 1 + add_y, nb))
            if tx <= mx:
                add_x_prev += 1
        if ty <= my:
            add_y_prev += 1
    return pieces


def tile_decompose(x_cells: int, y_cells: int, tiles: int, left: int = 1, bottom: int = 1) -> list:
    """clover_tile_decompose (driver/comms.cpp:158-253): the factor split with the
    reference's integer mesh ratio and integer factor ratio."""
    ratio = int(x_cells / y_cells)
    tile_x, tile_y, found = tiles, 1, False
    for t in range(1, tiles + 1):
        if tiles % t == 0:
            factor_x, factor_y = int(tiles / t), t
            if factor_x // factor_y <= ratio:
                tile_x, tile_y, found = tiles // t, t, True
                break
    if not found or tile_y == tiles:
        tile_x, tile_y = (tiles, 1) if ratio >= 1.0 else (1, tiles)
    return _split(x_cells, y_cells, tile_x, tile_y, left, bottom)


def chunk_decompose(x_cells: int, y_cells: int, chunks: int) -> list:
    """clover_decompose (driver/comms.cpp:64-156): the factor split with the
    reference's floating-point mesh and factor ratios."""
    ratio = x_cells / y_cells
    chunk_x, chunk_y, found = chunks, 1, False
    for c in range(1, chunks + 1):
        if chunks % c == 0:
            if (chunks / c) / c <= ratio:
                chunk_x, chunk_y, found = chunks // c, c, True
                break
    if not found or chunk_y == chunks:
        chunk_x, chunk_y = (chunks, 1) if ratio >= 1.0 else (1, chunks)
    return _split(x_cells, y_cells, chunk_x, chunk_y, 1, 1)


class Decomposed:
    """One problem on a decomposed domain: mode "tiles" (copies between tiles) or
    "chunks" (packed messages between chunks)."""

    def __init__(self, cfg: Config, pieces: list, mode: str = "tiles"):
        if mode not in ("tiles", "chunks"):
            raise ValueError("mode is 'tiles' or 'chunks'")
        covered = sum(p.nx * p.ny for p in pieces)
        if covered != cfg.x_cells * cfg.y_cells:
            raise ValueError("the pieces do not cover the mesh")
        self.cfg, self.pieces, self.mode = cfg, pieces, mode
        self.chunks = [Chunk(p.nx, p.ny) for p in pieces]
        self.st = HydroState()
        self.send = []
        if mode == "chunks":
            # one outgoing message per face, sized for every field at depth 2
            for p in pieces:
                cells = {FACE_LEFT: p.ny, FACE_RIGHT: p.ny, FACE_BOTTOM: p.nx, FACE_TOP: p.nx}
                self.send.append({f: DeviceBuffer(message_length(ALL_HALO, 2, n)) for f, n in cells.items()
                                  if p.neighbours[f] >= 0})

    def close(self):
        for s in self.send:
            for b in s.values():
                b.close()
        for c in self.chunks:
            c.close()
        self.chunks, self.send = [], []

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- halos ----
    def update_halo(self, fields, depth: int):
        """update_halo (src/serial/update_halo.cpp:690-718): the exchange between
        pieces, then the reflective halo on each piece's external faces."""
        if self.mode == "tiles":
            for faces in ((FACE_TOP, FACE_BOTTOM), (FACE_LEFT, FACE_RIGHT)):
                for p, ch in zip(self.pieces, self.chunks):
                    for f in faces:
                        if p.neighbours[f] >= 0:
                            ch.update_tile_halo(self.chunks[p.neighbours[f]], f, fields, depth)
        else:
            opposite = {FACE_LEFT: FACE_RIGHT, FACE_RIGHT: FACE_LEFT, FACE_BOTTOM: FACE_TOP, FACE_TOP: FACE_BOTTOM}
            for faces in ((FACE_LEFT, FACE_RIGHT), (FACE_BOTTOM, FACE_TOP)):
                for p, ch, snd in zip(self.pieces, self.chunks, self.send):
                    for f in faces:
                        if p.neighbours[f] >= 0:
                            ch.pack_fields(f, fields, depth, p.ny if f <= FACE_RIGHT else p.nx, 0, snd[f])
                for p, ch in zip(self.pieces, self.chunks):
                    for f in faces:
                        n = p.neighbours[f]
                        if n >= 0:  # the message the neighbour sent towards this piece
                            msg = self.send[n][opposite[f]]
                            ch.pack_fields(f, fields, depth, p.ny if f <= FACE_RIGHT else p.nx, 0, msg, unpack=True)
        for p, ch in zip(self.pieces, self.chunks):
            ch.update_halo_faces(fields, depth, [n < 0 for n in p.neighbours])

    # ---- driver ----
    def start(self):
        """driver/start.cpp:69-127 on every piece."""
        for p, ch in zip(self.pieces, self.chunks):
            ch.build_field()
            ch.initialise_tile(self.cfg, p.left, p.bottom)
            ch.generate_chunk(self.cfg)
            ch.ideal_gas(False)
        self.update_halo(["density0", "energy0", "pressure", "viscosity", "density1", "energy1", "xvel0", "yvel0", "xvel1",
                          "yvel1"], 2)
        self.st = HydroState()
        self.st.advect_x = 1
        self.st.dt = self.st.dtold = self.cfg.dtinit
        return self.st

    def timestep(self):
        """driver/timestep.cpp:37-108: dt = the minimum over the pieces."""
        cfg, st = self.cfg, self.st
        for ch in self.chunks:
            ch.ideal_gas(False)
        self.update_halo(["pressure", "energy0", "density0", "xvel0", "yvel0"], 1)
        for ch in self.chunks:
            ch.viscosity()
        self.update_halo(["viscosity"], 1)
        st.dt = G_BIG
        for ch in self.chunks:
            dtlp = ch.calc_dt(cfg.dtmin, cfg.dtc_safe, cfg.dtu_safe, cfg.dtv_safe, cfg.dtdiv_safe)[0]
            if dtlp <= st.dt:
                st.dt = dtlp
        b = st.dtold * cfg.dtrise
        m = b if b < st.dt else st.dt
        st.dt = cfg.dtmax if cfg.dtmax < m else m
        st.dtold = st.dt
        if st.dt < cfg.dtmin:
            raise SmallTimestep("timestep: small timestep")

    def advection(self):
        """driver/advection.cpp:29-105."""
        direction = XDIR if self.st.advect_x else YDIR
        cell_fields = ["density1", "energy1", "xvel1", "yvel1", "mass_flux_x", "mass_flux_y"]
        self.update_halo(["energy1", "density1", "vol_flux_x", "vol_flux_y"], 2)
        for ch in self.chunks:
            ch.advec_cell(1, direction)
        self.update_halo(cell_fields, 2)
        for ch in self.chunks:
            ch.advec_mom(1, direction, 1)
            ch.advec_mom(2, direction, 1)
        direction = YDIR if self.st.advect_x else XDIR
        for ch in self.chunks:
            ch.advec_cell(2, direction)
        self.update_halo(cell_fields, 2)
        for ch in self.chunks:
            ch.advec_mom(1, direction, 2)
            ch.advec_mom(2, direction, 2)

    def step(self):
        """driver/hydro.cpp:57-89 with the PdV predictor chain of src/serial/PdV.cpp:113-132."""
        st = self.st
        st.step += 1
        self.timestep()
        dt = st.dt
        for ch in self.chunks:
            ch.PdV(True, dt)
        for ch in self.chunks:
            ch.ideal_gas(True)
        self.update_halo(["pressure"], 1)
        for ch in self.chunks:
            ch.revert()
        for ch in self.chunks:
            ch.accelerate(dt)
        for ch in self.chunks:
            ch.PdV(False, dt)
        for ch in self.chunks:
            ch.flux_calc(dt)
        self.advection()
        for ch in self.chunks:
            ch.reset_field()
        st.advect_x = 0 if st.advect_x else 1
        st.time += dt
        st.complete = 1 if (st.time + G_SMALL > self.cfg.end_time or st.step >= self.cfg.end_step) else 0
        return st

    def summary(self) -> list:
        """field_summary: the sum over pieces of each piece's sums."""
        tot = np.zeros(5)
        for ch in self.chunks:
            tot += np.array(ch.summary())
        return tot.tolist()

    def gather(self, name: str, node: bool = False) -> np.ndarray:
        """The mesh-wide interior of a field: cells (y_cells, x_cells), or with node=True
        nodes (y_cells + 1, x_cells + 1), assembled from the pieces."""
        e = 1 if node else 0
        out = np.full((self.cfg.y_cells + e, self.cfg.x_cells + e), np.nan)
        for p, ch in zip(self.pieces, self.chunks):
            a = ch.download(name)
            out[p.bottom - 1:p.top + e, p.left - 1:p.right + e] = a[2:p.ny + 2 + e, 2:p.nx + 2 + e]
        return out
