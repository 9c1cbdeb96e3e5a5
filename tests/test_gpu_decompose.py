"""Decomposed domains on the GPU, through the C ABI: clv_update_tile_halo,
clv_update_halo_faces, clv_pack_* / clv_unpack_*, and cloverleaf_amd.decompose.

Pinned by:
- each kernel bit-identical to the oracle on random arrays, ghosts included (the
  oracle is itself bit-identical to the reference's update_tile_halo_kernel.cpp,
  pack_kernel.cpp and update_halo.cpp with non-external faces:
  tests/test_oracle_vs_ref.py);
- whole runs: the two-state problem split into tiles (copies between tiles) or
  chunks (packed messages between chunks) equals the oracle run on the same
  pieces bit for bit (dt of every step, every piece's fields); tile copies and
  packed messages give the same run bit for bit; against the undecomposed run
  the initial state is bit-identical and 15 steps later the fields differ only
  where the reference's kernels clamp at a piece's own edge (advec_cell's upwind
  index, src/serial/advec_cell.cpp:77,169), by round-off-sized amounts.
"""
import ctypes as C
import zlib

import numpy as np
import pytest

import clover_testlib as T
import cloverleaf_amd as clv
from cloverleaf_amd.decompose import Decomposed, chunk_decompose, tile_decompose

pytestmark = pytest.mark.gpu

HALO = T.FIELDS[:15]
EXT_PATTERNS = [(1, 1, 1, 1), (0, 1, 1, 1), (1, 0, 1, 1), (1, 1, 0, 1), (1, 1, 1, 0), (0, 0, 1, 1), (1, 1, 0, 0), (0, 1, 0, 1),
                (0, 0, 0, 0)]


def _chunk(nx, ny, f, m):
    ch = clv.Chunk(nx, ny)
    for n in T.FIELDS:
        ch.upload(n, f[n])
    for n in T.MESH:
        ch.upload_mesh(n, m[n])
    return ch


def _same(ch, want: dict, what):
    for n in T.FIELDS:
        got = ch.download(n)
        assert np.array_equal(got, want[n]), f"{what}: {n} differs at {np.argwhere(got != want[n])[:4].tolist()}"


@pytest.mark.parametrize("nx,ny", [(8, 8), (13, 7), (3, 5)])
@pytest.mark.parametrize("depth", [1, 2])
@pytest.mark.parametrize("ext", EXT_PATTERNS)
def test_update_halo_faces_vs_oracle(gpu, nx, ny, depth, ext):
    f, m = T.random_chunk_arrays(nx, ny, seed=zlib.crc32(repr((nx, ny, depth, ext)).encode()))
    with _chunk(nx, ny, f, m) as ch:
        ch.update_halo_faces(HALO, depth, ext)
        o = T.oracle_on_arrays(nx, ny, f, m)
        o.L.orc_update_halo_ext(C.byref(o.c), T.mask(HALO), depth, (C.c_int * 4)(*ext))
        _same(ch, {n: o.f(n) for n in T.FIELDS}, f"update_halo_faces {ext}")


@pytest.mark.parametrize("face", [0, 1, 2, 3])
@pytest.mark.parametrize("depth", [1, 2])
@pytest.mark.parametrize("fields", [HALO, ["density0", "xvel1", "vol_flux_y", "mass_flux_x"]])
def test_tile_halo_vs_oracle(gpu, face, depth, fields):
    nx, ny = 11, 9
    nnx, nny = (7, ny) if face <= 1 else (nx, 6)
    seed = zlib.crc32(repr((face, depth, len(fields))).encode())
    f, m = T.random_chunk_arrays(nx, ny, seed=seed)
    nf, nm = T.random_chunk_arrays(nnx, nny, seed=seed + 1)
    with _chunk(nx, ny, f, m) as ch, _chunk(nnx, nny, nf, nm) as nb:
        ch.update_tile_halo(nb, face, fields, depth)
        o, on = T.oracle_on_arrays(nx, ny, f, m), T.oracle_on_arrays(nnx, nny, nf, nm)
        o.L.orc_tile_halo(C.byref(o.c), C.byref(on.c), face, T.mask(fields), depth)
        _same(ch, {n: o.f(n) for n in T.FIELDS}, "update_tile_halo")
        _same(nb, nf, "neighbour untouched")


@pytest.mark.parametrize("face", [0, 1, 2, 3])
@pytest.mark.parametrize("unpack", [False, True])
@pytest.mark.parametrize("field_type", [1, 2, 3, 4])
@pytest.mark.parametrize("depth", [1, 2])
def test_pack_message_vs_oracle(gpu, face, unpack, field_type, depth):
    nx, ny = 10, 7
    f, m = T.random_chunk_arrays(nx, ny, seed=zlib.crc32(repr((face, unpack, field_type, depth)).encode()))
    rng = np.random.default_rng(5)
    offset = 3 * depth * ((ny if face <= 1 else nx) + 5)
    buf0 = rng.uniform(-1, 1, offset + 2 * depth * (max(nx, ny) + 5))
    o = T.oracle_on_arrays(nx, ny, f, m)
    want_b = buf0.copy()
    P = C.POINTER(C.c_double)
    a = o.f("energy1")
    fn = o.L.orc_unpack_message if unpack else o.L.orc_pack_message
    fn(C.byref(o.c), a.ctypes.data_as(P), field_type, face, depth, want_b.ctypes.data_as(P), C.c_long(offset))
    with _chunk(nx, ny, f, m) as ch:
        buf = clv.DeviceBuffer(len(buf0))
        buf.upload(buf0)
        ch.pack_message(ch.field_ptr("energy1"), field_type, face, depth, buf, offset, unpack=unpack)
        ch.synchronize()
        assert np.array_equal(buf.download(), want_b)
        assert np.array_equal(ch.download("energy1"), a)
        buf.close()


@pytest.mark.parametrize("face", [0, 1, 2, 3])
@pytest.mark.parametrize("depth", [1, 2])
def test_pack_fields_layout_and_round_trip(gpu, face, depth):
    """clover_pack_<face> of a tile inside a wider chunk face: the k-th requested field
    at k * depth * (chunk_cells + 5) + tile_offset, each as the single-array message;
    unpacking the message into a second chunk writes exactly those ghost layers."""
    nx, ny = 12, 9
    cells = ny if face <= 1 else nx
    chunk_cells, tile_offset = cells + 6, 4 * depth
    fields = ["density0", "xvel0", "vol_flux_x", "mass_flux_y", "soundspeed"]
    f, m = T.random_chunk_arrays(nx, ny, seed=face * 10 + depth)
    n = clv.message_length(fields, depth, chunk_cells)
    assert n == len(fields) * depth * (chunk_cells + 5)
    o = T.oracle_on_arrays(nx, ny, f, m)
    want = np.zeros(n)
    P = C.POINTER(C.c_double)
    ftype = {"density0": 1, "soundspeed": 1, "xvel0": 2, "vol_flux_x": 3, "mass_flux_y": 4}
    for k, name in enumerate(sorted(fields, key=T.FIELDS.index)):
        o.L.orc_pack_message(C.byref(o.c), o.f(name).ctypes.data_as(P), ftype[name], face, depth, want.ctypes.data_as(P),
                             C.c_long(k * depth * (chunk_cells + 5) + tile_offset))
    g, gm = T.random_chunk_arrays(nx, ny, seed=99)
    with _chunk(nx, ny, f, m) as ch, _chunk(nx, ny, g, gm) as other:
        buf = clv.DeviceBuffer(n)
        buf.upload(np.zeros(n))
        ch.pack_fields(face, fields, depth, chunk_cells, tile_offset, buf)
        ch.synchronize()
        assert np.array_equal(buf.download(), want)
        other.pack_fields(face, fields, depth, chunk_cells, tile_offset, buf, unpack=True)
        ot = T.oracle_on_arrays(nx, ny, g, gm)
        for k, name in enumerate(sorted(fields, key=T.FIELDS.index)):
            ot.L.orc_unpack_message(C.byref(ot.c), ot.f(name).ctypes.data_as(P), ftype[name], face, depth, want.ctypes.data_as(P),
                                    C.c_long(k * depth * (chunk_cells + 5) + tile_offset))
        _same(other, {k: ot.f(k) for k in T.FIELDS}, "unpack_fields")
        with pytest.raises(clv.CloverError, match="buffer_len"):
            small = clv.DeviceBuffer(n - 1)
            ch.pack_fields(face, fields, depth, chunk_cells, tile_offset, small)
        buf.close()


def test_decomposed_halo_rejects_bad_use(gpu):
    with clv.Chunk(8, 6) as a, clv.Chunk(5, 7) as b:
        with pytest.raises(clv.CloverError, match="share y_cells"):
            a.update_tile_halo(b, clv.FACE_LEFT, ["density0"], 1)
        with pytest.raises(clv.CloverError, match="own neighbour"):
            a.update_tile_halo(a, clv.FACE_TOP, ["density0"], 1)
        buf = clv.DeviceBuffer(8)
        with pytest.raises(clv.CloverError, match="does not fit"):
            a.pack_message(a.field_ptr("density0"), 1, clv.FACE_LEFT, 2, buf, 0)
        buf.close()


CELL_OUT = ["density0", "energy0", "pressure", "viscosity", "soundspeed", "density1", "energy1"]
NODE_OUT = ["xvel0", "yvel0", "xvel1", "yvel1"]
DECOMPOSITIONS = [(64, 48, 4), (61, 47, 6), (40, 90, 3), (33, 29, 9)]


def _oracle_mesh(ch, p, name):
    """A piece's mesh array at the oracle's length (x arrays nx + 5, y arrays ny + 5)."""
    a = ch.download_mesh(name)
    out = np.zeros((p.nx if name in ("cellx", "vertexx", "celldx", "vertexdx") else p.ny) + 5)
    out[:len(a)] = a
    return out


def _pieces(mode, nx, ny, n):
    return (tile_decompose if mode == "tiles" else chunk_decompose)(nx, ny, n)


@pytest.mark.parametrize("mode", ["tiles", "chunks"])
@pytest.mark.parametrize("nx,ny,n", DECOMPOSITIONS)
def test_decomposed_run_vs_oracle(gpu, mode, nx, ny, n):
    """The decomposed GPU run (tile copies, or packed messages between chunks) and the
    oracle on the same pieces (tile copies), from the same start: dt of every step
    and every piece's interior bit for bit."""
    cfg = clv.parse_deck(clv.bm_deck(nx, ny))
    pieces = _pieces(mode, nx, ny, n)
    print(f"# decomposed run vs oracle: {mode} {nx}x{ny} in {n} pieces", flush=True)
    with Decomposed(cfg, pieces, mode) as dec:
        dec.start()
        arrays = [({k: ch.download(k) for k in T.FIELDS}, {k: _oracle_mesh(ch, p, k) for k in T.MESH})
                  for p, ch in zip(pieces, dec.chunks)]
        orc = T.DecomposedOracle(T.parse_deck(clv.bm_deck(nx, ny)), pieces, arrays)
        for step in range(1, 16):
            dec.step()
            dt = orc.step()
            assert dec.st.dt == dt, (step, dec.st.dt, dt)
            for name in CELL_OUT + NODE_OUT:
                e = 1 if name in NODE_OUT else 0
                for k, (p, ch, o) in enumerate(zip(pieces, dec.chunks, orc.o)):
                    got = ch.download(name)[2:p.ny + 2 + e, 2:p.nx + 2 + e]
                    assert np.array_equal(got, o.f(name)[2:p.ny + 2 + e, 2:p.nx + 2 + e]), (step, name, k)


@pytest.mark.parametrize("mode", ["tiles", "chunks"])
@pytest.mark.parametrize("nx,ny,n", DECOMPOSITIONS)
def test_decomposed_run_vs_undecomposed(gpu, mode, nx, ny, n):
    """Against the undecomposed run: the same initial state bit for bit (each piece's
    own mesh and generate_chunk); after 15 steps the same problem.  The pieces'
    kernels clamp at their own edges as the reference's do (advec_cell's upwind
    index, src/serial/advec_cell.cpp:77,169), so fields agree to round-off-sized
    differences and the summed summaries to 1e-10."""
    cfg = clv.parse_deck(clv.bm_deck(nx, ny))
    pieces = _pieces(mode, nx, ny, n)
    with clv.Chunk(nx, ny) as single, Decomposed(cfg, pieces, mode) as dec:
        st, _ = single.start(cfg)
        dec.start()
        for name in ["density0", "energy0", "pressure", "soundspeed"]:
            assert np.array_equal(dec.gather(name), single.download(name)[2:ny + 2, 2:nx + 2]), name
        for _ in range(15):
            single.step(cfg, st, clv.PLAN_REFERENCE)
            dec.step()
        assert abs(dec.st.time - st.time) <= 1e-12 * st.time
        rho = single.download("density0")[2:ny + 2, 2:nx + 2]
        assert np.max(np.abs(dec.gather("density0") - rho)) <= 1e-10 * np.max(np.abs(rho))
        want, got = single.summary(), dec.summary()
        for w, g in zip(want, got):
            assert abs(w - g) <= 1e-10 * max(abs(w), 1e-300), (want, got)


def test_tiles_and_chunks_agree(gpu):
    """The same pieces exchanged by tile copies and by packed messages: bit for bit."""
    nx, ny = 45, 38
    cfg = clv.parse_deck(clv.bm_deck(nx, ny))
    pieces = tile_decompose(nx, ny, 6)
    with Decomposed(cfg, pieces, "tiles") as a, Decomposed(cfg, pieces, "chunks") as b:
        a.start()
        b.start()
        for _ in range(12):
            a.step()
            b.step()
            assert a.st.dt == b.st.dt
        for name in CELL_OUT + NODE_OUT:
            for ca, cb in zip(a.chunks, b.chunks):
                assert np.array_equal(ca.download(name), cb.download(name)), name
