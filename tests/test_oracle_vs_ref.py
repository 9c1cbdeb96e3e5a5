"""Each oracle kernel bit-for-bit against the reference serial kernel (no GPU).

The reference kernels come from oracle/_ref/librefkern.so, which oracle/Makefile
compiles from the reference's own src/serial sources where the reference tree is
present; elsewhere these tests skip (the oracle is then pinned by the golden
traces of test_oracle.py).  Inputs are random but physically shaped
(clover_testlib.random_chunk_arrays): positive densities/energies/volumes,
mixed-sign velocities and fluxes, nonuniform mesh spacing, so both upwind
branches, every limiter branch and the viscosity branch are taken.
"""
import ctypes as C
import zlib

import numpy as np
import pytest

import clover_testlib as T

pytestmark = pytest.mark.skipif(T.refkern_lib() is None, reason="oracle/_ref/librefkern.so not built (no reference tree)")

SIZES = [(8, 8), (13, 7), (24, 31)]


def _oracle_run(kernel, nx, ny, f2d, m1d, ia=(), da=()):
    o = T.Oracle(T.Deck(x_cells=nx, y_cells=ny, states=[T.State(density=1, energy=1)]))
    for n in T.FIELDS:
        o.f(n)[...] = f2d[n]
    for n in T.MESH:
        o.m(n)[...] = m1d[n]
    L, cp = o.L, C.byref(o.c)
    dout, iout = [0.0], [0, 0, 0, 0]
    if kernel == "ideal_gas":
        L.orc_ideal_gas(cp, ia[0])
    elif kernel == "viscosity":
        L.orc_viscosity(cp)
    elif kernel == "calc_dt":
        dtv = C.c_double()
        ints = [C.c_int() for _ in range(4)]
        L.orc_calc_dt(cp, *da, C.byref(dtv), *[C.byref(x) for x in ints])
        dout = [dtv.value]
        iout = [x.value for x in ints]
    elif kernel == "PdV":
        L.orc_PdV(cp, ia[0], da[0])
    elif kernel == "revert":
        L.orc_revert(cp)
    elif kernel == "accelerate":
        L.orc_accelerate(cp, da[0])
    elif kernel == "flux_calc":
        L.orc_flux_calc(cp, da[0])
    elif kernel == "advec_cell":
        L.orc_advec_cell(cp, ia[0], ia[1])
    elif kernel == "advec_mom":
        L.orc_advec_mom(cp, ia[0], ia[1], ia[2])
    elif kernel == "reset_field":
        L.orc_reset_field(cp)
    elif kernel == "update_halo":
        m = (C.c_int * 15)(*ia[1:])
        L.orc_update_halo(cp, m, ia[0])
    return {n: o.f(n).copy() for n in T.FIELDS}, dout, iout


def _assert_same(kernel, nx, ny, ref, orc):
    for n in T.FIELDS:
        ex, ey = T.EXTENT[n]
        a = ref[n][: ny + ey, : nx + ex]
        b = orc[n][: ny + ey, : nx + ex]
        same = (a == b) | (np.isnan(a) & np.isnan(b))
        assert same.all(), f"{kernel}: field {n} differs at {np.argwhere(~same)[:5].tolist()}"


CASES = [("ideal_gas", (0,), ()), ("ideal_gas", (1,), ()), ("viscosity", (), ()),
         ("PdV", (1,), (0.0123,)), ("PdV", (0,), (0.0123,)), ("revert", (), ()),
         ("accelerate", (), (0.0123,)), ("flux_calc", (), (0.0123,)), ("reset_field", (), ())]
CASES += [("advec_cell", (sweep, d), ()) for sweep in (1, 2) for d in (1, 2)]
CASES += [("advec_mom", (wv, sweep, d), ()) for wv in (1, 2) for sweep in (1, 2) for d in (1, 2)]


@pytest.mark.parametrize("nx,ny", SIZES)
@pytest.mark.parametrize("kernel,ia,da", CASES)
def test_kernel_bitwise(kernel, ia, da, nx, ny):
    f, m = T.random_chunk_arrays(nx, ny, seed=zlib.crc32(repr((kernel, ia, nx, ny)).encode()))
    if kernel == "advec_mom" and ia[0] == 2:
        # the yvel pass reads the node arrays the xvel pass left (src/serial/advec_mom.cpp:81-103)
        for w in ("work1", "work2", "work3"):
            f[w] = np.abs(f[w]) + 0.5
    if kernel == "advec_mom":
        # reference advec_mom takes (which_vel, sweep_number, direction) in this harness order
        ref, _, _ = T.refk_run(kernel, nx, ny, f, m, ia=ia)
        orc, _, _ = _oracle_run(kernel, nx, ny, f, m, ia=ia)
    else:
        ref, _, _ = T.refk_run(kernel, nx, ny, f, m, ia=ia, da=da)
        orc, _, _ = _oracle_run(kernel, nx, ny, f, m, ia=ia, da=da)
    _assert_same(kernel, nx, ny, ref, orc)


@pytest.mark.parametrize("nx,ny", SIZES)
@pytest.mark.parametrize("seed", [1, 2])
def test_calc_dt_bitwise(nx, ny, seed):
    f, m = T.random_chunk_arrays(nx, ny, seed=seed)
    da = (1e-7, 0.7, 0.5, 0.5, 0.7)
    _, dref, iref = T.refk_run("calc_dt", nx, ny, f, m, da=da)
    _, dorc, iorc = _oracle_run("calc_dt", nx, ny, f, m, da=da)
    assert dref[0] == dorc[0]
    assert iref == iorc  # control, jldt, kldt, small


@pytest.mark.parametrize("nx,ny", SIZES + [(10, 2), (3, 5)])
@pytest.mark.parametrize("depth", [1, 2])
@pytest.mark.parametrize("fields", [list(range(15)), [0, 2, 4, 7, 9], [11, 12, 13, 14], [6]])
def test_update_halo_bitwise(nx, ny, depth, fields):
    f, m = T.random_chunk_arrays(nx, ny, seed=nx * 100 + ny + depth)
    mask = [1 if i in fields else 0 for i in range(15)]
    ref, _, _ = T.refk_run("update_halo", nx, ny, f, m, ia=[depth] + mask)
    orc, _, _ = _oracle_run("update_halo", nx, ny, f, m, ia=[depth] + mask)
    _assert_same("update_halo", nx, ny, ref, orc)


# ---- decomposed halos: faces shared with a tile or chunk, tile copies, messages ---------------
HALO_SETS = [T.FIELDS[:15], ["density0", "energy0", "pressure", "xvel0", "yvel0"], ["vol_flux_x", "vol_flux_y", "mass_flux_x",
                                                                                     "mass_flux_y"], ["soundspeed"]]
EXT_PATTERNS = [(1, 1, 1, 1), (0, 1, 1, 1), (1, 0, 1, 1), (1, 1, 0, 1), (1, 1, 1, 0), (0, 0, 1, 1), (1, 1, 0, 0), (0, 1, 0, 1),
                (1, 0, 1, 0), (0, 0, 0, 0)]


@pytest.mark.parametrize("nx,ny", [(8, 8), (13, 7)])
@pytest.mark.parametrize("depth", [1, 2])
@pytest.mark.parametrize("ext", EXT_PATTERNS)
@pytest.mark.parametrize("fields", [0, 1, 2, 3])
def test_update_halo_external_faces_bitwise(nx, ny, depth, ext, fields):
    """src/serial/update_halo.cpp:33-684 with only the flagged faces external."""
    f, m = T.random_chunk_arrays(nx, ny, seed=zlib.crc32(repr((nx, ny, depth, ext, fields)).encode()))
    names = HALO_SETS[fields]
    mask = [1 if n in names else 0 for n in T.FIELDS[:15]]
    ref, _, _ = T.refk_run("update_halo_ext", nx, ny, f, m, ia=[depth] + mask + list(ext))
    o = T.oracle_on_arrays(nx, ny, f, m)
    o.L.orc_update_halo_ext(C.byref(o.c), T.mask(names), depth, (C.c_int * 4)(*ext))
    _assert_same("update_halo_ext", nx, ny, ref, {n: o.f(n) for n in T.FIELDS})


@pytest.mark.parametrize("face", [0, 1, 2, 3])
@pytest.mark.parametrize("depth", [1, 2])
@pytest.mark.parametrize("fields", [0, 1, 2])
def test_tile_halo_bitwise(face, depth, fields):
    """update_tile_halo_{l,r,b,t}_kernel (src/serial/update_tile_halo_kernel.cpp:30-804):
    the tile and its neighbour differ in size along the face normal."""
    nx, ny = 11, 9
    nnx, nny = (7, ny) if face in (0, 1) else (nx, 6)
    seed = zlib.crc32(repr((face, depth, fields)).encode())
    f, m = T.random_chunk_arrays(nx, ny, seed=seed)
    nf, nm = T.random_chunk_arrays(nnx, nny, seed=seed + 1)
    names = HALO_SETS[fields]
    ref = T.refk_tile_halo(face, nx, ny, f, m, nnx, nny, nf, nm, names, depth)
    o = T.oracle_on_arrays(nx, ny, f, m)
    n = T.oracle_on_arrays(nnx, nny, nf, nm)
    o.L.orc_tile_halo(C.byref(o.c), C.byref(n.c), face, T.mask(names), depth)
    _assert_same("tile_halo", nx, ny, ref, {k: o.f(k) for k in T.FIELDS})


@pytest.mark.parametrize("face", [0, 1, 2, 3])
@pytest.mark.parametrize("unpack", [0, 1])
@pytest.mark.parametrize("field_type", [1, 2, 3, 4])
@pytest.mark.parametrize("depth", [1, 2])
def test_pack_message_bitwise(face, unpack, field_type, depth):
    """clover_pack_message_* / clover_unpack_message_* (src/serial/pack_kernel.cpp:27-303):
    message layout and the layers read or written, with a nonzero buffer offset."""
    nx, ny = 10, 7
    rng = np.random.default_rng(zlib.crc32(repr((face, unpack, field_type, depth)).encode()))
    field = rng.uniform(-1, 1, (ny + 5, nx + 5))
    offset = 3 * depth * ((ny if face in (0, 1) else nx) + 5)
    buf = rng.uniform(-1, 1, offset + 2 * depth * (max(nx, ny) + 5))
    ref_a, ref_b = T.refk_pack(face, unpack, nx, ny, field, field_type, depth, buf, offset)
    o = T.oracle_on_arrays(nx, ny, *T.random_chunk_arrays(nx, ny, seed=1))
    a = field.copy()
    b = buf.copy()
    P = C.POINTER(C.c_double)
    fn = o.L.orc_unpack_message if unpack else o.L.orc_pack_message
    fn(C.byref(o.c), a.ctypes.data_as(P), field_type, face, depth, b.ctypes.data_as(P), C.c_long(offset))
    assert np.array_equal(a, ref_a) and np.array_equal(b, ref_b)
