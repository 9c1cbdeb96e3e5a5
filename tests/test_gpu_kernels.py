"""Each gfx950 kernel, through the C ABI, bit-for-bit against the oracle (GPU).

Random but physically shaped inputs (clover_testlib.random_chunk_arrays) are
uploaded to a chunk, one entry point runs, and every array is downloaded and
compared with the oracle kernel run on the same inputs: bit-identical over the
whole (ny+5) x (nx+5) extent, ghost layers included.
"""
import ctypes as C
import zlib

import numpy as np
import pytest

import clover_testlib as T
import cloverleaf_amd as clv

pytestmark = pytest.mark.gpu

SIZES = [(8, 8), (13, 7), (70, 33), (257, 130)]


def _load(ch, f, m):
    for n in T.FIELDS:
        ch.upload(n, f[n])
    for n in T.MESH:
        ch.upload_mesh(n, m[n])


def _oracle(nx, ny, f, m):
    o = T.Oracle(T.Deck(x_cells=nx, y_cells=ny, states=[T.State(density=1, energy=1)]))
    for n in T.FIELDS:
        o.f(n)[...] = f[n]
    for n in T.MESH:
        o.m(n)[...] = m[n]
    return o


def _same(ch, o, what):
    bad = []
    for n in T.FIELDS:
        a, b = ch.download(n), o.f(n)
        same = (a == b) | (np.isnan(a) & np.isnan(b))
        if not same.all():
            bad.append((n, np.argwhere(~same)[:3].tolist()))
    assert not bad, f"{what}: {bad}"


CASES = [("ideal_gas", (0,)), ("ideal_gas", (1,)), ("viscosity", ()), ("PdV", (1, 0.0123)), ("PdV", (0, 0.0123)),
         ("revert", ()), ("accelerate", (0.0123,)), ("flux_calc", (0.0123,)), ("reset_field", ())]
CASES += [("advec_cell", (sweep, d)) for sweep in (1, 2) for d in (1, 2)]
CASES += [("advec_mom", (wv, d, sweep)) for wv in (1, 2) for d in (1, 2) for sweep in (1, 2)]


@pytest.mark.parametrize("nx,ny", SIZES)
@pytest.mark.parametrize("kernel,args", CASES)
def test_kernel_matches_oracle(gpu, kernel, args, nx, ny):
    f, m = T.random_chunk_arrays(nx, ny, seed=zlib.crc32(repr((kernel, args, nx, ny)).encode()))
    if kernel == "advec_mom" and args[0] == 2:
        for w in ("work1", "work2", "work3"):
            f[w] = np.abs(f[w]) + 0.5
    o = _oracle(nx, ny, f, m)
    oc = C.byref(o.c)
    with clv.Chunk(nx, ny) as ch:
        _load(ch, f, m)
        if kernel == "ideal_gas":
            ch.ideal_gas(args[0]); o.L.orc_ideal_gas(oc, args[0])
        elif kernel == "viscosity":
            ch.viscosity(); o.L.orc_viscosity(oc)
        elif kernel == "PdV":
            ch.PdV(args[0], args[1]); o.L.orc_PdV(oc, args[0], args[1])
        elif kernel == "revert":
            ch.revert(); o.L.orc_revert(oc)
        elif kernel == "accelerate":
            ch.accelerate(args[0]); o.L.orc_accelerate(oc, args[0])
        elif kernel == "flux_calc":
            ch.flux_calc(args[0]); o.L.orc_flux_calc(oc, args[0])
        elif kernel == "reset_field":
            ch.reset_field(); o.L.orc_reset_field(oc)
        elif kernel == "advec_cell":
            ch.advec_cell(args[0], args[1]); o.L.orc_advec_cell(oc, args[0], args[1])
        elif kernel == "advec_mom":
            wv, d, sweep = args
            ch.advec_mom(wv, d, sweep); o.L.orc_advec_mom(oc, wv, sweep, d)
        _same(ch, o, f"{kernel}{args} {nx}x{ny}")


@pytest.mark.parametrize("nx,ny", SIZES + [(10, 2), (3, 5), (4, 4)])
@pytest.mark.parametrize("depth", [1, 2])
@pytest.mark.parametrize("fields", [list(range(15)), [0, 2, 4, 7, 9], [11, 12, 13, 14], [6]])
def test_update_halo_matches_oracle(gpu, nx, ny, depth, fields):
    f, m = T.random_chunk_arrays(nx, ny, seed=nx * 1000 + ny * 10 + depth)
    o = _oracle(nx, ny, f, m)
    with clv.Chunk(nx, ny) as ch:
        _load(ch, f, m)
        ch.update_halo(fields, depth)
        o.L.orc_update_halo(C.byref(o.c), T.mask([T.FIELDS[i] for i in fields]), depth)
        _same(ch, o, f"update_halo {fields} depth {depth} {nx}x{ny}")


@pytest.mark.parametrize("nx,ny", SIZES)
def test_calc_dt_matches_oracle(gpu, nx, ny):
    f, m = T.random_chunk_arrays(nx, ny, seed=nx + ny)
    o = _oracle(nx, ny, f, m)
    dtv = C.c_double()
    ints = [C.c_int() for _ in range(4)]
    o.L.orc_calc_dt(C.byref(o.c), 1e-7, 0.7, 0.5, 0.5, 0.7, C.byref(dtv), *[C.byref(x) for x in ints])
    with clv.Chunk(nx, ny) as ch:
        _load(ch, f, m)
        dt, ctl, jl, kl, small = ch.calc_dt(1e-7, 0.7, 0.5, 0.5, 0.7)
    assert dt == dtv.value
    assert [ctl, jl, kl, small] == [x.value for x in ints]


@pytest.mark.parametrize("nx,ny", SIZES)
def test_field_summary_close_to_oracle(gpu, nx, ny):
    """Fixed-order tree sum vs the serial order: same terms, rounding-level gap."""
    f, m = T.random_chunk_arrays(nx, ny, seed=3 * nx + ny)
    o = _oracle(nx, ny, f, m)
    ref = (C.c_double * 5)()
    o.L.orc_field_summary_kernel(C.byref(o.c), ref)
    with clv.Chunk(nx, ny) as ch:
        _load(ch, f, m)
        got = ch.field_summary()
        again = ch.field_summary()
    assert got == again  # reproducible
    for a, b in zip(got, ref):
        assert abs(a - b) <= 1e-13 * abs(b), (a, b)


def test_setup_kernels_match_oracle(gpu):
    """build_field + initialise_chunk + generate_chunk (rectangle, circle, point states)."""
    d = T.load_deck("small_geom_mix")
    cfg = clv.parse_deck((T.DECKS / "small_geom_mix.in").read_text())
    o = T.Oracle(d)
    cp = C.byref(o.c)
    o.L.orc_build_field(cp)
    o.L.orc_initialise_chunk(cp, C.byref(o.cfg))
    o.L.orc_generate_chunk(cp, C.byref(o.cfg))
    with clv.Chunk(d.x_cells, d.y_cells) as ch:
        ch.build_field()
        ch.initialise_chunk(cfg)
        ch.generate_chunk(cfg)
        _same(ch, o, "set-up")
        for n in T.MESH:
            got = ch.download_mesh(n)
            assert np.array_equal(got, o.m(n)[:len(got)]), n


def test_measurement_probes(gpu):
    """The copy-bandwidth and access-pattern probes run and time something; the
    pattern probe accepts only the fused kernels' shapes."""
    assert clv.bandwidth_probe(64 << 20, 2) > 0.0
    for r, w in [(4, 1), (5, 6), (6, 5), (5, 5), (7, 1), (8, 6), (7, 5)]:
        assert clv.march_probe(300, 200, r, w, 16, 2) > 0.0
    with pytest.raises(clv.CloverError, match="reads, writes"):
        clv.march_probe(300, 200, 3, 3, 16, 1)
    with pytest.raises(clv.CloverError, match="bad argument"):
        clv.march_probe(0, 200, 4, 1, 16, 1)


def test_bad_arguments_are_rejected(gpu):
    with clv.Chunk(16, 16) as ch:
        with pytest.raises(clv.CloverError):
            ch.advec_cell(3, 1)
        with pytest.raises(clv.CloverError):
            ch.update_halo([0], 3)
        with pytest.raises(clv.CloverError):
            ch.advec_mom(1, 0, 1)
