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


@pytest.mark.parametrize("depth", [1, 2])
@pytest.mark.parametrize("nx,ny", [(4, 4), (8, 8), (13, 7), (70, 33), (257, 130)])
def test_timestep_kernels_equal_the_entry_point_sequence(gpu, nx, ny, depth):
    """clv_timestep_kernels (ideal_gas, viscosity and calc_dt in one pass, then the
    pressure and viscosity halos) against driver/timestep.cpp:47-69's sequence of
    entry points on the same arrays: every array bit for bit, and the same dt and
    bookkeeping.  Both start with the density0, energy0, xvel0, yvel0 halos done, as
    the drop-in does before it runs the one-pass form."""
    f, m = T.random_chunk_arrays(nx, ny, seed=7 * nx + ny + depth)
    first = ["energy0", "density0", "xvel0", "yvel0"]
    with clv.Chunk(nx, ny) as a, clv.Chunk(nx, ny) as b:
        _load(a, f, m)
        _load(b, f, m)
        a.update_halo(first, 1)
        b.update_halo(first, 1)
        got = a.timestep_kernels(depth)
        b.ideal_gas(False)
        b.update_halo(["pressure"], depth)
        b.viscosity()
        b.update_halo(["viscosity"], 1)
        want = b.calc_dt()
        assert got == want
        for n in T.FIELDS:
            assert np.array_equal(a.download(n), b.download(n), equal_nan=True), n


@pytest.mark.parametrize("nx,ny", [(4, 4), (8, 8), (13, 7), (70, 33), (257, 130)])
def test_timestep_kernels_without_pressure_then_lagrange(gpu, nx, ny):
    """clv_timestep_kernels with pressure_halo_depth 0 stores neither pressure nor
    soundspeed (viscosity and dt as with depth 1); followed by clv_lagrange_kernels,
    which rewrites both, every array equals the depth-1 form followed by the same
    call, bit for bit."""
    f, m = T.random_chunk_arrays(nx, ny, seed=3 * nx + ny + 1)
    dt = 1.1e-3
    with clv.Chunk(nx, ny) as a, clv.Chunk(nx, ny) as b:
        for ch in (a, b):
            _load(ch, f, m)
            ch.update_halo(["energy0", "density0", "xvel0", "yvel0"], 1)
        got = a.timestep_kernels(0)
        want = b.timestep_kernels(1)
        assert got == want
        for n in T.FIELDS:
            if n in ("pressure", "soundspeed"):
                assert np.array_equal(a.download(n), f[n], equal_nan=True), n  # left as they were
            else:
                assert np.array_equal(a.download(n), b.download(n), equal_nan=True), n
        a.lagrange_kernels(dt)
        b.lagrange_kernels(dt)
        for n in T.FIELDS:
            if n.startswith("work"):
                continue
            assert np.array_equal(a.download(n), b.download(n), equal_nan=True), n


@pytest.mark.parametrize("nx,ny", [(4, 4), (8, 8), (13, 7), (70, 33), (257, 130)])
def test_lagrange_kernels_equal_the_entry_point_sequence(gpu, nx, ny):
    """clv_lagrange_kernels against PdV(predict) chain, accelerate, PdV(correct),
    flux_calc on the same arrays, after the same timestep (its halos and pressure):
    every array bit for bit, the predictor's pressure, its halo and soundspeed
    included."""
    f, m = T.random_chunk_arrays(nx, ny, seed=11 * nx + ny)
    dt = 1.3e-3
    with clv.Chunk(nx, ny) as a, clv.Chunk(nx, ny) as b:
        for ch in (a, b):
            _load(ch, f, m)
            ch.update_halo(["energy0", "density0", "xvel0", "yvel0"], 1)
            ch.timestep_kernels(1)
        a.lagrange_kernels(dt)
        b.PdV_driver(True, dt)
        b.accelerate(dt)
        b.PdV_driver(False, dt)
        b.flux_calc(dt)
        for n in T.FIELDS:
            if n.startswith("work"):
                continue  # scratch of the reference's loops
            assert np.array_equal(a.download(n), b.download(n), equal_nan=True), n


# 4100 columns: 67 strips of the y pass, past the width where its blocks hold 8
# walks, so the last block has 5 walks past the chunk
ADVECTION_SIZES = [(4, 4), (8, 8), (13, 7), (70, 33), (257, 130), (4100, 24)]


@pytest.mark.parametrize("sweep,direction", [(1, 1), (1, 2), (2, 1), (2, 2)])
@pytest.mark.parametrize("nx,ny", ADVECTION_SIZES)
def test_advection_pass_equals_the_entry_point_sequence(gpu, nx, ny, sweep, direction):
    """clv_advection_pass against advec_cell, update_halo(density1, energy1, xvel1,
    yvel1, mass fluxes, 2), advec_mom(xvel), advec_mom(yvel) on the same arrays after
    the same halos: every array but the work arrays bit for bit over its extent
    (driver/definitions.cpp:4-41), ghosts included."""
    f, m = T.random_chunk_arrays(nx, ny, seed=13 * nx + ny + 7 * sweep + direction)
    with clv.Chunk(nx, ny) as a, clv.Chunk(nx, ny) as b:
        for ch in (a, b):
            _load(ch, f, m)
            ch.update_halo(["energy1", "density1", "vol_flux_x", "vol_flux_y", "xvel1", "yvel1"], 2)
        a.advection_pass(sweep, direction)
        b.advec_cell(sweep, direction)
        b.update_halo(["density1", "energy1", "xvel1", "yvel1", "mass_flux_x", "mass_flux_y"], 2)
        b.advec_mom(1, direction, sweep)
        b.advec_mom(2, direction, sweep)
        bad = []
        for n in T.FIELDS:
            if n.startswith("work"):
                continue  # scratch of the reference's loops
            # the reference's extent of the field (the exchanged buffers hold other
            # values in the padding past it)
            ex, ey = T.EXTENT[n]
            x, y = a.download(n)[:ny + ey, :nx + ex], b.download(n)[:ny + ey, :nx + ex]
            same = (x == y) | (np.isnan(x) & np.isnan(y))
            if not same.all():
                bad.append((n, [(tuple(int(v) for v in ij), float(x[tuple(ij)]), float(y[tuple(ij)])) for ij in np.argwhere(~same)[:4]]))
        assert not bad, bad


@pytest.mark.parametrize("sweep,direction", [(1, 1), (1, 2), (2, 1), (2, 2)])
@pytest.mark.parametrize("nx,ny", ADVECTION_SIZES)
def test_advection_pass_reset_equals_pass_then_reset_field(gpu, nx, ny, sweep, direction):
    """clv_advection_pass_reset against clv_advection_pass followed by clv_reset_field
    (src/serial/reset_field.cpp:29-42): every array but the work arrays bit for bit
    over its extent, the ghost frames of density0, energy0, xvel0 and yvel0 (which
    reset_field leaves alone) included."""
    f, m = T.random_chunk_arrays(nx, ny, seed=11 * nx + ny + 5 * sweep + direction)
    with clv.Chunk(nx, ny) as a, clv.Chunk(nx, ny) as b:
        for ch in (a, b):
            _load(ch, f, m)
            ch.update_halo(["energy1", "density1", "vol_flux_x", "vol_flux_y", "xvel1", "yvel1"], 2)
        a.advection_pass_reset(sweep, direction)
        b.advection_pass(sweep, direction)
        b.reset_field()
        bad = []
        for n in T.FIELDS:
            if n.startswith("work"):
                continue
            ex, ey = T.EXTENT[n]
            x, y = a.download(n)[:ny + ey, :nx + ex], b.download(n)[:ny + ey, :nx + ex]
            same = (x == y) | (np.isnan(x) & np.isnan(y))
            if not same.all():
                bad.append((n, [(tuple(int(v) for v in ij), float(x[tuple(ij)]), float(y[tuple(ij)])) for ij in np.argwhere(~same)[:4]]))
        assert not bad, bad
        # the copy really happened: density0 now holds the advected density
        x0, x1 = a.download("density0"), a.download("density1")
        assert np.array_equal(x0[2:ny + 2, 2:nx + 2], x1[2:ny + 2, 2:nx + 2])


@pytest.mark.parametrize("nx,ny", SIZES)
def test_summary_one_pass_equals_ideal_gas_then_sums(gpu, nx, ny):
    """clv_hydro_summary (ideal_gas and the sums in one pass) against clv_ideal_gas
    followed by clv_field_summary: the same sums, pressure and soundspeed bit for
    bit, and the oracle's ideal_gas arrays."""
    f, m = T.random_chunk_arrays(nx, ny, seed=5 * nx + ny)
    f["pressure"][:] = -1.0
    f["soundspeed"][:] = -1.0
    o = _oracle(nx, ny, f, m)
    o.L.orc_ideal_gas(C.byref(o.c), 0)
    with clv.Chunk(nx, ny) as a, clv.Chunk(nx, ny) as b:
        _load(a, f, m)
        _load(b, f, m)
        one = a.summary()
        b.ideal_gas(False)
        two = b.field_summary()
        assert one == two
        for name in ("pressure", "soundspeed"):
            assert np.array_equal(a.download(name), b.download(name), equal_nan=True), name
            assert np.array_equal(a.download(name), o.f(name), equal_nan=True), name


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
    for r, w in [(4, 1), (5, 6), (6, 5), (5, 5), (7, 1), (8, 6), (7, 5), (7, 0), (7, 6), (4, 0), (4, 6), (7, 4), (6, 4), (5, 4)]:
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


# ---- CLV_ENTRY_FUSED: advec_cell, advec_mom and the PdV predictor chain in one launch ----
def _interior_rows_cols(n, ext):
    return slice(2, n + 2 + ext)


def _same_except(ch, o, what, regions):
    """Every array equal to the oracle's over its whole extent, except those in
    `regions` (compared over the given (rows, cols) region only) and the work arrays."""
    bad = []
    for n in T.FIELDS:
        if n.startswith("work"):
            continue
        a, b = ch.download(n), o.f(n)
        if n in regions:
            a, b = a[regions[n]], b[regions[n]]
        same = (a == b) | (np.isnan(a) & np.isnan(b))
        if not same.all():
            bad.append((n, np.argwhere(~same)[:3].tolist()))
    assert not bad, f"{what}: {bad}"


ENTRY_SIZES = SIZES + [(4, 4), (5, 9), (63, 66), (65, 3)]


@pytest.mark.parametrize("nx,ny", ENTRY_SIZES)
@pytest.mark.parametrize("sweep,d", [(1, 1), (1, 2), (2, 1), (2, 2)])
def test_entry_fused_advec_cell_matches_oracle(gpu, nx, ny, sweep, d):
    """One launch: density1 and energy1 on the interior cells, the mass flux over its
    whole array (the reference's face range written, the rest kept) and every other
    array bit-identical to the oracle's three loops; density1/energy1 now live in
    other arrays (their ghosts wait for the driver's next halo update)."""
    f, m = T.random_chunk_arrays(nx, ny, seed=zlib.crc32(repr(("entry_cell", sweep, d, nx, ny)).encode()))
    o = _oracle(nx, ny, f, m)
    with clv.Chunk(nx, ny) as ch:
        _load(ch, f, m)
        ch.set_entry_plan(clv.ENTRY_FUSED)
        before = ch.field_ptr("density1")
        ch.advec_cell(sweep, d)
        o.L.orc_advec_cell(C.byref(o.c), sweep, d)
        assert ch.field_ptr("density1") != before
        inner = (slice(2, ny + 2), slice(2, nx + 2))
        _same_except(ch, o, f"entry advec_cell({sweep}, {d}) {nx}x{ny}", {"density1": inner, "energy1": inner})


@pytest.mark.parametrize("nx,ny", ENTRY_SIZES)
@pytest.mark.parametrize("d,sweep", [(1, 1), (2, 1), (1, 2), (2, 2)])
def test_entry_fused_advec_mom_matches_oracle(gpu, nx, ny, d, sweep):
    """xvel then yvel (the reference's yvel call reuses the xvel call's node arrays; the
    fused yvel call recomputes them): both velocities on the interior and boundary
    nodes and every other array bit-identical to the oracle."""
    f, m = T.random_chunk_arrays(nx, ny, seed=zlib.crc32(repr(("entry_mom", sweep, d, nx, ny)).encode()))
    o = _oracle(nx, ny, f, m)
    with clv.Chunk(nx, ny) as ch:
        _load(ch, f, m)
        ch.set_entry_plan(clv.ENTRY_FUSED)
        nodes = (slice(2, ny + 3), slice(2, nx + 3))
        ch.advec_mom(1, d, sweep)
        o.L.orc_advec_mom(C.byref(o.c), 1, sweep, d)
        _same_except(ch, o, f"entry advec_mom(xvel, {d}, {sweep}) {nx}x{ny}", {"xvel1": nodes})
        ch.advec_mom(2, d, sweep)
        o.L.orc_advec_mom(C.byref(o.c), 2, sweep, d)
        _same_except(ch, o, f"entry advec_mom(yvel, {d}, {sweep}) {nx}x{ny}", {"xvel1": nodes, "yvel1": nodes})


@pytest.mark.parametrize("nx,ny", ENTRY_SIZES)
def test_entry_fused_pdv_predict_chain_matches_oracle(gpu, nx, ny):
    """PdV predictor + ideal_gas(predict) + pressure halo + revert in one launch: every
    array bit-identical to the oracle's chain over its whole extent."""
    f, m = T.random_chunk_arrays(nx, ny, seed=zlib.crc32(repr(("entry_pdv", nx, ny)).encode()))
    o = _oracle(nx, ny, f, m)
    with clv.Chunk(nx, ny) as ch:
        _load(ch, f, m)
        ch.set_entry_plan(clv.ENTRY_FUSED)
        ch.PdV_driver(1, 0.0123)
        o.L.orc_PdV_driver(C.byref(o.c), 1, 0.0123)
        _same(ch, o, f"entry PdV predict chain {nx}x{ny}")
        # the corrector is not fused: the same as the reference plan
        ch.PdV_driver(0, 0.0123)
        o.L.orc_PdV_driver(C.byref(o.c), 0, 0.0123)
        _same(ch, o, f"entry PdV corrector {nx}x{ny}")


@pytest.mark.parametrize("deck", ["small_bm_64x48", "small_geom_mix", "clover_sodxy"])
def test_entry_fused_whole_runs_match_oracle(gpu, deck):
    """A run through the per-kernel entry points under CLV_ENTRY_FUSED, sequenced like
    driver/hydro.cpp (timestep, PdV chain, accelerate, PdV, flux_calc, advection with
    its halo updates, reset_field): dt at every step and every persistent array equal
    to the oracle's."""
    d = T.load_deck(deck)
    cfg = clv.parse_deck((T.DECKS / f"{deck}.in").read_text())
    o = T.Oracle(d)
    o.start()
    nx, ny = d.x_cells, d.y_cells
    with clv.Chunk(nx, ny) as ch:
        st, _ = ch.start(cfg)
        ch.set_entry_plan(clv.ENTRY_FUSED)
        advect_x = True
        for step in range(1, 21):
            # driver/timestep.cpp:47-60
            ch.ideal_gas(False)
            ch.update_halo(["pressure", "energy0", "density0", "xvel0", "yvel0"], 1)
            ch.viscosity()
            ch.update_halo(["viscosity"], 1)
            dtc = ch.calc_dt(d.dtmin, d.dtc_safe, d.dtu_safe, d.dtv_safe, d.dtdiv_safe)[0]
            assert o.step() == 0
            dt = min(dtc, st.dtold * d.dtrise, d.dtmax)
            assert dt == o.g.dt, step
            st.dtold = dt
            ch.PdV_driver(1, dt)
            ch.accelerate(dt)
            ch.PdV_driver(0, dt)
            ch.flux_calc(dt)
            # driver/advection.cpp:29-105
            dirs = (1, 2) if advect_x else (2, 1)
            ch.update_halo(["energy1", "density1", "vol_flux_x", "vol_flux_y"], 2)
            ch.advec_cell(1, dirs[0])
            mom = ["density1", "energy1", "xvel1", "yvel1", "mass_flux_x", "mass_flux_y"]
            ch.update_halo(mom, 2)
            ch.advec_mom(1, dirs[0], 1)
            ch.advec_mom(2, dirs[0], 1)
            ch.advec_cell(2, dirs[1])
            ch.update_halo(mom, 2)
            ch.advec_mom(1, dirs[1], 2)
            ch.advec_mom(2, dirs[1], 2)
            ch.reset_field()
            advect_x = not advect_x
        cells = (slice(2, ny + 2), slice(2, nx + 2))
        nodes = (slice(2, ny + 3), slice(2, nx + 3))
        for n, reg in (("density0", cells), ("energy0", cells), ("xvel0", nodes), ("yvel0", nodes), ("density1", cells),
                       ("energy1", cells), ("xvel1", nodes), ("yvel1", nodes), ("viscosity", None), ("pressure", None),
                       ("vol_flux_x", None), ("vol_flux_y", None), ("mass_flux_x", None), ("mass_flux_y", None)):
            a, b = ch.download(n), o.f(n)
            if reg is not None:
                a, b = a[reg], b[reg]
            assert np.array_equal(a, b), n
