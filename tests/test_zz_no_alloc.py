"""The allocation-free entry point clv_field_summary_noalloc, through the C ABI.

It runs the kernels of clv_field_summary (the fixed tree order) on a caller's view
with its partials in the library's fixed per-device scratch array, so its five sums
are the same doubles as clv_field_summary's on the same arrays.  Checked on two
calls in a row, the second on a smaller grid (fewer block partials than the first
left in the scratch array), and against the oracle's sums of the same arrays.
"""
import ctypes as C

import numpy as np
import pytest

import clover_testlib as T
import cloverleaf_amd as clv

pytestmark = pytest.mark.gpu


def _deck(nx, ny, steps):
    return T.DECKS.joinpath("small_geom_mix.in").read_text().replace("x_cells=40", f"x_cells={nx}").replace(
        "y_cells=56", f"y_cells={ny}").replace("end_step=80", f"end_step={steps}")


def _exact(ch):
    """math.fsum of the reference's per-cell terms (src/serial/field_summary.cpp:76-90)
    over the chunk's arrays."""
    import math
    nx, ny = ch.nx, ch.ny
    f = {n: ch.download(n) for n in ("volume", "density0", "energy0", "pressure", "xvel0", "yvel0")}
    vol = f["volume"][2:ny + 2, 2:nx + 2]
    mass = vol * f["density0"][2:ny + 2, 2:nx + 2]
    vs = np.zeros_like(vol)
    for kv in (0, 1):
        for jv in (0, 1):
            u = f["xvel0"][2 + kv:ny + 2 + kv, 2 + jv:nx + 2 + jv]
            v = f["yvel0"][2 + kv:ny + 2 + kv, 2 + jv:nx + 2 + jv]
            vs = vs + 0.25 * (u * u + v * v)
    terms = [vol, mass, mass * f["energy0"][2:ny + 2, 2:nx + 2], mass * 0.5 * vs, vol * f["pressure"][2:ny + 2, 2:nx + 2]]
    return [math.fsum(t.ravel().tolist()) for t in terms]


@pytest.mark.parametrize("sizes", [((300, 2100), (64, 48)), ((97, 130), (9, 5))], ids=["2048-blocks-then-48", "130-then-5"])
def test_noalloc_summary_equals_field_summary(gpu, sizes):
    for call, (nx, ny) in enumerate(sizes):
        print(f"# noalloc call {call + 1}: {nx}x{ny}", flush=True)
        text = _deck(nx, ny, 6)
        cfg = clv.parse_deck(text)
        with clv.Chunk(nx, ny) as ch:
            st, _ = ch.start(cfg)
            while not st.complete:
                ch.step(cfg, st, clv.PLAN_REFERENCE)
            ch.ideal_gas(False)
            want = ch.field_summary()
            got = ch.field_summary_noalloc()
            assert got == want, (call, got, want)
            again = ch.field_summary_noalloc()
            assert again == want
            ex = _exact(ch)
        for g, x in zip(got, ex):
            assert abs(g - x) <= 1e-13 * abs(x), (call, g, x)


def test_noalloc_refuses_bad_views(gpu):
    L = clv.lib()
    out = (C.c_double * 5)()
    with clv.Chunk(16, 12) as ch:
        v = ch.view()
        bad = clv.ChunkView.from_buffer_copy(v)
        bad.pitch = 16 + 4
        assert L.clv_field_summary_noalloc(C.byref(bad), None, out) == -1
        assert b"pitch" in L.clv_last_error()
        bad = clv.ChunkView.from_buffer_copy(v)
        bad.field[clv.FID["pressure"]] = None
        assert L.clv_field_summary_noalloc(C.byref(bad), None, out) == -1
        assert L.clv_field_summary_noalloc(None, None, out) == -1
