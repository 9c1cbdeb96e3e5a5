"""The batched form (clv_batch_*) on the GPU.

A batch step runs the fused step of every member with one persistent launch per
phase over all members' tiles, dt per member on the device, replayed as a
hipGraph.  Pinned by:
- the plain path: each member's state after a batched run is bit-identical to
  the same deck stepped alone with clv_hydro_step(..., CLV_PLAN_FUSED), for
  uneven sizes (strip, segment and XCD-split edges), members finishing at
  different steps, graph and direct launches, every schedule, runs split into
  pieces, and an autotune in the middle of a run;
- the reference: the 64 problems of the seeded workload (tests/golden/batch_seed355.json,
  written by the reference's serial model) -- dt of every step bit-identical, and
  the five field_summary sums bit-identical in the serial summation order;
- failure handling: a member whose dt falls below dtmin stops, the others go on.
"""
import json
import re

import numpy as np
import pytest

import cloverleaf_amd as clv

pytestmark = pytest.mark.gpu

SEED = 355
GOLDEN = clv.PKG.parent / "tests" / "golden" / f"batch_seed{SEED}.json"


def _deck(index, nx, ny, end_step):
    """A deck of the seeded workload with its grid and end_step replaced."""
    t = clv.batch_deck(SEED, index)
    t = re.sub(r"x_cells=\d+", f"x_cells={nx}", t)
    t = re.sub(r"y_cells=\d+", f"y_cells={ny}", t)
    return re.sub(r"end_step=\d+", f"end_step={end_step}", t)


# uneven members: strip edges (62/61/60/64 columns), row-segment edges, tiny grids
SMALL = [(4, 5, 15), (37, 61, 15), (64, 48, 9), (130, 97, 15), (249, 33, 12), (61, 200, 15), (250, 129, 7), (5, 300, 15)]


def _state_regions(nx, ny):
    """Arrays that carry the fused step's state (interior cells, interior and
    boundary nodes, and the volume fluxes whole; q and the mass fluxes are
    recomputed where they are used)."""
    return {"density0": (slice(2, ny + 2), slice(2, nx + 2)), "energy0": (slice(2, ny + 2), slice(2, nx + 2)),
            "xvel0": (slice(2, ny + 3), slice(2, nx + 3)), "yvel0": (slice(2, ny + 3), slice(2, nx + 3)),
            "vol_flux_x": (slice(None), slice(None)), "vol_flux_y": (slice(None), slice(None))}


def _start(decks, geometry=clv.GEOMETRY_ARRAYS):
    cfgs = [clv.parse_deck(t) for t in decks]
    chunks, states = [], []
    for c in cfgs:
        ch = clv.Chunk(c.x_cells, c.y_cells, geometry)
        st, _ = ch.start(c)
        chunks.append(ch)
        states.append(st)
    return cfgs, chunks, states


def _plain(decks, max_steps=None):
    """Each deck alone on the plain fused path: (state, {field: array}, dts)."""
    out = []
    for t in decks:
        cfg = clv.parse_deck(t)
        with clv.Chunk(cfg.x_cells, cfg.y_cells) as ch:
            st, _ = ch.start(cfg)
            dts = []
            while not st.complete and (max_steps is None or st.step < max_steps):
                ch.step(cfg, st, clv.PLAN_FUSED)
                dts.append(st.dt)
            arrs = {n: ch.download(n)[r] for n, r in _state_regions(cfg.x_cells, cfg.y_cells).items()}
            out.append((st, arrs, dts))
    return out


def _assert_same(batch, chunks, cfgs, plain):
    states = batch.states()
    for p, (ch, cfg, (pst, parr, pdts)) in enumerate(zip(chunks, cfgs, plain)):
        st = states[p]
        got = (st.step, st.time, st.dt, st.dtold, st.advect_x, st.complete, st.jdt, st.kdt, st.dt_control)
        want = (pst.step, pst.time, pst.dt, pst.dtold, pst.advect_x, pst.complete, pst.jdt, pst.kdt, pst.dt_control)
        assert got == want, (p, got, want)
        assert np.array_equal(batch.dt_history(p), np.array(pdts)), p
        for n, r in _state_regions(cfg.x_cells, cfg.y_cells).items():
            a = ch.download(n)[r]
            if not np.array_equal(a, parr[n]):
                raise AssertionError(f"member {p} ({cfg.x_cells}x{cfg.y_cells}) {n} differs at "
                                     f"{np.argwhere(a != parr[n])[:3].tolist()}")


@pytest.fixture(scope="module")
def small_decks():
    return [_deck(i, nx, ny, es) for i, (nx, ny, es) in enumerate(SMALL)]


@pytest.fixture(scope="module")
def small_plain(small_decks):
    return _plain(small_decks)


@pytest.mark.parametrize("geometry", [clv.GEOMETRY_ARRAYS, clv.GEOMETRY_MESH], ids=["arrays", "mesh"])
@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_bitwise_vs_plain(gpu, small_decks, small_plain, use_graph, geometry):
    """Both geometry modes: the volume and face areas read from their arrays, or
    computed from the mesh (the plain runs read the arrays: the same doubles)."""
    cfgs, chunks, states = _start(small_decks, geometry)
    with clv.Batch(chunks, cfgs, states, hist_len=64) as b:
        done = b.run(100, use_graph=use_graph)
        assert done == 15  # the largest end_step: no member runs past its own
        _assert_same(b, chunks, cfgs, small_plain)


SCHEDULES = [
    dict(seg=8, blocks_per_cu=1, order=clv.ORDER_GIVEN),
    dict(seg=32, blocks_per_cu=0, order=clv.ORDER_LARGEST_FIRST),
    dict(seg=128, blocks_per_cu=8, order=clv.ORDER_GIVEN),
    dict(seg=4096, blocks_per_cu=2, order=clv.ORDER_LARGEST_FIRST),
]


@pytest.mark.parametrize("sched", SCHEDULES, ids=lambda s: f"seg{s['seg']}-bpc{s['blocks_per_cu']}-o{s['order']}")
def test_batch_schedules_agree(gpu, small_decks, small_plain, sched):
    cfgs, chunks, states = _start(small_decks)
    with clv.Batch(chunks, cfgs, states, hist_len=64) as b:
        s = b.schedule()
        for f in s.fam:
            f.seg, f.blocks_per_cu, f.order = sched["seg"], sched["blocks_per_cu"], sched["order"]
        b.set_schedule(s)
        assert b.schedule().as_dict()["lagrange"] == sched
        b.run(100)
        _assert_same(b, chunks, cfgs, small_plain)


def test_batch_pieces_and_autotune(gpu, small_decks, small_plain):
    """Runs split at odd step counts (the graph parity carries over) with an
    autotune and a schedule change in between."""
    cfgs, chunks, states = _start(small_decks)
    with clv.Batch(chunks, cfgs, states, hist_len=64) as b:
        assert b.run(3) == 3
        sched, ms = b.autotune(reps=2)
        assert all(m > 0 for m in ms)
        assert all(8 <= f.seg <= 4096 for f in sched.fam)
        assert b.run(5, use_graph=False) == 5
        # members' own views are kept current between runs: a summary reads them
        s_mid = chunks[3].summary()
        assert np.isfinite(s_mid).all()
        b.run(100)
        _assert_same(b, chunks, cfgs, small_plain)


def test_batch_rejects_bad_use(gpu, small_decks):
    cfgs, chunks, states = _start(small_decks[:3])
    with pytest.raises(clv.CloverError, match="advect_x"):
        bad = list(states)
        bad[1] = clv.HydroState.from_buffer_copy(states[1])
        bad[1].advect_x = 0
        clv.Batch(chunks, cfgs, bad)
    with pytest.raises(clv.CloverError, match="twice"):
        clv.Batch([chunks[0], chunks[0]], [cfgs[0], cfgs[0]], [states[0], states[0]])
    with clv.Batch(chunks, cfgs, states) as b:
        s = b.schedule()
        s.fam[0].seg = 4
        with pytest.raises(clv.CloverError, match="seg"):
            b.set_schedule(s)
        b.run(2)
        # stepping a member outside the batch is detected at the next run
        st = clv.HydroState.from_buffer_copy(b.states()[1])
        chunks[1].step(cfgs[1], st, clv.PLAN_FUSED)
        with pytest.raises(clv.CloverError, match="outside the batch"):
            b.run(1)
    # in mesh mode the batch computes the members' volumes and areas from their meshes:
    # once a member's arrays may have changed (a view was taken) it refuses to run
    cfgs, chunks, states = _start(small_decks[:2], clv.GEOMETRY_MESH)
    with clv.Batch(chunks, cfgs, states) as b:
        b.run(1)
        chunks[0].field_ptr("volume")
        with pytest.raises(clv.CloverError, match="face areas may have changed"):
            b.run(1)


def test_batch_small_dt_member_stops(gpu, small_decks, small_plain):
    decks = small_decks[:4]
    cfgs, chunks, states = _start(decks)
    cfgs[2].dtmin = 1.0  # every dt is below this: the member fails at step 1
    with clv.Batch(chunks, cfgs, states, hist_len=64) as b:
        with pytest.raises(clv.SmallTimestep, match="member 2"):
            b.run(100)
        st = b.states()
        assert st[2].step == 1 and st[2].complete == 1 and st[2].time == 0.0
        others = [0, 1, 3]
        for p in others:
            assert st[p].step == small_plain[p][0].step and st[p].time == small_plain[p][0].time
    # the plain path fails the same way
    with clv.Chunk(cfgs[2].x_cells, cfgs[2].y_cells) as ch:
        s, _ = ch.start(cfgs[2])
        with pytest.raises(clv.SmallTimestep):
            ch.step(cfgs[2], s, clv.PLAN_FUSED)


def test_batch_seeded_workload_vs_reference(gpu):
    """The 64-problem seeded workload against the reference serial model."""
    g = json.loads(GOLDEN.read_text())
    assert g["seed"] == SEED and len(g["problems"]) == 64
    import hashlib
    decks = [clv.batch_deck(SEED, i) for i in range(64)]
    for i, t in enumerate(decks):
        assert hashlib.sha256(t.encode()).hexdigest() == g["problems"][i]["deck_sha256"], i
    print("# seeded workload: 64 problems, building", flush=True)
    cfgs, chunks, states = _start(decks)
    for ch in chunks:
        ch.set_summary_order(clv.SUMMARY_SERIAL)
    with clv.Batch(chunks, cfgs, states, hist_len=100) as b:
        b.autotune(reps=2)
        done_at = 0
        for want_step in (1, 50, 100):
            done_at += b.run(want_step - done_at)
            assert done_at == want_step
            print(f"# seeded workload: step {want_step}", flush=True)
            sts = b.states()
            for p, prob in enumerate(g["problems"]):
                rec = next(r for r in prob["summaries"] if r["step"] == want_step)
                assert sts[p].step == want_step
                assert sts[p].time == rec["time"], (p, want_step)
                s = chunks[p].summary()
                want = [rec[k] for k in ("vol", "mass", "ie", "ke", "press")]
                assert s == want, (p, want_step, s, want)
        for p, prob in enumerate(g["problems"]):
            assert b.dt_history(p).tolist() == prob["dt"], p
            assert b.states()[p].complete == 1
    for ch in chunks:
        ch.close()
