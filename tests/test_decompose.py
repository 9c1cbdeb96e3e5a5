"""Mesh and tile decompositions (cloverleaf_amd.decompose) against the reference's
own clover_decompose / clover_tile_decompose (driver/comms.cpp:64-253, compiled
into oracle/_ref/librefkern.so), no GPU.

The ranges must be the reference's.  Neighbours: the reference numbers chunks
from 1 (a rank + 1) and its tile code reuses that formula; the pieces here are
0-based indices, so each neighbour is the reference's number - 1.  The
reference's tile_bottom and tile_top share one slot (driver/definitions.h:58):
for tiles between the first and the last row that slot holds the top
neighbour, which is what can be compared; elsewhere it reads external.
"""
import ctypes as C

import numpy as np
import pytest

import clover_testlib as T
from cloverleaf_amd.decompose import chunk_decompose, tile_decompose

CASES = [(960, 960, n) for n in (1, 2, 3, 4, 5, 6, 8, 9, 12, 16)] + [(3840, 960, 6), (100, 400, 6), (61, 47, 4), (97, 13, 7),
                                                                         (13, 97, 12), (2048, 1024, 10)]


def _coverage(pieces, x_cells, y_cells):
    count = np.zeros((y_cells, x_cells), dtype=np.int32)
    for p in pieces:
        count[p.bottom - 1:p.top, p.left - 1:p.right] += 1
    assert (count == 1).all()


def _reciprocal(pieces):
    opposite = {0: 1, 1: 0, 2: 3, 3: 2}
    for k, p in enumerate(pieces):
        for f, n in enumerate(p.neighbours):
            if n >= 0:
                q = pieces[n]
                assert q.neighbours[opposite[f]] == k
                if f <= 1:
                    assert (q.bottom, q.top) == (p.bottom, p.top)
                    assert (q.right + 1 == p.left) if f == 0 else (p.right + 1 == q.left)
                else:
                    assert (q.left, q.right) == (p.left, p.right)
                    assert (q.top + 1 == p.bottom) if f == 2 else (p.top + 1 == q.bottom)


@pytest.mark.parametrize("x_cells,y_cells,n", CASES)
def test_chunk_decompose_matches_reference(x_cells, y_cells, n):
    pieces = chunk_decompose(x_cells, y_cells, n)
    assert len(pieces) == n
    _coverage(pieces, x_cells, y_cells)
    _reciprocal(pieces)
    L = T.refkern_lib()
    if L is None:
        pytest.skip("oracle/_ref/librefkern.so not built (no reference tree)")
    for task, p in enumerate(pieces):
        out = (C.c_int * 8)()
        assert L.refk_chunk_decompose(x_cells, y_cells, n, task, out) == 0
        assert (p.left, p.right, p.bottom, p.top) == tuple(out[:4]), task
        assert tuple(nb if nb < 0 else nb + 1 for nb in p.neighbours) == tuple(out[4:8]), task


@pytest.mark.parametrize("x_cells,y_cells,n", CASES)
@pytest.mark.parametrize("left,bottom", [(1, 1), (41, 7)])
def test_tile_decompose_matches_reference(x_cells, y_cells, n, left, bottom):
    tiles = tile_decompose(x_cells, y_cells, n, left, bottom)
    assert len(tiles) == n
    for p in tiles:
        assert left <= p.left and p.right <= left + x_cells - 1 and bottom <= p.bottom and p.top <= bottom + y_cells - 1
    _coverage([type(p)(p.left - left + 1, p.right - left + 1, p.bottom - bottom + 1, p.top - bottom + 1, p.neighbours) for p in tiles],
              x_cells, y_cells)
    _reciprocal(tiles)
    L = T.refkern_lib()
    if L is None:
        pytest.skip("oracle/_ref/librefkern.so not built (no reference tree)")
    out = (C.c_int * (12 * n))()
    assert L.refk_tile_decompose(x_cells, y_cells, n, left, bottom, out) == 0
    rows = sorted({p.bottom for p in tiles})
    for k, p in enumerate(tiles):
        r = out[12 * k:12 * k + 12]
        assert (p.left, p.right, p.bottom, p.top) == tuple(r[:4]), k
        for f in (0, 1):  # left / right: the reference's number - 1, or external (-1) with its mask set
            assert p.neighbours[f] == (r[4 + f] - 1 if r[4 + f] != -1 else -1), (k, f)
            assert r[8 + f] == (1 if p.neighbours[f] < 0 else 0)
        if p.bottom not in (rows[0], rows[-1]):  # the shared slot holds the top neighbour
            assert p.neighbours[3] == r[7] - 1, k


def test_decomposed_oracle_on_one_piece_is_the_oracle():
    """The decomposed replay of the GPU tests (clover_testlib.DecomposedOracle) on a
    single piece: the oracle's own step, bit for bit (every face external)."""
    import cloverleaf_amd as clv
    deck = T.parse_deck(clv.bm_deck(24, 20))
    single = T.Oracle(deck)
    single.start()
    arrays = [({k: single.f(k).copy() for k in T.FIELDS}, {k: single.m(k).copy() for k in T.MESH})]
    d = T.DecomposedOracle(deck, tile_decompose(24, 20, 1), arrays)
    for _ in range(6):
        single.step()
        dt = d.step()
        assert dt == single.g.dt
        for k in T.FIELDS[:15]:
            assert np.array_equal(d.o[0].f(k), single.f(k)), k
