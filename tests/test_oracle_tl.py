"""The oracle's tangent-linear model (orc_tl_*, forward-mode dual numbers), on the CPU.

Pinned by:
- values: the dual-number run carries exactly the oracle's values (bit-identical
  fields after every kernel and after whole fixed-dt steps);
- derivatives: every kernel's output derivatives against central finite
  differences of the oracle kernel on random, physically shaped inputs with a
  random direction for every input array (both upwind branches, limiter and
  viscosity branches taken); whole runs: the directional derivative of the
  final kinetic energy against central finite differences of oracle runs;
- linearity: the derivative of a sum of directions is the sum of derivatives.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import clover_testlib as T
import cloverleaf_amd as clv

SIZES = [(9, 7), (24, 19)]
DT = 0.004
# kernel, arguments of orc_<kernel> / orc_tl_<kernel>
KERNELS = [("ideal_gas", (0,)), ("ideal_gas", (1,)), ("viscosity", ()), ("PdV", (1, DT)), ("PdV", (0, DT)),
           ("revert", ()), ("accelerate", (DT,)), ("flux_calc", (DT,)),
           ("advec_cell", (1, 1)), ("advec_cell", (1, 2)), ("advec_cell", (2, 1)), ("advec_cell", (2, 2)),
           ("mom_pair", (1, 1)), ("mom_pair", (1, 2)), ("mom_pair", (2, 1)), ("mom_pair", (2, 2)),
           ("reset_field", ())]


def _run_double(o, kernel, args):
    import ctypes as C
    cp = C.byref(o.c)
    if kernel == "mom_pair":
        o.L.orc_advec_mom(cp, 1, *args)
        o.L.orc_advec_mom(cp, 2, *args)
    else:
        getattr(o.L, f"orc_{kernel}")(cp, *args)


def _run_tl(t, kernel, args):
    if kernel == "mom_pair":
        t.run("advec_mom", 1, *args)
        t.run("advec_mom", 2, *args)
    else:
        t.run(kernel, *args)


def _direction(nx, ny, seed):
    rng = np.random.default_rng(seed)
    return {n: rng.uniform(-1.0, 1.0, (ny + 5, nx + 5)) for n in T.FIELDS}


def _outputs(nx, ny, f, m, kernel, args):
    o = T.oracle_on_arrays(nx, ny, f, m)
    _run_double(o, kernel, args)
    return {n: o.f(n).copy() for n in T.FIELDS}


@pytest.mark.parametrize("nx,ny", SIZES)
@pytest.mark.parametrize("kernel,args", KERNELS, ids=[f"{k}{a[:2]}" for k, a in KERNELS])
def test_kernel_derivatives_vs_central_differences(nx, ny, kernel, args):
    seed = hash((kernel, args, nx)) % 2**32
    f, m = T.random_chunk_arrays(nx, ny, seed)
    xdot = _direction(nx, ny, seed + 1)
    # tangent-linear run
    o = T.oracle_on_arrays(nx, ny, f, m)
    t = T.TLOracle.from_oracle(o)
    for n in T.FIELDS:
        t.der(n)[...] = xdot[n]
    _run_tl(t, kernel, args)
    # values are the oracle's
    ref = _outputs(nx, ny, f, m, kernel, args)
    for n in T.FIELDS:
        assert np.array_equal(t.val(n), ref[n]), n
    # derivatives against central differences
    eps = 1e-7
    fp = _outputs(nx, ny, {n: f[n] + eps * xdot[n] for n in T.FIELDS}, m, kernel, args)
    fm = _outputs(nx, ny, {n: f[n] - eps * xdot[n] for n in T.FIELDS}, m, kernel, args)
    for n in T.FIELDS:
        fd = (fp[n] - fm[n]) / (2 * eps)
        scale = max(1.0, float(np.abs(fd).max()))
        bad = ~np.isclose(t.der(n), fd, rtol=1e-6, atol=1e-6 * scale)
        assert not bad.any(), (n, np.argwhere(bad)[:3].tolist(), t.der(n)[bad][:3], fd[bad][:3])


def test_halo_derivative_is_the_reflection():
    nx, ny = 11, 9
    f, m = T.random_chunk_arrays(nx, ny, 5)
    o = T.oracle_on_arrays(nx, ny, f, m)
    t = T.TLOracle.from_oracle(o)
    xdot = _direction(nx, ny, 6)
    for n in T.FIELDS:
        t.der(n)[...] = xdot[n]
    names = T.FIELDS[:15]
    t.update_halo(names, 2)
    # the halo is linear: derivatives transform exactly like values
    o2 = T.oracle_on_arrays(nx, ny, xdot, m)
    o2.L.orc_update_halo(__import__("ctypes").byref(o2.c), T.mask(names), 2)
    for n in names:
        assert np.array_equal(t.der(n), o2.f(n)), n


def _two_state(nx, ny):
    text = (T.DECKS / "clover_bm_short.in").read_text().replace("x_cells=960", f"x_cells={nx}").replace(
        "y_cells=960", f"y_cells={ny}")
    return T.parse_deck(text)


def _perturbation(nx, ny, seed):
    """A smooth field on the interior cells: a few low sine modes."""
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    p = np.zeros((ny, nx))
    for _ in range(4):
        kx, ky = rng.integers(1, 4, 2)
        p += rng.uniform(-1, 1) * np.outer(np.sin(np.pi * ky * y), np.sin(np.pi * kx * x))
    return p


def _ke_after(o_start, steps, dt, drho=None, de=None):
    o = T.Oracle(o_start.deck)
    for n in T.FIELDS:
        o.f(n)[...] = o_start.f(n)
    for n in T.MESH:
        o.m(n)[...] = o_start.m(n)
    nx, ny = o.nx, o.ny
    if drho is not None:
        o.f("density0")[2:ny + 2, 2:nx + 2] += drho
        o.f("energy0")[2:ny + 2, 2:nx + 2] += de
    adv = 1
    for _ in range(steps):
        T.oracle_step_fixed(o, dt, adv)
        adv = 1 - adv
    return o.summary()[3]


def _started(nx, ny):
    """The two-state problem with a smooth 1 % perturbation of density and energy:
    a generic point (piecewise-constant fields sit exactly on the limiters' kinks,
    where central differences see one-sided slopes)."""
    o = T.Oracle(_two_state(nx, ny))
    o.start()
    o.f("density0")[2:ny + 2, 2:nx + 2] *= 1.0 + 1e-2 * _perturbation(nx, ny, 7)
    o.f("energy0")[2:ny + 2, 2:nx + 2] *= 1.0 + 1e-2 * _perturbation(nx, ny, 8)
    return o


@pytest.mark.parametrize("nx,ny,steps,dt", [(24, 20, 12, 0.02), (33, 17, 20, 0.01)])
def test_run_derivative_of_kinetic_energy_vs_central_differences(nx, ny, steps, dt):
    o = _started(nx, ny)
    drho, de = _perturbation(nx, ny, 1), _perturbation(nx, ny, 2)
    # tangent-linear run
    t = T.TLOracle.from_oracle(o)
    t.der("density0")[2:ny + 2, 2:nx + 2] = drho
    t.der("energy0")[2:ny + 2, 2:nx + 2] = de
    adv = 1
    for _ in range(steps):
        t.run("step_fixed", dt, adv)
        adv = 1 - adv
    ke, dke = t.summary()[3]
    assert ke == _ke_after(o, steps, dt)
    # a step of 1e-7 keeps the +-eps runs on one side of every limiter kink here
    # (at 1e-6 a few cells cross one by step 12); round-off is ~1e-9 relative
    eps = 1e-7
    fd = (_ke_after(o, steps, dt, eps * drho, eps * de) - _ke_after(o, steps, dt, -eps * drho, -eps * de)) / (2 * eps)
    assert abs(dke - fd) <= 1e-6 * abs(fd), (dke, fd)


def test_linearity_in_the_direction():
    nx, ny, steps, dt = 20, 16, 6, 0.02
    o = _started(nx, ny)
    dirs = [(_perturbation(nx, ny, s), _perturbation(nx, ny, s + 10)) for s in (3, 4)]

    def dke(dr, de):
        t = T.TLOracle.from_oracle(o)
        t.der("density0")[2:ny + 2, 2:nx + 2] = dr
        t.der("energy0")[2:ny + 2, 2:nx + 2] = de
        adv = 1
        for _ in range(steps):
            t.run("step_fixed", dt, adv)
            adv = 1 - adv
        return t.summary()[3][1]

    a, b = dke(*dirs[0]), dke(*dirs[1])
    ab = dke(dirs[0][0] + dirs[1][0], dirs[0][1] + dirs[1][1])
    assert abs(ab - (a + b)) <= 1e-12 * (abs(a) + abs(b))


def test_fd_kink_fixture_matches_the_adjoint_workload():
    """tests/golden/adjoint_fd_kinks.json (make_adjoint_kinks.py) is the oracle's
    tangent-linear derivative and kink census for the 32 directions bench.py's
    adjoint leg checks: same workload, same seeds, the leg's central-difference step."""
    g = json.loads((Path(__file__).parent / "golden" / "adjoint_fd_kinks.json").read_text())
    assert g["workload"] == clv.ADJOINT_WORKLOAD
    assert g["fd_eps"] == 1e-8
    assert [r["k"] for r in g["directions"]] == list(range(32))
    for r in g["directions"]:
        assert (r["seed_density"], r["seed_energy"]) == (300 + r["k"], 400 + r["k"])
        assert np.isfinite(r["dke"]) and r["kinks"] > 0


def test_fd_scan_fixture_accounts_for_the_outliers():
    """tests/golden/adjoint_fd_scan.json (make_adjoint_fd_scan.py): the oracle's central
    differences of the directions whose eps = 1e-8 difference misses 1e-6, with ke
    summed exactly, over eps 1e-9 .. 1e-6.  Their error does not shrink with eps from
    1e-6 to 1e-8 (switch crossings, not rounding) and is below 1e-6 at eps = 1e-9."""
    g = json.loads((Path(__file__).parent / "golden" / "adjoint_fd_scan.json").read_text())
    assert g["workload"] == clv.ADJOINT_WORKLOAD
    kinks = json.loads((Path(__file__).parent / "golden" / "adjoint_fd_kinks.json").read_text())
    tl = {r["k"]: r["dke"] for r in kinks["directions"]}
    by = {}
    for r in g["records"]:
        assert r["tangent_linear"] == tl[r["k"]]
        assert r["rel_diff_exact_sum"] == pytest.approx(abs(r["fd_exact_sum"] - tl[r["k"]]) / abs(tl[r["k"]]))
        by.setdefault(r["k"], {})[r["eps"]] = r["rel_diff_exact_sum"]
    for k in (25, 28):
        e = by[k]
        assert e[1e-9] < 1e-6
        assert e[1e-8] > 1e-6 and max(e[1e-7], e[1e-6]) > 1e-6
