"""The adjoint form on the GPU (clv_adj_*, clv_adjoint_*).

Pinned by:
- every adjoint kernel against the oracle's tangent-linear model (forward-mode
  dual numbers, itself pinned against central differences in
  test_oracle_tl.py) through the dot-product identity
      <ybar, J xdot> = <J^T ybar, xdot>
  over all arrays, on random physically shaped inputs (both upwind branches,
  limiter and viscosity branches taken) and random directions;
- whole checkpointed runs: the gradient's directional derivative equals the
  tangent-linear oracle's d ke / d(direction), and central finite differences of
  GPU forward runs, on random smooth directions;
- checkpoint spacing does not change the gradient.
"""
import numpy as np
import pytest

import clover_testlib as T
import cloverleaf_amd as clv

pytestmark = pytest.mark.gpu

DT = 0.004
WORK = {f"work{i}" for i in range(1, 8)}
GEOM = {"volume", "xarea", "yarea"}
# kernel name -> (oracle tangent-linear runner, adjoint args)
KERNELS = [("ideal_gas", (0,)), ("ideal_gas", (1,)), ("viscosity", ()), ("PdV", (1, DT)), ("PdV", (0, DT)), ("revert", ()),
           ("accelerate", (DT,)), ("flux_calc", (DT,)), ("advec_cell", (1, 1)), ("advec_cell", (1, 2)), ("advec_cell", (2, 1)),
           ("advec_cell", (2, 2)), ("advec_mom", (1, 1)), ("advec_mom", (1, 2)), ("advec_mom", (2, 1)), ("advec_mom", (2, 2)),
           ("reset_field", ())]
SIZES = [(13, 11), (37, 29)]


def _tl(t, kernel, args):
    if kernel == "advec_mom":  # both velocities of one sweep
        t.run("advec_mom", 1, *args)
        t.run("advec_mom", 2, *args)
    else:
        t.run(kernel, *args)


def _chunk(nx, ny, f, m):
    ch = clv.Chunk(nx, ny)
    for n in T.FIELDS:
        ch.upload(n, f[n])
    for n in T.MESH:
        ch.upload_mesh(n, m[n])
    return ch


def _dot(a, b, names):
    return sum(float(np.sum(a[n] * b[n])) for n in names)


def _dot_test(nx, ny, seed, kernel, args):
    f, m = T.random_chunk_arrays(nx, ny, seed)
    rng = np.random.default_rng(seed + 7)
    shape = (ny + 5, nx + 5)
    xdot = {n: (np.zeros(shape) if n in WORK | GEOM | {"soundspeed"} else rng.uniform(-1, 1, shape)) for n in T.FIELDS}
    ybar = {n: (np.zeros(shape) if n in WORK | {"soundspeed"} else rng.uniform(-1, 1, shape)) for n in T.FIELDS}
    # J xdot by the tangent-linear oracle
    t = T.TLOracle.from_oracle(T.oracle_on_arrays(nx, ny, f, m))
    for n in T.FIELDS:
        t.der(n)[...] = xdot[n]
    _tl(t, kernel, args)
    jx = {n: t.der(n).copy() for n in T.FIELDS}
    # J^T ybar on the GPU
    p = _chunk(nx, ny, f, m)
    a = _chunk(nx, ny, ybar, m)
    clv.adjoint_kernel(kernel, p, a, *args)
    xbar = {n: a.download(n) for n in T.FIELDS}
    names = [n for n in T.FIELDS if n not in WORK]
    lhs, rhs = _dot(ybar, jx, names), _dot(xbar, xdot, names)
    scale = sum(float(np.sum(np.abs(ybar[n] * jx[n]))) for n in names)
    return lhs, rhs, scale


@pytest.mark.parametrize("nx,ny", SIZES)
@pytest.mark.parametrize("kernel,args", KERNELS, ids=[f"{k}{a[:2]}" for k, a in KERNELS])
def test_adjoint_kernel_dot_product(gpu, nx, ny, kernel, args):
    lhs, rhs, scale = _dot_test(nx, ny, 1000 + nx, kernel, args)
    assert abs(lhs - rhs) <= 1e-11 * scale, (lhs, rhs, scale)


@pytest.mark.parametrize("depth", [1, 2])
def test_adjoint_halo_dot_product(gpu, depth):
    nx, ny = 17, 12
    f, m = T.random_chunk_arrays(nx, ny, 3)
    rng = np.random.default_rng(4)
    shape = (ny + 5, nx + 5)
    names = [n for n in T.FIELDS[:15] if n != "soundspeed"]
    xdot = {n: rng.uniform(-1, 1, shape) for n in names}
    ybar = {n: rng.uniform(-1, 1, shape) for n in names}
    o = T.oracle_on_arrays(nx, ny, {**f, **xdot}, m)
    import ctypes as C
    o.L.orc_update_halo(C.byref(o.c), T.mask(names), depth)  # the halo is linear: J xdot = halo(xdot)
    jx = {n: o.f(n).copy() for n in names}
    a = _chunk(nx, ny, {**f, **ybar}, m)
    clv.adjoint_kernel("update_halo", None, a, names, depth)
    xbar = {n: a.download(n) for n in names}
    lhs, rhs = _dot(ybar, jx, names), _dot(xbar, xdot, names)
    assert abs(lhs - rhs) <= 1e-12 * sum(float(np.sum(np.abs(ybar[n] * jx[n]))) for n in names)


def test_adjoint_kinetic_energy_dot_product(gpu):
    nx, ny = 21, 16
    f, m = T.random_chunk_arrays(nx, ny, 9)
    rng = np.random.default_rng(10)
    shape = (ny + 5, nx + 5)
    xdot = {n: (rng.uniform(-1, 1, shape) if n in ("density0", "xvel0", "yvel0") else np.zeros(shape)) for n in T.FIELDS}
    t = T.TLOracle.from_oracle(T.oracle_on_arrays(nx, ny, f, m))
    for n in T.FIELDS:
        t.der(n)[...] = xdot[n]
    dke = t.summary(with_ideal_gas=False)[3][1]
    p = _chunk(nx, ny, f, m)
    a = _chunk(nx, ny, {n: np.zeros(shape) for n in T.FIELDS}, m)
    clv.adjoint_kernel("kinetic_energy", p, a, 0.75)
    rhs = sum(float(np.sum(a.download(n) * xdot[n])) for n in ("density0", "xvel0", "yvel0"))
    assert abs(0.75 * dke - rhs) <= 1e-12 * abs(rhs), (0.75 * dke, rhs)


def test_adjoint_rejects_bad_arguments(gpu):
    p, a, b = clv.Chunk(8, 8), clv.Chunk(8, 8), clv.Chunk(9, 8)
    with pytest.raises(clv.CloverError, match="differ"):
        clv.adjoint_kernel("viscosity", p, b)
    with pytest.raises(clv.CloverError, match="sweep_number"):
        clv.adjoint_kernel("advec_cell", p, a, 3, 1)
    with pytest.raises(clv.CloverError, match="depth"):
        clv.adjoint_kernel("update_halo", None, a, ["density0"], 3)


# ---- whole runs ---------------------------------------------------------------------------
def _deck(nx, ny):
    text = (T.DECKS / "clover_bm_short.in").read_text().replace("x_cells=960", f"x_cells={nx}").replace(
        "y_cells=960", f"y_cells={ny}")
    return text


def _smooth(nx, ny, seed, modes=4):
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    p = np.zeros((ny, nx))
    for _ in range(modes):
        kx, ky = rng.integers(1, 4, 2)
        p += rng.uniform(-1, 1) * np.outer(np.sin(np.pi * ky * y), np.sin(np.pi * kx * x))
    return p


def _started(nx, ny):
    """Two-state start with a smooth 1 % perturbation of density and energy (a
    generic point: piecewise-constant fields sit on the limiters' kinks)."""
    text = _deck(nx, ny)
    cfg = clv.parse_deck(text)
    ch = clv.Chunk(nx, ny)
    ch.start(cfg)
    rho = ch.download("density0")[2:ny + 2, 2:nx + 2] * (1 + 1e-2 * _smooth(nx, ny, 7))
    e = ch.download("energy0")[2:ny + 2, 2:nx + 2] * (1 + 1e-2 * _smooth(nx, ny, 8))
    return text, ch, rho, e


@pytest.mark.parametrize("nx,ny,steps,dt,every", [(40, 32, 12, 0.01, 5), (29, 23, 9, 0.02, 9)])
def test_run_gradient_vs_tangent_linear_and_differences(gpu, nx, ny, steps, dt, every):
    text, ch, rho, e = _started(nx, ny)
    with clv.Adjoint(ch, dt, steps, every) as g:
        g.set_initial(rho, e)
        s = g.forward()
        gr, ge = g.backward(1.0)
        # the tangent-linear oracle from the same initial state
        o = T.Oracle(T.parse_deck(text))
        o.start()
        o.f("density0")[2:ny + 2, 2:nx + 2] = rho
        o.f("energy0")[2:ny + 2, 2:nx + 2] = e
        o.L.orc_update_halo(__import__("ctypes").byref(o.c), T.mask(["density0", "energy0"]), 2)
        for seed in (1, 2, 3):
            dr, de = _smooth(nx, ny, 10 * seed), _smooth(nx, ny, 10 * seed + 1)
            t = T.TLOracle.from_oracle(o)
            t.der("density0")[2:ny + 2, 2:nx + 2] = dr
            t.der("energy0")[2:ny + 2, 2:nx + 2] = de
            adv = 1
            for _ in range(steps):
                t.run("step_fixed", dt, adv)
                adv = 1 - adv
            ke, dke = t.summary()[3]
            assert abs(ke - s[3]) <= 1e-13 * abs(ke)  # same trajectory (tree vs serial sum order)
            adj = float(np.sum(gr * dr) + np.sum(ge * de))
            assert abs(adj - dke) <= 1e-9 * abs(dke), (seed, adj, dke)
            # central differences of GPU forward runs
            eps = 1e-7
            g.set_initial(rho + eps * dr, e + eps * de)
            kp = g.forward()[3]
            g.set_initial(rho - eps * dr, e - eps * de)
            km = g.forward()[3]
            fd = (kp - km) / (2 * eps)
            assert abs(adj - fd) <= 1e-6 * abs(fd), (seed, adj, fd)
            g.set_initial(rho, e)
            g.forward()


def test_checkpoint_spacing_does_not_change_the_gradient(gpu):
    nx, ny, steps, dt = 24, 20, 10, 0.01
    _, ch, rho, e = _started(nx, ny)
    grads = []
    for every in (1, 3, 10):
        with clv.Adjoint(ch, dt, steps, every) as g:
            g.set_initial(rho, e)
            g.forward()
            grads.append(g.backward(1.0))
    for gr, ge in grads[1:]:
        assert np.array_equal(gr, grads[0][0]) and np.array_equal(ge, grads[0][1])


# ---- the adjoint workload against the oracle's tangent-linear sweep -------------------------
GOLDEN_C4 = clv.PKG.parent / "tests" / "golden" / "adjoint_c4.json"


@pytest.mark.skipif(not GOLDEN_C4.exists(), reason="adjoint golden not generated")
def test_adjoint_workload_vs_tangent_linear_golden(gpu):
    """1920^2, 200 fixed-dt steps, checkpoints every 20: the gradient along the
    golden's directions equals the oracle's forward-mode derivative."""
    import json
    g = json.loads(GOLDEN_C4.read_text())
    w = g["workload"]
    assert w == clv.ADJOINT_WORKLOAD
    nx, ny = w["nx"], w["ny"]
    ch = clv.Chunk(nx, ny)
    ch.start(clv.parse_deck(clv.bm_deck(nx, ny)))
    rho, e = clv.adjoint_initial(ch.download("density0")[2:ny + 2, 2:nx + 2], ch.download("energy0")[2:ny + 2, 2:nx + 2],
                                 w["amplitude"], w["seed"])
    print("# adjoint workload: forward", flush=True)
    with clv.Adjoint(ch, w["dt"], w["steps"], w["checkpoint_every"]) as adj:
        adj.set_initial(rho, e)
        s = adj.forward()
        assert abs(s[3] - g["ke"]) <= 1e-12 * abs(g["ke"])
        print("# adjoint workload: backward", flush=True)
        gr, ge = adj.backward(1.0)
    for rec in g["directions"]:
        dr, de = clv.smooth_field(nx, ny, rec["seed_density"]), clv.smooth_field(nx, ny, rec["seed_energy"])
        an = float(np.sum(gr * dr) + np.sum(ge * de))
        assert abs(an - rec["dke"]) <= 1e-9 * abs(rec["dke"]), (rec, an)
