"""Throughput of the MI355X CloverLeaf hydro step on one GPU.

One step = one pass of driver/hydro.cpp:57-89 (timestep with the calc_dt
minimum, PdV predict/correct, accelerate, flux_calc, both advection sweeps,
reset_field) over the whole chunk, everything resident in HBM, with the
field_summary of driver/hydro.cpp:92-94 every summary_frequency steps inside the
timed loop (the reference's grind time includes it).  Default workload:
InputDecks/clover_bm256_short.in (15360 x 15360 cells, the two-state ideal-gas
box, FP64), the volume and face areas read from their arrays.  Metric:
cell-steps/s = x_cells * y_cells * steps / wall time (driver/hydro.cpp:213-217).

Prints ONE JSON line.  `roofline` is the dominant kernel (largest share of the
timed region by HIP events on the chunk's stream): achieved = its algorithmic
bytes per launch / its mean launch duration, against 8 TB/s.
`geometry_from_mesh` (unless --no-mesh) is the same run with the geometry
computed from the 1D spacings, a separately labelled figure.  `dropin` (unless
--no-dropin) runs the reference's own driver with the MI355X plugin model on the
same deck trimmed to --dropin-steps and reports the driver's own grind time.
Unless --no-probe, `pattern_ceiling` times each fused kernel's access pattern
without arithmetic (clv_march_probe: the same march, arrays read and written),
the ceiling of that kernel on this GPU.  `cpu_baseline` is the reference's own
omp model (oracle/_ref/omp-cloverleaf, compiled from the reference's sources) on
a bounded sample, with the OpenMP CPU oracle on the same sample beside it
(`cpu_baseline.port`).

`batched` (unless --no-batch) runs the seeded 64-problem workload (decks from
clv_batch_deck, seed 355, 100 steps each) one problem at a time on the plain
path and as one batch, and reports both times.

`adjoint` (unless --no-adjoint) times the checkpointed gradient of the final
kinetic energy of a 200-step fixed-dt run on 1920^2 (cloverleaf_amd.ADJOINT_WORKLOAD)
against its forward pass, and checks it against central differences.

`decomposed` (unless --no-decomposed) runs clover_bm16_short (3840^2) as one
chunk, as 4 tiles and as 4 chunks on the one GPU, 50 steps each.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--plan fused|reference]
                       [--deck NAME] [--no-cpu] [--no-mesh] [--no-dropin] [--no-batch]
                       [--batch-only] [--no-adjoint] [--adjoint-only]
                       [--adjoint-directions D] [--no-decomposed] [--decomposed-only]
                       [--profile-out FILE]
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, public spec

# Algorithmic bytes per launch: distinct FP64 2D arrays each launch reads or
# writes once over its own iteration range (reference kernel granularity,
# src/serial/<kernel>.cpp loop ranges), x 8 bytes.  Halos, 1D mesh arrays and
# reduction partials are excluded.  Extents: C = nx*ny cells, N = (nx+1)(ny+1)
# nodes, G = (nx+4)(ny+4) with ghosts, X = (nx+2)*ny / nx*(ny+2) flux faces,
# NX = (nx+4)(ny+1) node-flux range, NM = (nx+3)(ny+1) node-mass range,
# MF = (nx+2)(ny+1) momentum-flux range.
def _extents(nx, ny):
    return {"C": nx * ny, "N": (nx + 1) * (ny + 1), "G": (nx + 4) * (ny + 4), "X": (nx + 2) * ny,
            "Y": nx * (ny + 2), "NX": (nx + 4) * (ny + 1), "NY": (nx + 1) * (ny + 4), "NMX": (nx + 3) * (ny + 1),
            "NMY": (nx + 1) * (ny + 3), "MFX": (nx + 2) * (ny + 1), "MFY": (nx + 1) * (ny + 2)}


# name -> list of (arrays, extent) terms per launch, counted with the volume and face
# areas read from their arrays (the default geometry mode)
ALG_TERMS = {
    "ideal_gas": [(4, "C")],                   # rho, e -> p, c
    "viscosity": [(5, "C")],                   # u0, v0, p, rho0 -> q
    "calc_dt": [(8, "C")],                     # c, q, rho0, u0, v0, xarea, yarea, vol
    "PdV_predict": [(11, "C")],                # areas, vol, rho0, e0, p, q, u0, v0 -> rho1, e1
    "PdV_correct": [(13, "C")],                # + u1, v1
    "revert": [(4, "C")],
    "accelerate": [(10, "N")],                 # rho0, vol, areas, p, q, u0, v0 -> u1, v1
    "flux_calc": [(8, "N")],                   # areas, u0, v0, u1, v1 -> vfx, vfy
    "reset_field": [(4, "C"), (4, "N")],
    "advec_cell_vol": [(5, "G")],              # vol, vfx, vfy -> pre, post (sweep 2: 4)
    "advec_cell_flux_x": [(6, "X")],           # vfx, pre, rho1, e1 -> mfx, ener_flux
    "advec_cell_flux_y": [(6, "Y")],
    "advec_cell_update_x": [(8, "C")],         # rho1, e1, pre, mf, ef, vf -> rho1, e1
    "advec_cell_update_y": [(8, "C")],
    "advec_mom_vol": [(5, "G")],               # vol, vfx, vfy -> post, pre (sweep 2: 4)
    "advec_mom_node_flux_x": [(2, "NX")],
    "advec_mom_node_flux_y": [(2, "NY")],
    "advec_mom_node_mass_x": [(5, "NMX")],     # rho1, post_vol, node_flux -> post, pre
    "advec_mom_node_mass_y": [(5, "NMY")],
    "advec_mom_flux_x": [(4, "MFX")],          # node_flux, mass_pre, vel -> mom_flux
    "advec_mom_flux_y": [(4, "MFY")],
    "advec_mom_update_x": [(5, "N")],          # vel, pre, mom_flux, post -> vel
    "advec_mom_update_y": [(5, "N")],
    "field_summary": [(6, "C")],               # vol, rho0, e0, p, u0, v0
    "field_summary_gas": [(7, "C")],           # vol, rho0, e0, u0, v0 -> p, c (ideal_gas in the same pass)
    # fused plan: compulsory traffic of each fused kernel (every input array read once,
    # every output written once over the interior)
    "f_timestep": [(7, "C")],                  # rho0, e0, u0, v0, vol, xarea, yarea (q, p, c not stored)
    "f_timestep_q": [(8, "C")],                # the same -> q (the per-entry-point path)
    "f_timestep_pc": [(10, "C")],              # the same -> q, p, c (the per-entry-point path)
    "f_lagrange_pc": [(16, "C")],              # rho0, e0, q, u0, v0, vol, areas -> f_lagrange's + p', c' (the per-entry-point path)
    "f_lagrange": [(13, "C")],                 # rho0, e0, u0, v0, vol, xarea, yarea -> u1, v1, rho1, e1, vfx, vfy (q recomputed)
    "f_cm_x1": [(11, "C")],                    # vol, vfx, vfy, rho1, e1, u1, v1 -> rho1, e1, u1, v1 (mass flux not stored)
    "f_cm_y1": [(11, "C")],                    # vol, vfx, vfy, rho1, e1, u1, v1 -> rho1, e1, u1, v1
    "f_cm_x2": [(10, "C")],                    # vol, vfx, rho1, e1, u1, v1 -> rho1, e1, u1, v1
    "f_cm_y2": [(10, "C")],                    # vol, vfy, rho1, e1, u1, v1 -> rho1, e1, u1, v1
    "f_cm_x1_mf": [(12, "C")],                 # the same -> + mfx (the per-entry-point path)
    "f_cm_y1_mf": [(12, "C")],
    "f_cm_x2_mf": [(11, "C")],
    "f_cm_y2_mf": [(11, "C")],
}
# of those, the volume / face-area arrays a kernel does not stream when the chunk's
# geometry comes from the mesh (CLV_GEOMETRY_MESH), on the first term's extent
GEOM_ARRAYS = {"calc_dt": 3, "PdV_predict": 3, "PdV_correct": 3, "accelerate": 3, "flux_calc": 2, "advec_cell_vol": 1,
               "advec_mom_vol": 1, "field_summary": 1, "field_summary_gas": 1, "f_timestep": 3, "f_timestep_q": 3, "f_timestep_pc": 3,
               "f_lagrange": 3, "f_lagrange_pc": 3, "f_cm_x1": 1, "f_cm_y1": 1, "f_cm_x2": 1, "f_cm_y2": 1, "f_cm_x1_mf": 1,
               "f_cm_y1_mf": 1, "f_cm_x2_mf": 1, "f_cm_y2_mf": 1}
# compulsory traffic of one fused step, array sweeps: timestep 7 + lagrange 13 + the
# first advection pass 11 + the second 10 (geometry read); 4 + 10 + 10 + 9 from the mesh
B_FUSED_PER_CELL_STEP = {False: 41 * 8.0, True: 33 * 8.0}
B_ALG_PER_CELL_STEP = 1440.0  # 180 array sweeps of 8 B per interior cell-step at reference kernel granularity


# arrays read and written per interior cell by each fused kernel (the march probe's shape)
PATTERN = {False: {"f_timestep": (7, 0), "f_lagrange": (7, 6), "f_cm_x1": (7, 4), "f_cm_y1": (7, 4), "f_cm_x2": (6, 4),
                   "f_cm_y2": (6, 4)},
           True: {"f_timestep": (4, 0), "f_lagrange": (4, 6), "f_cm_x1": (6, 4), "f_cm_y1": (6, 4), "f_cm_x2": (5, 4),
                  "f_cm_y2": (5, 4)}}


def alg_bytes(name, nx, ny, geom_mesh=False):
    ext = _extents(nx, ny)
    terms = ALG_TERMS.get(name)
    if terms is None:
        return None
    total = sum(k * ext[e] for k, e in terms)
    if geom_mesh and name in GEOM_ARRAYS:
        total -= GEOM_ARRAYS[name] * ext[terms[0][1]]
    return total * 8.0


def cpu_baseline(deck_text, steps):
    """The OpenMP CPU oracle (oracle/liboracle.so, this repository's restatement of the
    reference) on a bounded sample: same problem (deck geometry and states) on a
    3840^2 grid, `steps` steps."""
    sys.path.insert(0, str(REPO / "tests"))
    import clover_testlib as T
    d = T.parse_deck(deck_text)
    d.x_cells = d.y_cells = 3840
    # the 1/100-cell state shrink depends on the cell count: re-read with the sample size
    text = deck_text.replace(f"x_cells={T.parse_deck(deck_text).x_cells}", "x_cells=3840").replace(
        f"y_cells={T.parse_deck(deck_text).y_cells}", "y_cells=3840")
    d = T.parse_deck(text)
    o = T.Oracle(d)
    o.start()
    o.step()  # first-touch / thread start-up outside the timed sample
    t0 = time.perf_counter()
    for _ in range(steps):
        if o.step():
            raise RuntimeError("oracle: small timestep")
    dt = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    return {"value": d.x_cells * d.y_cells * steps / dt, "unit": "cell-steps/s", "cores": threads, "kind": "port",
            "sample": f"CPU oracle (OpenMP, {threads} threads), same two-state problem on 3840x3840 cells, "
                      f"steps 2..{steps + 1}, {dt:.1f} s"}


BATCH_SEED = 355


def batch_leg(clv, n=64, seed=BATCH_SEED, reps=3):
    """The seeded batch workload: n independent problems (decks from
    clv_batch_deck), each to end_step 100 with its own dt.  Times (a) the problems
    one at a time on the plain fused path, each with its host loop, and (b) all of
    them as one batch (persistent launch per phase, hipGraph per step, autotuned
    schedule); checks that both give every problem the same dt sequence."""
    decks = [clv.batch_deck(seed, i) for i in range(n)]
    cfgs = [clv.parse_deck(t) for t in decks]
    cell_steps = float(sum(c.x_cells * c.y_cells * c.end_step for c in cfgs))
    # (a) one at a time
    seq_s, seq_dts = 0.0, []
    for i, c in enumerate(cfgs):
        with clv.Chunk(c.x_cells, c.y_cells) as ch:
            st, _ = ch.start(c)
            ch.synchronize()
            dts = []
            t0 = time.perf_counter()
            while not st.complete:
                ch.step(c, st, clv.PLAN_FUSED)
                dts.append(st.dt)
            ch.synchronize()
            seq_s += time.perf_counter() - t0
            seq_dts.append(dts)
    print(f"# bench: batch workload one at a time: {seq_s:.3f} s", file=sys.stderr, flush=True)

    def started():
        chunks, states = [], []
        for c in cfgs:
            ch = clv.Chunk(c.x_cells, c.y_cells)
            st, _ = ch.start(c)
            chunks.append(ch)
            states.append(st)
        return chunks, states

    # (b) batched, graph replay; the schedule tuned first (untimed)
    chunks, states = started()
    b = clv.Batch(chunks, cfgs, states, hist_len=128)
    t0 = time.perf_counter()
    sched, fam_ms = b.autotune(reps=reps)
    tune_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    steps = b.run(1000, use_graph=True)
    bat_s = time.perf_counter() - t0
    same = all(b.dt_history(i).tolist() == seq_dts[i] for i in range(n))
    states_end = b.states()
    b.close()
    for ch in chunks:
        ch.close()
    print(f"# bench: batched (graph): {bat_s:.3f} s, dt identical: {same}", file=sys.stderr, flush=True)
    # (c) batched, direct launches (same schedule): the graph's share
    chunks, states = started()
    b = clv.Batch(chunks, cfgs, states, hist_len=0)
    b.set_schedule(sched)
    t0 = time.perf_counter()
    b.run(1000, use_graph=False)
    direct_s = time.perf_counter() - t0
    b.close()
    for ch in chunks:
        ch.close()
    sides = [c.x_cells for c in cfgs]
    return {
        "workload": f"{n} problems from seed {seed}: square grids, side uniform in {{256,384,...,2048}} "
                    f"(drawn: min {min(sides)}, max {max(sides)}), two-state ideal gas with a random state-2 rectangle, "
                    f"100 steps each with its own dt, FP64",
        "cell_steps": cell_steps,
        "one_at_a_time_s": seq_s,
        "batched_s": bat_s,
        "batched_direct_launch_s": direct_s,
        "speedup_vs_one_at_a_time": seq_s / bat_s,
        "batched_cell_steps_per_s": cell_steps / bat_s,
        "one_at_a_time_cell_steps_per_s": cell_steps / seq_s,
        "batch_steps": steps,
        "all_complete": all(s.complete for s in states_end),
        "dt_sequences_identical": same,
        "autotune_s": tune_s,
        "schedule": sched.as_dict(),
        "family_ms_per_launch": dict(zip(clv.BATCH_FAMILIES, [round(x, 4) for x in fam_ms])),
    }


def adjoint_leg(clv, directions=32, eps=1e-8):
    """The adjoint workload (cloverleaf_amd.ADJOINT_WORKLOAD: two-state problem on
    1920^2 with a seeded 1e-3 density perturbation, 200 fixed-dt steps, checkpoints
    every 20).  Times the forward pass, the checkpointed backward pass (gradient of
    the final kinetic energy w.r.t. the initial density and energy) and the
    product's fused forward of the same run; checks the gradient against central
    differences of GPU forward runs along `directions` seeded smooth directions."""
    w = clv.ADJOINT_WORKLOAD
    nx, ny, steps, dt = w["nx"], w["ny"], w["steps"], w["dt"]
    cfg = clv.parse_deck(clv.bm_deck(nx, ny))
    ch = clv.Chunk(nx, ny)
    ch.start(cfg)
    rho, e = clv.adjoint_initial(ch.download("density0")[2:ny + 2, 2:nx + 2], ch.download("energy0")[2:ny + 2, 2:nx + 2],
                                 w["amplitude"], w["seed"])
    adj = clv.Adjoint(ch, dt, steps, w["checkpoint_every"])
    adj.set_initial(rho, e)
    # warm-up: the first backward allocates the undo log and the adjoint arrays (after
    # the headline leg's 47 GB chunk is freed that alone adds ~0.2 s)
    adj.forward()
    adj.backward(1.0)
    t0 = time.perf_counter()
    s = adj.forward()
    fwd_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    gr, ge = adj.backward(1.0)
    bwd_s = time.perf_counter() - t0
    print(f"# bench: adjoint forward {fwd_s:.3f} s, backward {bwd_s:.3f} s", file=sys.stderr, flush=True)
    # the product's fused forward of the same fixed-dt run: dt = dtinit = dtmax and calc_dt never limiting
    fcfg = clv.parse_deck(clv.bm_deck(nx, ny))
    fcfg.dtinit = fcfg.dtmax = dt
    fcfg.dtc_safe = fcfg.dtu_safe = fcfg.dtv_safe = fcfg.dtdiv_safe = 1e30
    fcfg.end_step = steps
    def plain_forward(plan):
        with clv.Chunk(nx, ny) as fc:
            st, _ = fc.start(fcfg)
            for name, a in (("density0", rho), ("energy0", e)):
                full = fc.download(name)
                full[2:ny + 2, 2:nx + 2] = a
                fc.upload(name, full)
            fc.update_halo(["density0", "energy0"], 2)
            fc.synchronize()
            t0 = time.perf_counter()
            while not st.complete:
                fc.step(fcfg, st, plan)
            fc.synchronize()
            return time.perf_counter() - t0, fc.summary()[3]
    fused_s, ke_fused = plain_forward(clv.PLAN_FUSED)
    # and on the reference plan (one launch per reference loop), as the adjoint's forward ran before its
    # Lagrangian phase and advection sweeps became single passes
    ref_s, ke_ref = plain_forward(clv.PLAN_REFERENCE)
    # central differences along seeded smooth directions.  eps = 1e-8: over 200 steps
    # on 1920^2 a step of 1e-7 or more moves some van Leer limiter / upwind switch
    # across its kink, and the difference quotient then carries an O(1e-5) error
    # that does not shrink with eps (at 10-40 steps, 1e-6 and 1e-7 agree to 1e-8)
    rel, per_dir = [], []
    # the oracle's tangent-linear derivative and kink census of the same directions
    # (tests/golden/adjoint_fd_kinks.json, tests/golden/make_adjoint_kinks.py)
    tl = {}
    kinks_file = REPO / "tests" / "golden" / "adjoint_fd_kinks.json"
    if kinks_file.exists():
        g = json.loads(kinks_file.read_text())
        if g.get("fd_eps") == eps and g.get("workload") == {k: w[k] for k in g.get("workload", {})}:
            tl = {r["k"]: r for r in g["directions"]}
    for k in range(directions):
        dr, de = clv.smooth_field(nx, ny, 300 + k), clv.smooth_field(nx, ny, 400 + k)
        an = float(np.sum(gr * dr) + np.sum(ge * de))
        adj.set_initial(rho + eps * dr, e + eps * de)
        kp = adj.forward()[3]
        adj.set_initial(rho - eps * dr, e - eps * de)
        km = adj.forward()[3]
        fd = (kp - km) / (2 * eps)
        rel.append(abs(an - fd) / abs(fd))
        rec = {"k": k, "adjoint": an, "fd": fd, "fd_rel_diff": rel[-1]}
        if k in tl:
            t = tl[k]["dke"]
            rec.update({"tangent_linear": t, "adjoint_rel_diff_vs_tl": abs(an - t) / abs(t), "fd_rel_diff_vs_tl": abs(fd - t) / abs(t),
                        "kinks": tl[k]["kinks"]})
        per_dir.append(rec)
    adj.close()
    rel = np.array(rel)
    # the oracle's central differences of the same directions over a range of eps, with
    # ke summed exactly (tests/golden/adjoint_fd_scan.json, make_adjoint_fd_scan.py)
    scan = {}
    scan_file = REPO / "tests" / "golden" / "adjoint_fd_scan.json"
    if scan_file.exists():
        for r in json.loads(scan_file.read_text())["records"]:
            scan.setdefault(r["k"], {})[f"{r['eps']:g}"] = r["rel_diff_exact_sum"]
    return {
        "workload": f"two-state problem (InputDecks/clover_bm_short.in) on {nx}x{ny}, initial density x (1 + "
                    f"{w['amplitude']} s), s = seeded smooth modes; {steps} steps at fixed dt {dt} (0.004 diverges at "
                    f"this resolution); J = final kinetic energy; checkpoints every {w['checkpoint_every']} steps",
        "ke": s[3],
        "ke_fused_forward": ke_fused,
        "forward_s": fwd_s,
        "backward_s": bwd_s,
        "fused_forward_s": fused_s,
        "reference_plan_forward_s": ref_s,
        "ke_reference_plan_forward": ke_ref,
        "adjoint_cost_vs_forward": bwd_s / fwd_s,
        "adjoint_cost_vs_reference_plan_forward": bwd_s / ref_s,
        "adjoint_cost_vs_fused_forward": bwd_s / fused_s,
        "fd_directions": directions,
        "fd_eps": eps,
        "fd_max_rel_diff": float(rel.max()) if directions else None,
        "fd_median_rel_diff": float(np.median(rel)) if directions else None,
        "fd_within_1e-6": int(np.sum(rel <= 1e-6)),
        "tl_max_rel_diff": max((r["adjoint_rel_diff_vs_tl"] for r in per_dir if "tangent_linear" in r), default=None),
        "fd_kink_crossings": {
            "note": "kinks: piecewise choices of the oracle's tangent-linear run whose deciding difference x has |x| < "
                    "fd_eps |dx| -- taken differently by the +eps and -eps runs of the central difference (to first "
                    "order); tangent_linear: the oracle's exact directional derivative on the branches taken "
                    "(tests/golden/adjoint_fd_kinks.json)",
            "directions_above_1e-6": [{"k": r["k"], "fd_rel_diff": r["fd_rel_diff"], "kinks": r.get("kinks"),
                                       "adjoint_rel_diff_vs_tl": r.get("adjoint_rel_diff_vs_tl"),
                                       "oracle_fd_rel_diff_by_eps": scan.get(r["k"])}
                                      for r in per_dir if r["fd_rel_diff"] > 1e-6],
            "oracle_fd_scan_note": "oracle_fd_rel_diff_by_eps: the oracle's central difference of the same direction "
                                   "at each eps, ke summed exactly, against the tangent-linear derivative "
                                   "(tests/golden/adjoint_fd_scan.json): an error that does not shrink with eps and "
                                   "changes from one eps to the next comes from branch flips, not rounding",
            "min_kinks_per_direction": min((r["kinks"] for r in per_dir if "kinks" in r), default=None),
        },
        "directions": per_dir,
    }


def decomposed_leg(clv, deck="clover_bm16_short", pieces=4, steps=50, warmup=2):
    """The decomposed-domain path on one GPU: the same problem as one chunk, as
    `pieces` tiles of one chunk (tile-to-tile halo copies) and as `pieces` chunks of
    a mesh (packed halo messages, the reference's MPI layout), all on the
    reference-plan kernels that the decomposed driver runs per piece, plus the
    fused plan on one chunk.  Reports each run's rate, the decomposition's cost
    against the one-chunk reference plan, and how far its final state moved from
    the one-chunk run (advec_cell's upwind clamp at a piece's own edge, the
    reference's behaviour; bit-identical elsewhere)."""
    from cloverleaf_amd.decompose import Decomposed, chunk_decompose, tile_decompose
    text = (REPO / "tests" / "golden" / "decks" / f"{deck}.in").read_text()
    cfg = clv.parse_deck(text)
    nx, ny = cfg.x_cells, cfg.y_cells
    cfg.end_step, cfg.end_time = 1 << 30, 1e30
    out = {"workload": f"{deck}: {nx}x{ny}, {pieces} pieces, steps {warmup + 1}..{warmup + steps}", "steps": steps}

    def timed(step, sync):
        for _ in range(warmup):
            step()
        sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        sync()
        return time.perf_counter() - t0

    summaries = {}
    for plan_name, plan in (("one_chunk_reference_plan", clv.PLAN_REFERENCE), ("one_chunk_fused_plan", clv.PLAN_FUSED)):
        with clv.Chunk(nx, ny) as ch:
            st, _ = ch.start(cfg)
            wall = timed(lambda: ch.step(cfg, st, plan), ch.synchronize)
            summaries[plan_name] = ch.summary()
        out[plan_name] = {"s": wall, "cell_steps_per_s": nx * ny * steps / wall}
    for mode, split in (("tiles", tile_decompose), ("chunks", chunk_decompose)):
        with Decomposed(cfg, split(nx, ny, pieces), mode) as d:
            d.start()
            wall = timed(d.step, lambda: [c.synchronize() for c in d.chunks])
            s = d.summary()
        ref = summaries["one_chunk_reference_plan"]
        out[mode] = {"s": wall, "cell_steps_per_s": nx * ny * steps / wall,
                     "time_vs_one_chunk_reference_plan": wall / out["one_chunk_reference_plan"]["s"],
                     "ke_rel_diff_vs_one_chunk": abs(s[3] - ref[3]) / abs(ref[3]),
                     "mass_rel_diff_vs_one_chunk": abs(s[1] - ref[1]) / abs(ref[1])}
        print(f"# bench: decomposed {mode} {wall:.3f} s", file=sys.stderr, flush=True)
    return out


def headline_leg(clv, deck_text, plan, steps, warmup, geometry, profile_out):
    """The timed loop of driver/hydro.cpp:57-111 on one chunk: `warmup` untimed steps,
    then `steps` timed steps, each followed by field_summary when the step count is a
    multiple of the deck's summary_frequency (driver/hydro.cpp:92-94), as the
    reference's own grind time includes it.  Returns the rate and the per-kernel
    HIP-event times of the timed region."""
    cfg = clv.parse_deck(deck_text)
    nx, ny = cfg.x_cells, cfg.y_cells
    cfg.end_step = 1 << 30  # the bench decides the step count
    cfg.end_time = 1e30
    freq = cfg.summary_frequency
    gname = "arrays" if geometry == clv.GEOMETRY_ARRAYS else "mesh"
    print(f"# bench: {nx}x{ny} plan={plan} geometry={gname} warmup={warmup} steps={steps}", file=sys.stderr, flush=True)
    ch = clv.Chunk(nx, ny, geometry)
    st, _ = ch.start(cfg)
    assert ch.geometry_from_mesh() == (geometry == clv.GEOMETRY_MESH)
    for _ in range(warmup):
        ch.step(cfg, st, plan)
    ch.synchronize()
    ch.profile(True)
    ch.profile_reset()
    summaries = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        ch.step(cfg, st, plan)
        if freq and st.step % freq == 0:
            ch.summary()
            summaries += 1
    ch.synchronize()
    wall = time.perf_counter() - t0
    prof = ch.profile_read()
    ch.profile(False)
    s_end = ch.summary()
    ch.close()
    print(f"# bench: {steps} steps in {wall:.3f} s ({nx * ny * steps / wall / 1e9:.3f} Gcell-steps/s), last dt {st.dt:.6g}, "
          f"step {st.step}", file=sys.stderr, flush=True)
    kern = sorted(((ms, name, n) for name, (n, ms) in prof.items()), reverse=True)
    total_kernel_ms = sum(ms for ms, _, _ in kern)
    mesh = geometry == clv.GEOMETRY_MESH
    per_kernel = {}
    for ms, name, n in kern:
        b = alg_bytes(name, nx, ny, mesh)
        per_kernel[name] = {"launches": n, "total_ms": round(ms, 3), "share": round(ms / total_kernel_ms, 4),
                            "avg_us": round(ms / n * 1e3, 2),
                            "alg_GBps": round(b / (ms / n / 1e3) / 1e9, 1) if b else None}
    if profile_out:
        Path(profile_out).write_text(json.dumps({"plan": plan, "geometry": gname, "steps": steps, "wall_s": wall,
                                                 "kernels": per_kernel}, indent=1))
    return {"nx": nx, "ny": ny, "value": nx * ny * steps / wall, "wall_s": wall, "per_kernel": per_kernel,
            "dominant": kern[0][1], "kernel_ms": total_kernel_ms, "summaries": summaries, "summary_frequency": freq,
            "final_summary": {"step": st.step, "ke": s_end[3], "mass": s_end[1]}}


def _omp_env():
    env = dict(os.environ)
    threads = int(env.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    env.update({"OMP_NUM_THREADS": str(threads), "OMP_PLACES": "cores", "OMP_PROC_BIND": "true"})
    return env, threads


def _cpu_model():
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        info = dict(l.split(":", 1) for l in out.splitlines() if ":" in l)
        return {"model": info.get("Model name", "").strip(), "sockets": info.get("Socket(s)", "").strip(),
                "cores_per_socket": info.get("Core(s) per socket", "").strip(),
                "threads_per_core": info.get("Thread(s) per core", "").strip()}
    except Exception:
        return {}


def cpu_baseline_reference(steps, deck="clover_bm16_300"):
    """The reference's own omp model (driver/ + src/omp/ compiled from the reference's
    sources by oracle/Makefile into oracle/_ref/omp-cloverleaf) on a bounded sample:
    InputDecks/clover_bm16_300.in (3840^2) trimmed to `steps` steps, OMP_PLACES=cores,
    OMP_PROC_BIND=true.  The rate is the reference's own printed `Average time per cell`
    (driver/hydro.cpp:212-215).  None where the binary was not built."""
    import re
    import subprocess
    import tempfile
    exe = REPO / "oracle" / "_ref" / "omp-cloverleaf"
    if not exe.exists():
        return None
    text = (REPO / "tests" / "golden" / "decks" / f"{deck}.in").read_text()
    text = re.sub(r"end_step=\d+", f"end_step={steps}", text)
    env, threads = _omp_env()
    with tempfile.TemporaryDirectory() as td:
        (Path(td) / "deck.in").write_text(text)
        t0 = time.perf_counter()
        p = subprocess.run([str(exe), "--in", "deck.in", "--out", "clover.out"], cwd=td, env=env, capture_output=True,
                           text=True, timeout=600)
        whole = time.perf_counter() - t0
    walls = re.findall(r"^ Wall clock\s+([0-9.eE+-]+)", p.stdout, re.M)
    done = re.findall(r"^ Step (\d+) time", p.stdout, re.M)
    if p.returncode != 0 or not walls or not done or int(done[-1]) != steps:
        return {"value": None, "error": (p.stderr or p.stdout)[-500:]}
    wall = float(walls[-1])
    grind = wall / (steps * 3840.0 * 3840.0)
    return {"value": 1.0 / grind, "unit": "cell-steps/s", "cores": threads, "kind": "reference",
            "sample": f"reference omp model (oracle/_ref/omp-cloverleaf, built from the reference's sources), "
                      f"{deck}.in = 3840x3840 trimmed to {steps} steps, {threads} threads, OMP_PLACES=cores "
                      f"OMP_PROC_BIND=true; grind time from its final hydro 'Wall clock' {wall:.2f} s: {grind:.5g} s "
                      f"per cell-step; process {whole:.1f} s",
            "cpu": _cpu_model()}


def dropin_leg(deck_text, steps):
    """The drop-in as it runs by default (the model's chains of entry points as one
    launch each), with one launch per entry point (CLOVER_MI355X_CHAINS=off), and with
    one launch per reference loop (CLOVER_MI355X_ENTRY=reference)."""
    out = dropin_run(deck_text, steps, {})
    out["one_launch_per_entry_point"] = dropin_run(deck_text, steps, {"CLOVER_MI355X_CHAINS": "off"})
    out["one_launch_per_reference_loop"] = dropin_run(deck_text, steps, {"CLOVER_MI355X_ENTRY": "reference"})
    return out


def dropin_run(deck_text, steps, env_extra):
    """The drop-in: the reference's unchanged driver (driver/*.cpp) linked with the MI355X
    plugin model (cloverleaf_amd/model_mi355x/_dropin/mi355x-cloverleaf) on the headline
    deck trimmed to `steps` steps; the rate is the driver's own `Average time per cell`
    (driver/hydro.cpp:212-215: hydro wall clock over steps x cells, from the first step,
    field summaries included)."""
    import re
    import subprocess
    import tempfile
    exe = REPO / "cloverleaf_amd" / "model_mi355x" / "_dropin" / "mi355x-cloverleaf"
    if not exe.exists():
        return {"value": None, "note": "drop-in binary not built (needs the reference tree at build time)"}
    text = re.sub(r"end_step=\d+", f"end_step={steps}", deck_text)
    with tempfile.TemporaryDirectory() as td:
        (Path(td) / "deck.in").write_text(text)
        env = dict(os.environ)
        env.update(env_extra)
        p = subprocess.run([str(exe), "--in", "deck.in", "--out", "clover.out"], cwd=td, capture_output=True, text=True,
                           timeout=900, env=env)
    # the driver prints the grind lines after every step but the last; at the end it
    # prints the whole hydro wall clock (driver/hydro.cpp:108-118, 205-218)
    walls = re.findall(r"^ Wall clock\s+([0-9.eE+-]+)", p.stdout, re.M)
    grinds = re.findall(r"^ Average time per cell\s+([0-9.eE+-]+)", p.stdout, re.M)
    done = re.findall(r"^ Step (\d+) time", p.stdout, re.M)
    if p.returncode != 0 or not walls or not done or int(done[-1]) != steps:
        return {"value": None, "error": (p.stderr or p.stdout)[-800:]}
    nx, ny = (int(re.search(rf"{k}=(\d+)", text).group(1)) for k in ("x_cells", "y_cells"))
    wall = float(walls[-1])
    grind = wall / (steps * float(nx) * ny)
    return {"value": 1.0 / grind, "unit": "cell-steps/s", "steps": steps, "hydro_wall_s": wall,
            "last_printed_grind_s": float(grinds[-1]) if grinds else None,
            "effective_hbm_frac_at_1440B": B_ALG_PER_CELL_STEP / grind / (HBM_PEAK_GBS * 1e9),
            "workload": f"the headline deck trimmed to end_step={steps}, run by the reference's own driver with the "
                        f"MI355X model; rate = steps x cells / its final printed hydro 'Wall clock' (the grind time of "
                        f"driver/hydro.cpp:212-215, {grind:.5g} s per cell-step, field summaries included)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=87)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plan", choices=["fused", "reference"], default="fused")
    ap.add_argument("--deck", default="clover_bm256_short")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-probe", action="store_true", help="skip the HBM copy-bandwidth probe")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-problem batched workload")
    ap.add_argument("--batch-only", action="store_true", help="only the 64-problem batched workload")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint workload")
    ap.add_argument("--adjoint-only", action="store_true", help="only the adjoint workload")
    ap.add_argument("--adjoint-directions", type=int, default=32,
                    help="finite-difference directions of the adjoint check (0: timings only)")
    ap.add_argument("--no-decomposed", action="store_true", help="skip the decomposed-domain workload")
    ap.add_argument("--decomposed-only", action="store_true", help="only the decomposed-domain workload")
    ap.add_argument("--probe-only", action="store_true",
                    help="only run the copy kernel (4 GiB, 20 launches): the counter calibration run")
    ap.add_argument("--cpu-steps", type=int, default=60, help="steps of the CPU baselines on the 3840^2 sample")
    ap.add_argument("--no-mesh", action="store_true", help="skip the timed run with the geometry from the mesh")
    ap.add_argument("--no-dropin", action="store_true", help="skip the drop-in (reference driver) leg")
    ap.add_argument("--dropin-steps", type=int, default=87, help="steps of the drop-in run (87: the deck as it is)")
    ap.add_argument("--profile-out", default=None)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: runs on one GPU only (one chunk per GPU); --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import cloverleaf_amd as clv
    if args.probe_only:
        gbps = clv.bandwidth_probe(4 << 30, 20)
        print(json.dumps({"probe": "copy", "bytes_per_launch_read": 4 << 30, "bytes_per_launch_write": 4 << 30,
                          "launches": 21, "copy_GBps": gbps}))
        return
    if args.batch_only:
        print(json.dumps({"batched": batch_leg(clv)}))
        return
    if args.adjoint_only:
        print(json.dumps({"adjoint": adjoint_leg(clv, args.adjoint_directions)}))
        return
    if args.decomposed_only:
        print(json.dumps({"decomposed": decomposed_leg(clv)}))
        return
    deck_path = REPO / "tests" / "golden" / "decks" / f"{args.deck}.in"
    deck_text = deck_path.read_text()
    plan = clv.PLAN_FUSED if args.plan == "fused" else clv.PLAN_REFERENCE

    # the headline: volume and face areas read from their arrays (driver/clover_leaf.cpp:31-36)
    head = headline_leg(clv, deck_text, plan, args.steps, args.warmup, clv.GEOMETRY_ARRAYS, args.profile_out)
    nx, ny = head["nx"], head["ny"]
    value = head["value"]
    per_kernel, dom = head["per_kernel"], head["dominant"]
    dom_bytes = alg_bytes(dom, nx, ny)
    avg_s = per_kernel[dom]["avg_us"] / 1e6
    achieved = dom_bytes / avg_s / 1e9 if dom_bytes else None
    traffic = None
    pmc = REPO / "profiles" / f"{dom}_{args.deck}_{args.plan}_traffic.json"
    if pmc.exists():
        traffic = json.loads(pmc.read_text()).get("bytes_per_launch")
    out = {
        "metric": "Gcell-steps/s + % HBM peak at 15360^2 FP64 (cell-steps/s, clover_bm256_short)",
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["wall_s"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": f"synthetic (deck {args.deck}.in generated on the GPU)",
        "config": {"workload": f"{args.deck}: {nx}x{ny} cells two-state ideal gas, FP64, steps "
                               f"{args.warmup + 1}..{args.warmup + args.steps} with field_summary every "
                               f"{head['summary_frequency']} steps; volume and face areas read from their arrays",
                   "x_cells": nx, "y_cells": ny, "plan": args.plan, "geometry": "arrays", "parallelism": "1 GPU, 1 chunk"},
        "gcell_steps_per_s": value / 1e9,
        "effective_hbm_frac_at_1440B": value * B_ALG_PER_CELL_STEP / (HBM_PEAK_GBS * 1e9),
        "fused_compulsory_B_per_cell_step": B_FUSED_PER_CELL_STEP[False] if args.plan == "fused" else None,
        "compulsory_hbm_frac_fused": value * B_FUSED_PER_CELL_STEP[False] / (HBM_PEAK_GBS * 1e9) if args.plan == "fused" else None,
        "summaries_in_timed_region": head["summaries"],
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
                     "alg_bytes_per_launch": dom_bytes, "avg_launch_us": avg_s * 1e6,
                     "share_of_kernel_time": per_kernel[dom]["share"]},
        "kernel_time_fraction_of_wall": head["kernel_ms"] / 1e3 / head["wall_s"],
        "kernels": per_kernel,
        "final_summary": head["final_summary"],
    }
    if not args.no_mesh:
        # the same run with the geometry computed from the 1D spacings (CLV_GEOMETRY_MESH):
        # a separately labelled figure, not the headline
        mesh = headline_leg(clv, deck_text, plan, args.steps, args.warmup, clv.GEOMETRY_MESH, None)
        out["geometry_from_mesh"] = {
            "note": "volume = celldx(i) * celldy(j), xarea = celldy(j), yarea = celldx(i) computed from the 1D mesh "
                    "instead of read (the same doubles); not the headline",
            "cell_steps_per_s": mesh["value"], "ms_per_step": mesh["wall_s"] / args.steps * 1e3,
            "compulsory_B_per_cell_step": B_FUSED_PER_CELL_STEP[True] if args.plan == "fused" else None,
            "final_summary_identical": mesh["final_summary"] == head["final_summary"],
            "kernels": {k: {"avg_us": v["avg_us"], "alg_GBps": alg_bytes(k, nx, ny, True) / (v["avg_us"] / 1e6) / 1e9
                            if alg_bytes(k, nx, ny, True) else None} for k, v in mesh["per_kernel"].items()}}
    if not args.no_dropin:
        print("# bench: drop-in (reference driver + MI355X model) ...", file=sys.stderr, flush=True)
        out["dropin"] = dropin_leg(deck_text, args.dropin_steps)
    if not args.no_batch:
        print("# bench: batched workload ...", file=sys.stderr, flush=True)
        out["batched"] = batch_leg(clv)
    if not args.no_adjoint:
        print("# bench: adjoint workload ...", file=sys.stderr, flush=True)
        out["adjoint"] = adjoint_leg(clv, args.adjoint_directions)
    if not args.no_decomposed:
        print("# bench: decomposed workload ...", file=sys.stderr, flush=True)
        out["decomposed"] = decomposed_leg(clv)
    # the probes last: their many multi-GB allocations and frees leave the allocator in a
    # state that slows later legs' kernels (the adjoint backward by up to 20 %)
    if not args.no_probe:
        out["hbm_copy_GBps_measured"] = clv.bandwidth_probe(4 << 30, 20)
        if achieved:
            out["roofline"]["frac_of_measured_copy"] = achieved / out["hbm_copy_GBps_measured"]
        if args.plan == "fused":
            # the ceiling of each fused kernel's access pattern: the same march over the
            # same grid with its arrays read and written and no arithmetic
            ceil = {}
            for name, (r, w) in PATTERN[False].items():
                if name in per_kernel:
                    probe_ms = clv.march_probe(nx, ny, r, w, 16, 10)
                    ceil[name] = {"arrays_read": r, "arrays_written": w, "probe_ms": round(probe_ms, 3),
                                  "kernel_ms": round(per_kernel[name]["avg_us"] / 1e3, 3),
                                  "kernel_over_probe": round(per_kernel[name]["avg_us"] / 1e3 / probe_ms, 3)}
            out["pattern_ceiling"] = ceil
            if dom in ceil:
                out["roofline"]["frac_of_pattern_ceiling"] = ceil[dom]["probe_ms"] / ceil[dom]["kernel_ms"]
    if not args.no_cpu:
        print("# bench: CPU baseline (reference omp model) ...", file=sys.stderr, flush=True)
        ref = cpu_baseline_reference(args.cpu_steps)
        print("# bench: CPU baseline (OpenMP oracle) ...", file=sys.stderr, flush=True)
        port = cpu_baseline(deck_text, args.cpu_steps)
        if ref is not None:
            ref["port"] = port
            out["cpu_baseline"] = ref
        else:
            out["cpu_baseline"] = port
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
