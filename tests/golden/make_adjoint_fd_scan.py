"""Writes tests/golden/adjoint_fd_scan.json: central differences of the final kinetic
energy of the adjoint workload (cloverleaf_amd.ADJOINT_WORKLOAD: 1920^2, 200
fixed-dt steps) computed with the oracle, for a few of bench.py's finite-difference
directions (density cloverleaf_amd.smooth_field seed 300 + k, energy seed 400 + k)
over a range of steps eps, against the oracle's tangent-linear derivative of the
same direction (tests/golden/adjoint_fd_kinks.json).

ke is summed two ways: exactly (math.fsum of the per-cell terms of
src/serial/field_summary.cpp:80-90, rounded once), and in the serial model's
left-to-right order (the oracle's field_summary).  With the exact sum the only
differences between the +eps and -eps runs are the runs' own: rounding in the
200 steps and the piecewise choices (upwind, limiter, fmin/fmax, viscosity switch)
that the two runs take differently.  Rounding in the runs would grow as eps shrinks;
an error that does not shrink as eps shrinks, or changes sign from one eps to the
next, comes from the choices.

usage: python tests/golden/make_adjoint_fd_scan.py [--directions 25,28,0] [--eps 1e-9,1e-8,1e-7,1e-6]
"""
import argparse
import ctypes as C
import json
import math
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
sys.path.insert(0, str(REPO))
sys.path.insert(0, str(REPO / "tests"))
sys.path.insert(0, str(HERE))

import clover_testlib as T  # noqa: E402
import cloverleaf_amd as clv  # noqa: E402
from make_adjoint_golden import initial_oracle  # noqa: E402


def ke_exact(o, nx, ny):
    """src/serial/field_summary.cpp:80-90's ke terms, summed exactly (math.fsum)."""
    u, v = o.f("xvel0"), o.f("yvel0")
    vsqrd = np.zeros((ny, nx))
    for kv in (0, 1):
        for jv in (0, 1):
            uu = u[2 + kv:ny + 2 + kv, 2 + jv:nx + 2 + jv]
            vv = v[2 + kv:ny + 2 + kv, 2 + jv:nx + 2 + jv]
            vsqrd = vsqrd + 0.25 * (uu * uu + vv * vv)
    mass = o.f("volume")[2:ny + 2, 2:nx + 2] * o.f("density0")[2:ny + 2, 2:nx + 2]
    return math.fsum((mass * 0.5 * vsqrd).ravel().tolist())


def final_ke(w, k, s):
    nx, ny = w["nx"], w["ny"]
    o = initial_oracle(w)
    o.f("density0")[2:ny + 2, 2:nx + 2] += s * clv.smooth_field(nx, ny, 300 + k)
    o.f("energy0")[2:ny + 2, 2:nx + 2] += s * clv.smooth_field(nx, ny, 400 + k)
    o.L.orc_update_halo(C.byref(o.c), T.mask(["density0", "energy0"]), 2)
    adv = 1
    for _ in range(w["steps"]):
        T.oracle_step_fixed(o, w["dt"], adv)
        adv = 1 - adv
    return ke_exact(o, nx, ny), o.summary()[3]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--directions", default="25,28,0")
    ap.add_argument("--eps", default="1e-9,1e-8,1e-7,1e-6")
    a = ap.parse_args()
    w = clv.ADJOINT_WORKLOAD
    kinks = json.loads((HERE / "adjoint_fd_kinks.json").read_text())
    tl = {r["k"]: r["dke"] for r in kinks["directions"]}
    recs = []
    for k in [int(x) for x in a.directions.split(",")]:
        for eps in [float(x) for x in a.eps.split(",")]:
            t0 = time.time()
            (xp, sp), (xm, sm) = final_ke(w, k, eps), final_ke(w, k, -eps)
            fd_exact, fd_serial = (xp - xm) / (2 * eps), (sp - sm) / (2 * eps)
            r = {"k": k, "eps": eps, "tangent_linear": tl[k], "fd_exact_sum": fd_exact, "fd_serial_sum": fd_serial,
                 "rel_diff_exact_sum": abs(fd_exact - tl[k]) / abs(tl[k]),
                 "rel_diff_serial_sum": abs(fd_serial - tl[k]) / abs(tl[k])}
            recs.append(r)
            print(f"direction {k} eps {eps:g}: exact-sum rel diff {r['rel_diff_exact_sum']:.3e}, serial-sum "
                  f"{r['rel_diff_serial_sum']:.3e} ({time.time() - t0:.0f} s)", flush=True)
    out = {"workload": w, "records": recs}
    (HERE / "adjoint_fd_scan.json").write_text(json.dumps(out, indent=1) + "\n")
    print("wrote", HERE / "adjoint_fd_scan.json")


if __name__ == "__main__":
    main()
