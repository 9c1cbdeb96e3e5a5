"""Writes tests/golden/adjoint_c4.json: the oracle's tangent-linear model on the
adjoint workload (cloverleaf_amd.ADJOINT_WORKLOAD: the two-state problem on a
1920^2 grid with a seeded 1e-3 density perturbation, 200 fixed-dt steps).

Records the final kinetic energy and, for a few seeded smooth directions
(cloverleaf_amd.smooth_field), its directional derivative d ke / d(direction)
from a forward-mode (dual-number) sweep of the oracle (orc_tl_step_fixed).
The oracle's tangent-linear model is pinned against central differences in
tests/test_oracle_tl.py.

usage: python tests/golden/make_adjoint_golden.py [--directions K]
"""
import argparse
import ctypes as C
import json
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
sys.path.insert(0, str(REPO))
sys.path.insert(0, str(REPO / "tests"))

import clover_testlib as T  # noqa: E402
import cloverleaf_amd as clv  # noqa: E402


def initial_oracle(w):
    nx, ny = w["nx"], w["ny"]
    o = T.Oracle(T.parse_deck(clv.bm_deck(nx, ny)))
    o.start()
    rho, e = clv.adjoint_initial(o.f("density0")[2:ny + 2, 2:nx + 2].copy(), o.f("energy0")[2:ny + 2, 2:nx + 2].copy(),
                                 w["amplitude"], w["seed"])
    o.f("density0")[2:ny + 2, 2:nx + 2] = rho
    o.f("energy0")[2:ny + 2, 2:nx + 2] = e
    o.L.orc_update_halo(C.byref(o.c), T.mask(["density0", "energy0"]), 2)
    return o


def direction_seeds(k):
    return 100 + k, 200 + k


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--directions", type=int, default=4)
    a = ap.parse_args()
    w = clv.ADJOINT_WORKLOAD
    nx, ny, steps, dt = w["nx"], w["ny"], w["steps"], w["dt"]
    o = initial_oracle(w)
    recs = []
    ke = None
    for k in range(a.directions):
        t0 = time.time()
        t = T.TLOracle.from_oracle(o)
        sr, se = direction_seeds(k)
        t.der("density0")[2:ny + 2, 2:nx + 2] = clv.smooth_field(nx, ny, sr)
        t.der("energy0")[2:ny + 2, 2:nx + 2] = clv.smooth_field(nx, ny, se)
        adv = 1
        for _ in range(steps):
            t.run("step_fixed", dt, adv)
            adv = 1 - adv
        v, dv = t.summary()[3]
        ke = v
        recs.append({"seed_density": sr, "seed_energy": se, "dke": dv})
        print(f"direction {k}: dke {dv:.17e} ({time.time() - t0:.0f} s)", flush=True)
    out = {"workload": w, "ke": ke, "directions": recs}
    (HERE / "adjoint_c4.json").write_text(json.dumps(out, indent=1) + "\n")
    print("wrote", HERE / "adjoint_c4.json")


if __name__ == "__main__":
    main()
