"""Writes c1_roof_path.json: the oracle's Crisfield path of the C1 roof (CPU, FP64).

C1: shallow cylindrical panel R = 2.5, half-angle 0.1 rad, length 0.5, t = 0.0127, E = 3.1e3,
nu = 0.3, degree 3, 32 x 32 elements; straight edges hinged, curved edges free, central point load
(0, 0, -10) (benchmarks/benchmark_Roof.cpp:198-236); Crisfield, Scaling 0, AngleMethod 0, MaxIter 20,
TolF 1e-8, TolU 1e-6, determinant stability (:359-377); 80 steps of the reference's arc length
0.5 mm on its 45-dof mesh, in metres and scaled by sqrt(n_free / 45).

Run: python tests/golden/make_c1_roof.py   (about four minutes on eight cores)
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from gsstructuralanalysis_amd import geometry as g  # noqa: E402
from oracle.solvers import run_crisfield  # noqa: E402

STEPS = 80


def main():
    pt = g.roof_problem(p=3, n=32)
    ds = g.roof_arc_length(pt.nfree, 1.0)
    t0 = time.time()
    r = run_crisfield(pt, STEPS, ds, max_iter=20, tol_f=1e-8, tol_u=1e-6, scaling=0.0)
    # centre deflection: the z dof of the centre control point (35 x 35 net, odd: one centre point)
    mx = pt.mx
    cp = (mx // 2) + mx * (mx // 2)
    free = np.flatnonzero(pt.fixed == 0)
    zc = int(np.flatnonzero(free == 3 * cp + 2)[0])
    out = {"case": "C1 roof p=3 32x32, Crisfield, Scaling 0, TolF 1e-8", "nfree": int(pt.nfree), "arc_length": ds,
           "steps": int(r["steps_done"]), "halvings": int(r["halvings"]), "centre_dof": zc,
           "lambda": r["lambda"].tolist(), "unorm": r["unorm"].tolist(),
           "w_centre": [float(u[zc]) for u in r["U"]], "negatives": [int(x) for x in r["negatives"]],
           "iterations": [int(x) for x in r["iterations"]], "seconds": round(time.time() - t0, 1)}
    with open(os.path.join(HERE, "c1_roof_path.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"{out['steps']} steps, negatives max {max(out['negatives'])}, {out['seconds']} s")


if __name__ == "__main__":
    main()
