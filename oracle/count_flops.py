"""Count the floating-point operations of the reference algorithm on the
benchmark flux passes (TEST/MEASUREMENT INFRASTRUCTURE: runs the oracle's
counting build, oracle/libflops.so; see oracle/flopcount.cpp for the counting
rule).

    python oracle/count_flops.py [--out profiles/flop_counts.json]

Workloads (the same seeded meshes and states bench.py times):
  3d: 2,097,152 Poisson generators in the 3D periodic unit box (seed
      20240604) in Hilbert order, Sedov blast, exact solver;
  2d: 512 x 512 jittered generators in the 2D periodic unit box (seed
      20240603) in Hilbert order, Kelvin-Helmholtz, HLLC.
Per workload: operations and transcendentals of the gradients (per cell),
the face fluxes (per face, without the scatter) and the scatter/gather
additions (per gather entry), all FP64 as the reference computes them.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import oracle as O  # noqa: E402
from shadowfax_amd import mesh as M  # noqa: E402


def workload(name, nthreads):
    if name == "3d":
        m = M.voronoi_periodic(M.hilbert_sorted(M.poisson_points(2097152, seed=20240604)),
                               nthreads=nthreads)
        return M.prepare(m, M.sedov_ics(m), 5.0 / 3.0, kind=0), 0
    m = M.voronoi_periodic(M.hilbert_sorted(M.jittered_grid(512, 0.25, seed=20240603)),
                           nthreads=nthreads)
    return M.prepare(m, M.kh_ics(m), 5.0 / 3.0, kind=1), 1


def count(b, kind):
    C, F = b["C"], b["F"]
    t0 = time.perf_counter()
    grad, g_ops, g_tr = O.count_flops(O.gradients, kind, b)
    bb = dict(b)
    bb["grad"] = grad
    dQL = np.zeros((b["ndim"] + 2, F))
    st = np.zeros(F, np.uint8)

    def faces():
        m, keep = O.mesh_struct(bb)
        s = O.params(kind, b["ndim"])
        O.lib().sfxo_face_flux_range(O.ct.byref(s), O.ct.byref(m), 0, F, O._p(dQL),
                                     O._p(st, O._u8p), None)

    _, f_ops, f_tr = O.count_flops(faces)
    dQ = np.zeros((b["ndim"] + 2, C))

    def gather():
        m, keep = O.mesh_struct(bb)
        O.lib().sfxo_gather(O.ct.byref(m), O._p(O._c(b["gather_off"], np.int64), O._i64p),
                            O._p(O._c(b["gather_entry"], np.int32), O._i32p), O._p(dQL), 0, C,
                            O._p(dQ))

    _, q_ops, q_tr = O.count_flops(gather)
    G = int(b["gather_off"][-1])
    return {"cells": int(C), "faces": int(F), "gather_entries": G,
            "gradients": {"ops": g_ops, "transcendentals": g_tr, "per_cell": g_ops / C},
            "faces_kernel": {"ops": f_ops, "transcendentals": f_tr, "per_face": f_ops / F,
                             "transcendentals_per_face": f_tr / F},
            "gather": {"ops": q_ops, "per_entry": q_ops / G},
            "pass_ops": g_ops + f_ops + q_ops, "pass_ops_per_face": (g_ops + f_ops + q_ops) / F,
            "count_s": time.perf_counter() - t0}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(ROOT, "profiles", "flop_counts.json"))
    ap.add_argument("--threads", type=int, default=8, help="host tessellation threads")
    args = ap.parse_args()
    out = {"rule": "+ - * / sqrt = 1, pow cbrt exp log = 1 (also tallied as transcendentals), "
                   "comparisons/negation/fabs/fmin/fmax = 0; FP64 (the reference algorithm, "
                   "oracle/libflops.so)",
           "command": "python oracle/count_flops.py"}
    for name in ("2d", "3d"):
        b, kind = workload(name, args.threads)
        out[name] = count(b, kind)
        out[name]["solver"] = ("exact", "hllc")[kind]
        print(name, json.dumps(out[name]), flush=True)
        del b
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
