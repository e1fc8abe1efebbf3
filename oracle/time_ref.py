"""Times the oracle's batched solve_for_flux against the reference solvers
compiled from source (oracle/_ref, built by oracle/Makefile) on the same
inputs, on one thread and on every core of the affinity mask: a check that
the oracle, which stands in for the reference on the GPU box's CPU
baseline, runs at the reference's speed.  Test infrastructure: reads only
oracle/ and the workload generator.

    python oracle/time_ref.py [--runs 5]

Batches: the 262,144 problems of the uniform batch (states log-uniform over
six decades) and the first 1,048,576 of the 4,194,304-problem batch with 5%
near-vacuum and 5% 1e6-shock problems, each solved exactly and with HLLC.
Threads take contiguous slices (ctypes releases the interpreter lock);
each figure is the median of --runs timed runs after one warm-up run."""
import argparse
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))
import oracle as O  # noqa: E402
from shadowfax_amd import workloads as WK  # noqa: E402


def sliced(fn, b, threads):
    """fn over `threads` contiguous slices of batch b, in parallel."""
    N = b["WL"].shape[1]
    cuts = np.linspace(0, N, threads + 1).astype(np.int64)
    parts = [(b["WL"][:, a:e], b["WR"][:, a:e], b["n"][:, a:e]) for a, e in zip(cuts, cuts[1:])]
    if threads == 1:
        return [fn(*parts[0])]
    with ThreadPoolExecutor(threads) as ex:
        return list(ex.map(lambda p: fn(*p), parts))


def median_time(fn, b, threads, runs):
    sliced(fn, b, threads)
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        sliced(fn, b, threads)
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--runs", type=int, default=5)
    args = ap.parse_args()
    if not O.ref_available(3):
        sys.exit("oracle/_ref/libsfxref3d.so is not built (make -C oracle ref)")
    cores = len(os.sched_getaffinity(0))
    batches = [("uniform, 262,144", WK.riemann_batch(262144, seed=20240601, family="uniform")),
               ("5% vacuum + 5% shock, first 1,048,576 of 4,194,304",
                WK.riemann_batch(1048576, seed=20240602, family="mixed"))]
    print("host: %s, %d cores in the affinity mask" % (cpu_model(), cores))
    print("| batch | solver | threads | oracle (M solves/s) | reference (M solves/s) | oracle / reference |")
    print("|---|---|---|---|---|---|")
    for name, b in batches:
        N = b["WL"].shape[1]
        for kind, kname in ((0, "exact"), (1, "HLLC")):
            ora = lambda WL, WR, n, k=kind: O.solve_for_flux(k, 3, WL, WR, n)
            ref = lambda WL, WR, n, k=kind: O.ref_solve_for_flux(k, 3, WL, WR, n)
            # the two agree bit for bit before either is timed
            Fo = np.concatenate([r[0] for r in sliced(ora, b, cores)], axis=1)
            Fr = np.concatenate([r[0] for r in sliced(ref, b, cores)], axis=1)
            assert np.array_equal(Fo, Fr), (name, kname)
            for threads in (1, cores):
                to = median_time(ora, b, threads, args.runs)
                tr = median_time(ref, b, threads, args.runs)
                print("| %s | %s | %d | %.2f | %.2f | %.2f |"
                      % (name, kname, threads, N / to / 1e6, N / tr / 1e6, tr / to))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()
