"""Individual time steps for the flux-pass tests: two calls of the oracle's
TimeLine::calculate_timestep (src/TimeLine.cpp:233-320) on a prepared mesh.

  first call, t = 0: every cell is active and takes the power-of-two step
    under its CFL time.  With two_levels, the time line's largest step is
    set just under the median CFL time and its smallest step to half that
    (the clamps of TimeLine::calculate_timestep): the cells above the median
    take the larger step 2 S, the others S.  Without, steps follow the CFL
    times down the whole hierarchy (a blast's hot cells many levels down);
  second call, t = S (the smallest step): the cells with step S end their
    step and start the next one (active); the others are inside theirs
    (inactive).

Returns, per call, the new start and end times, the per-cell dt
(get_realtime(end - start), what the face pass reads) and the active flags,
and the system step.
"""
import numpy as np

import oracle as O

MAXTIME = 1 << 40


def cfl_times(b, kind, gamma, cfl=0.4):
    """h / (|v - v_gen| + c) * cfl per cell (src/TimeLine.cpp:248-262)."""
    nd = b["ndim"]
    W, vel, V = b["W"], b["vel"], b["volume"]
    v = np.sqrt(((W[1:1 + nd] - vel) ** 2).sum(0))
    with np.errstate(divide="ignore", invalid="ignore"):
        c = np.sqrt(gamma * W[-1] / W[0])
    c = np.where(W[0] != 0, c, 0.0)
    h = np.sqrt(V / np.pi) if nd == 2 else np.cbrt(3.0 * V / (4.0 * np.pi))
    return cfl * h / (v + c)


def timeline(tottime, now, step, smin=1, smax=MAXTIME, cfl=0.4):
    return O.timeline(integer_maxtime=MAXTIME, current_time=now, timestep=step,
                      max_timestep=smax, min_timestep=smin, tottime=tottime, mintime=0.0,
                      cfl=cfl, global_timestep=0)


def hierarchy(b, kind, gamma=5.0 / 3.0, level=12, scale=1.0, two_levels=True):
    """The two calls above on bundle b (mesh.prepare).  scale multiplies the
    CFL number (scale > 1 makes every step that much longer than a CFL number
    of 0.4 allows).  Returns a dict with first/second: (start, end, dt,
    active) and S (the smallest step, integer), tottime."""
    cfl = 0.4 * scale
    S = cfl_times(b, kind, gamma, cfl)
    med = np.median(S[np.isfinite(S)])
    big = MAXTIME >> level
    # the median CFL time a hair above the realtime of `big` ticks
    tottime = med * 2.0 ** level * (1.0 - 1e-9)
    smin, smax = (big // 2, big) if two_levels else (1, MAXTIME)
    C = b["C"]
    W = np.ascontiguousarray(b["W"].T)
    vel = np.ascontiguousarray(b["vel"].T)
    zero = np.zeros(C, np.uint64)
    nd = b["ndim"]
    s1, e1, dt1, a1, ss1, na1, st1 = O.hydro_timestep(
        kind, nd, W, vel, b["volume"], timeline(tottime, 0, 0, smin, smax, cfl), zero, zero,
        gamma=gamma)
    assert st1.max() == 0 and na1 == C
    s2, e2, dt2, a2, ss2, na2, st2 = O.hydro_timestep(
        kind, nd, W, vel, b["volume"], timeline(tottime, 0, ss1, smin, smax, cfl), s1, e1,
        gamma=gamma)
    assert st2.max() == 0
    return {"first": (s1, e1, dt1, a1), "second": (s2, e2, dt2, a2), "S": int(ss1),
            "S2": int(ss2), "tottime": tottime, "smin": smin, "smax": smax}
