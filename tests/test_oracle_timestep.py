"""CPU tests pinning the oracle's TimeLine::calculate_timestep (gas cells, no
gravity; src/TimeLine.cpp:233-320,418-431).  The reference's TimeLine.cpp
needs Boost.property_tree, absent here, so no reference run pins it; the
checks are independent of the oracle's loops:
  * hand-worked cases (closed forms of h, c and the power-of-two floor);
  * brute force: for every active cell, the step is the FIRST of
    maxtime >> j, j = 0, 1, ..., that fits the CFL bound and divides the time
    left (maxtime - endtime) -- a scan over all candidates rather than the
    reference's two halving loops -- clamped to [min, max], or max_timestep
    when even that fits;
  * invariants: inactive cells keep their times in individual mode, every
    cell takes the smallest step in global mode, dt = realtime(end - start).
"""
import math

import numpy as np
import pytest

import oracle as O

G = 5.0 / 3.0


def tl(**kw):
    d = dict(integer_maxtime=1 << 20, current_time=0, timestep=0, max_timestep=1 << 20,
             min_timestep=1, tottime=1.0, mintime=0.0, cfl=0.4, global_timestep=0)
    d.update(kw)
    return O.timeline(**d), d


def cells(rng, C, nd, zero_rho=False):
    W = np.empty((C, nd + 2))
    W[:, 0] = 10 ** rng.uniform(-2, 2, C)
    W[:, 1:1 + nd] = rng.uniform(-2, 2, (C, nd))
    W[:, -1] = 10 ** rng.uniform(-3, 1, C)
    if zero_rho:
        W[::17, 0] = 0.0
    vel = W[:, 1:1 + nd] + rng.uniform(-0.5, 0.5, (C, nd))
    V = 10 ** rng.uniform(-7, -3, C)
    return W, vel, V


def cfl_step(kind, nd, W, vel, V, cfl, gamma=G):
    """The real CFL step of every cell: h/(|v - v_gen| + c), times cfl."""
    v = W[:, 1:1 + nd] - vel
    vi = np.sqrt((v * v).sum(1))
    with np.errstate(divide="ignore", invalid="ignore"):
        c = np.sqrt(gamma * W[:, -1] / W[:, 0])
    if kind == O.EXACT:
        c = np.where(W[:, 0] != 0, c, 0.0)
    h = np.cbrt(V * 3.0 / 4.0 / math.pi) if nd == 3 else np.sqrt(V / math.pi)
    return cfl * (h / (vi + c))


def brute_step(S, end, d):
    M = d["integer_maxtime"]
    interval = lambda t: d["tottime"] * (t / M)
    Smax = d["max_timestep"]
    if interval(Smax) < S:
        return Smax
    for j in range(64):
        P = M >> j
        if P == 0:
            return None
        if interval(P) <= S and (M - end) % P == 0:
            return min(max(P, d["min_timestep"]), Smax)
    return None


def test_hand_worked_2d():
    """rho = p = 1, v = v_gen = 0, V = 1e-4: h = sqrt(1e-4/pi) = 5.6419e-3,
    c = sqrt(5/3) = 1.29099, S = 0.4 h / c = 1.7481e-3 of a unit time line of
    2^20 ticks = 1833.1 ticks -> the largest power of two below: 1024."""
    t, d = tl()
    W = np.array([[1.0, 0.0, 0.0, 1.0]])
    s, e, dt, act, ss, na, st = O.hydro_timestep(O.EXACT, 2, W, np.zeros((1, 2)),
                                                 np.array([1e-4]), t, [0], [0])
    assert e[0] == 1024 and s[0] == 0 and ss == 1024 and na == 1 and act[0] == 1
    assert dt[0] == 1024 / 2 ** 20


def test_hand_worked_sync_and_clamps():
    """A cell ending at t = 3 * 256 on a 2^12 line can only take steps that
    divide 4096 - 768 = 3328 = 13 * 256, so a CFL-admissible 512 drops to
    256; min_timestep raises it back; a huge S gives max_timestep."""
    t, d = tl(integer_maxtime=1 << 12, max_timestep=1 << 12, current_time=700, timestep=68)
    W = np.array([[1.0, 0.0, 0.0, 1.0]])
    # S = 0.13 of the unit line -> 532 ticks -> 512 by the CFL floor
    c = math.sqrt(G)
    V = math.pi * (0.13 / 0.4 * c) ** 2
    s, e, dt, act, ss, na, st = O.hydro_timestep(O.EXACT, 2, W, np.zeros((1, 2)),
                                                 np.array([V]), t, [512], [768])
    assert act[0] == 1 and s[0] == 768 and e[0] == 768 + 256 and ss == 256
    t2, _ = tl(integer_maxtime=1 << 12, max_timestep=1 << 12, min_timestep=400,
               current_time=700, timestep=68)
    s, e, *_ = O.hydro_timestep(O.EXACT, 2, W, np.zeros((1, 2)), np.array([V]), t2,
                                [512], [768])
    assert e[0] == 768 + 400
    t3, _ = tl(integer_maxtime=1 << 12, max_timestep=1 << 10)
    s, e, dt, act, ss, *_ = O.hydro_timestep(O.EXACT, 2, W, np.zeros((1, 2)),
                                             np.array([1e6]), t3, [0], [0])
    assert e[0] == 1 << 10 and ss == 1 << 10


@pytest.mark.parametrize("nd", [2, 3])
@pytest.mark.parametrize("kind", [O.EXACT, O.HLLC])
@pytest.mark.parametrize("bits", [12, 20, 60])
def test_brute_force_individual(nd, kind, bits):
    rng = np.random.default_rng(100 + nd + 7 * kind + bits)
    C = 3000
    W, vel, V = cells(rng, C, nd, zero_rho=(kind == O.EXACT))
    M = 1 << bits
    t, d = tl(integer_maxtime=M, max_timestep=M >> 2, min_timestep=1 if bits < 20 else 8,
              tottime=0.05 if bits == 12 else 1e-3, current_time=M >> 3, timestep=M >> 4)
    now = d["current_time"] + d["timestep"]
    # half the cells end now (active), ending times multiples of small powers
    end = np.where(rng.uniform(size=C) < 0.5, now, now + (M >> 6)).astype(np.uint64)
    start = end - np.uint64(M >> 8)
    s, e, dt, act, ss, na, st = O.hydro_timestep(kind, nd, W, vel, V, t, start, end)
    S = cfl_step(kind, nd, W, vel, V, d["cfl"])
    assert np.array_equal(act, (end == now).astype(np.uint8))
    assert na == int((end == now).sum())
    smin = d["max_timestep"]
    for i in range(C):
        if end[i] != now:
            assert s[i] == start[i] and e[i] == end[i]
            continue
        want = brute_step(S[i], int(end[i]), d)
        if want is None:
            assert st[i] == 16
            continue
        assert st[i] == 0 and s[i] == end[i] and int(e[i] - s[i]) == want, i
        smin = min(smin, want)
    assert ss == smin
    M_ = float(M)
    assert np.array_equal(dt, d["tottime"] * ((e - s).astype(np.float64) / M_) + d["mintime"])


@pytest.mark.parametrize("nd", [2, 3])
def test_global_mode_takes_the_smallest_step(nd):
    rng = np.random.default_rng(7 + nd)
    C = 500
    W, vel, V = cells(rng, C, nd)
    M = 1 << 30
    t, d = tl(integer_maxtime=M, max_timestep=M, global_timestep=1, tottime=1e-2,
              mintime=0.25, current_time=0, timestep=0)
    start = np.zeros(C, np.uint64)
    end = np.zeros(C, np.uint64)
    s, e, dt, act, ss, na, st = O.hydro_timestep(O.HLLC, nd, W, vel, V, t, start, end)
    S = cfl_step(O.HLLC, nd, W, vel, V, d["cfl"])
    want = min(brute_step(x, 0, d) for x in S)
    assert ss == want and na == C
    assert np.all(s == 0) and np.all(e == want)
    # get_realtime includes _mintime (the reference's face dt does)
    assert np.all(dt == d["tottime"] * (want / float(M)) + 0.25)


def test_no_active_cell_gives_max_step():
    t, d = tl(current_time=10, timestep=5)
    W, vel, V = cells(np.random.default_rng(1), 10, 2)
    s, e, dt, act, ss, na, st = O.hydro_timestep(O.EXACT, 2, W, vel, V, t,
                                                 np.zeros(10), np.full(10, 99))
    assert ss == d["max_timestep"] and na == 0 and not act.any()
    assert np.all(e == 99)
