"""CPU tests pinning the oracle's handling of individual time steps in the
flux pass and the update (no reference run: VorFace::calculate_flux needs a
TimeLine, whose constructors need Boost):

  * gating and scatter (src/VorFace.cpp:672-683, 913-955), by brute force over
    every activity pattern of tiny periodic and mirror-walled meshes: a face
    is computed unless its area is zero, or both of its cells are inactive,
    or its right side is a mirror and its left cell is inactive; a computed
    face's flux does not depend on the pattern; each cell's dQ is the
    face-order sum of the computed faces' fluxes, left cells always, right
    cells when they are real and not periodic copies;
  * time steps (src/VorFace.cpp:695-699, 840-866): a face integrates over the
    smaller of its cells' steps, and the mass-flux limiter caps the outflow
    at the area share of the cell's mass times dt / max(dt_L, dt_R);
  * the update (src/Simulation.cpp:476-483, 792-796): Q -= dQ everywhere, W
    rebuilt only for the cells whose step ends at the new time.
"""
import itertools
import os
import sys

import numpy as np
import pytest

import oracle as O
import timesteps as TS
from shadowfax_amd import mesh as M

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
G = 5.0 / 3.0


def tiny(name):
    """Meshes of 9 (2D periodic), 9 (2D with mirror walls) and 12 (3D
    periodic) cells with varied states and moving generators."""
    rng = np.random.default_rng(7)
    if name == "p2":
        m = M.voronoi_periodic(M.jittered_grid(3, 0.25, seed=3))
    elif name == "m2":
        import make_golden as MG
        m = MG.mirrorize(M.voronoi_periodic(M.jittered_grid(3, 0.25, seed=5)))
    else:
        m = M.voronoi_periodic(M.poisson_points(12, seed=6, ndim=3))
    nd, C = m["ndim"], m["C"]
    W = np.zeros((nd + 2, C))
    W[0] = 1.0 + rng.random(C)
    W[1:1 + nd] = 0.3 * rng.normal(size=(nd, C))
    W[-1] = 1.0 + 2.0 * rng.random(C)
    vel = W[1:1 + nd] + 0.05 * rng.normal(size=(nd, C))
    dt = 2e-3 * 0.5 ** rng.integers(0, 3, C)
    return m, W, vel, dt


def expected_dQ(b, dQL_all, act):
    """The reference's rules, written out per face in face order."""
    L, R, gh = b["left"], b["right"], b["ghost"]
    dQ = np.zeros((b["ndim"] + 2, b["C"]))
    computed = np.zeros(b["F"], bool)
    for f in range(b["F"]):
        if b["area"][f] == 0:
            continue
        if R[f] < 0:
            if not act[L[f]]:
                continue
        elif not act[R[f]] and not act[L[f]]:
            continue
        computed[f] = True
        dQ[:, L[f]] += dQL_all[:, f]
        if R[f] >= 0 and not gh[f]:
            dQ[:, R[f]] += -dQL_all[:, f]
    return dQ, computed


@pytest.mark.parametrize("name", ["p2", "m2", "p3"])
@pytest.mark.parametrize("kind", [O.EXACT, O.HLLC])
def test_gating_and_scatter_brute_force(name, kind):
    m, W, vel, dt = tiny(name)
    C = m["C"]
    b = M.prepare(m, W, G, kind=kind, vel=vel, dt=dt)
    b["grad"] = O.gradients(kind, b)
    _, dQL_all, st, _ = O.face_flux_pass(kind, b)
    assert st.max() == 0
    if name == "m2":
        assert (b["right"] < 0).sum() > 0
    else:
        assert (b["ghost"] == 1).sum() > 0
    assert len(np.unique(dt)) > 1
    npat = 0
    for pat in itertools.product((0, 1), repeat=C):
        act = np.array(pat, np.uint8)
        bb = dict(b)
        bb["active"] = act
        dQ, dQL, st, _ = O.face_flux_pass(kind, bb)
        want, computed = expected_dQ(b, dQL_all, act)
        assert np.array_equal(dQL[:, computed], dQL_all[:, computed]), pat
        assert np.all(dQL[:, ~computed] == 0), pat
        assert np.array_equal(dQ, want), pat
        npat += 1
    assert npat == 2 ** C


@pytest.mark.parametrize("nd", [2, 3])
def test_hierarchy_puts_about_half_the_cells_inactive(nd):
    if nd == 2:
        m = M.voronoi_periodic(M.jittered_grid(32, 0.25, seed=9))
        W = M.kh_ics(m)
    else:
        m = M.voronoi_periodic(M.poisson_points(3000, seed=9))
        rng = np.random.default_rng(9)
        W = np.zeros((5, m["C"]))
        W[0] = 1.0 + 0.2 * rng.random(m["C"])
        W[1:4] = 0.1 * rng.normal(size=(3, m["C"]))
        W[4] = 1.0 + 0.2 * rng.random(m["C"])
    b = M.prepare(m, W, G, kind=1)
    h = TS.hierarchy(b, 1)
    s1, e1, dt1, a1 = h["first"]
    s2, e2, dt2, a2 = h["second"]
    assert a1.all()
    assert 0.3 < a2.mean() < 0.7, a2.mean()
    # the active cells are the ones whose first step was the small one
    assert np.array_equal(a2 == 1, e1 == h["S"])
    assert np.all(dt2[a2 == 0] == 2 * dt2[a2 == 1].min())
    assert np.all(s2[a2 == 1] == h["S"]) and np.all(s2[a2 == 0] == 0)


@pytest.mark.parametrize("kind", [O.EXACT, O.HLLC])
def test_mass_flux_limiter_with_individual_steps(kind):
    """Steps 40 times longer than the CFL time on a half-active mesh: the
    outflow through every face is capped at the area share of the cell's
    mass times dt / max(dt_L, dt_R), and the cap binds on faces whose cells
    have different steps."""
    m = M.voronoi_periodic(M.jittered_grid(24, 0.25, seed=2))
    W = M.kh_ics(m)
    vel = np.zeros((2, m["C"]))        # a static mesh: the shear flow crosses faces
    b = M.prepare(m, W, G, kind=kind, vel=vel)
    h = TS.hierarchy(b, kind, scale=40.0)
    s2, e2, dt2, a2 = h["second"]
    b = M.prepare(m, W, G, kind=kind, vel=vel, dt=dt2, active=a2)
    b["grad"] = O.gradients(kind, b)
    dQ, dQL, st, _ = O.face_flux_pass(kind, b)
    L, R = b["left"], b["right"]
    dt = np.minimum(dt2[L], dt2[R])
    tf = dt / np.maximum(dt2[L], dt2[R])
    capL = b["area"] / b["total_area"][L] * b["mass"][L] * tf
    capR = b["area"] / b["total_area"][R] * b["mass"][R] * tf
    assert np.all(dQL[0] <= capL * (1 + 1e-14))
    assert np.all(-dQL[0] <= capR * (1 + 1e-14))
    atcap = (np.abs(dQL[0] - capL) <= 1e-14 * capL) | (np.abs(-dQL[0] - capR) <= 1e-14 * capR)
    assert (atcap & (tf < 1)).sum() > 10
    assert (atcap & (tf == 1)).sum() > 10


@pytest.mark.parametrize("kind", [O.EXACT, O.HLLC])
def test_update_rebuilds_only_cells_whose_step_ends(kind):
    m = M.voronoi_periodic(M.jittered_grid(20, 0.25, seed=4))
    W = M.kh_ics(m)
    b = M.prepare(m, W, G, kind=kind)
    h = TS.hierarchy(b, kind)
    s1, e1, dt1, a1 = h["first"]
    b = M.prepare(m, W, G, kind=kind, dt=dt1, active=a1)
    b["grad"] = O.gradients(kind, b)
    dQ, dQL, st, _ = O.face_flux_pass(kind, b)
    tnow = h["S"]                      # the time after the first step
    ends = e1 == tnow
    assert 0.2 < ends.mean() < 0.8
    # a pressure floor in a few rebuilt and a few kept cells
    Qin = b["Q"].copy()
    low = np.arange(b["C"]) % 37 == 0
    Qin[-1, low] = 0.5 * (Qin[1, low] ** 2 + Qin[2, low] ** 2) / Qin[0, low] + dQ[-1, low]
    assert (low & ends).sum() > 0 and (low & ~ends).sum() > 0
    Q, Wn, cst = O.update(kind, 2, Qin, dQ, b["volume"], W=W, endtime=e1, tnow=tnow)
    Qd = Qin - dQ
    Wr, Qr = O.get_W(kind, 2, b["volume"], Qd)
    assert np.array_equal(Wn[:, ends], Wr[:, ends])
    assert np.array_equal(Wn[:, ~ends], W[:, ~ends])
    assert np.array_equal(Q[:, ends], Qr[:, ends])
    assert np.array_equal(Q[:, ~ends], Qd[:, ~ends])
    # without endtime every cell is rebuilt
    Q2, W2, _ = O.update(kind, 2, Qin, dQ, b["volume"])
    assert np.array_equal(W2, Wr) and np.array_equal(Q2, Qr)
