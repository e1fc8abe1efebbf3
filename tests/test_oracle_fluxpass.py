"""CPU tests pinning the oracle's flux pass.

  * gradients (VorCell::estimate_gradient), face velocities (VorFace::set_v)
    and normals (VorFace::get_normal): golden vectors from the reference
    classes compiled from source, on periodic and mirror-face meshes;
    bit-exact;
  * VorFace::limit: hand-computed values of the GIZMO pairwise limiter
    (Hopkins 2015, appendix B: psi1 = 0.5, psi2 = 0.25);
  * VorFace::calculate_flux (not runnable from the reference here: it needs
    a TimeLine, whose constructors need Boost): invariants -- zero gradients
    reduce each face to solve_for_flux on the boosted states (bit-exact),
    uniform comoving flow gives zero net change, the periodic pass conserves
    to round-off, the mass-flux limiter caps the outflow at the area share of
    the cell's mass, the scatter is the ordered per-cell sum of the face
    fluxes, and the update step is update_Q + get_W.
"""
import os

import numpy as np
import pytest

import oracle as O
from shadowfax_amd import mesh as M

HERE = os.path.dirname(os.path.abspath(__file__))
G = 5.0 / 3.0
KEYS = ("pos", "volume", "centroid", "left", "right", "area", "midpoint", "shift", "ghost",
        "twin", "cf_off", "cf_face", "W", "vel")


MESHES = ["per12", "mir10", "p3d", "m3d"]


def golden_mesh(name):
    nd = 3 if name.endswith("3d") else 2
    z = np.load(os.path.join(HERE, "golden", "mesh_%dd.npz" % nd))
    m = {k: z[name + "_" + k] for k in KEYS}
    m["ndim"] = nd
    m["C"] = m["volume"].shape[0]
    m["F"] = m["area"].shape[0]
    m["box"] = np.ones(nd)
    ref = {k: z[name + "_ref_" + k] for k in ("grad", "vface", "normal")}
    return m, ref


def bundle(m, kind=1, dt=None):
    return M.prepare(m, m["W"], G, kind=kind, vel=m["vel"], dt=dt)


@pytest.mark.parametrize("name", MESHES)
def test_gradients_match_reference_bitwise(name):
    m, ref = golden_mesh(name)
    assert np.array_equal(O.gradients(1, m), ref["grad"])


@pytest.mark.parametrize("name", MESHES)
def test_face_velocity_and_normal_match_reference_bitwise(name):
    m, ref = golden_mesh(name)
    v, n = O.face_velocities(m)
    assert np.array_equal(v, ref["vface"])
    assert np.array_equal(n, ref["normal"])


# (phimid0, phiL, phiR, |d|/r) -> limited value, worked by hand from
# phibar = phiL + |d|/r (phiR - phiL), delta1 = |phiL - phiR|/2,
# delta2 = |phiL - phiR|/4, phi+- = phimax/min +- delta1 unless that flips
# the sign (then phi/(1 + delta1/|phi|)).
LIMITER_CASES = [
    (3.0, 1.0, 2.0, 0.5, 1.75),      # capped at phibar + delta2 = 1.5 + 0.25
    (1.2, 1.0, 2.0, 0.5, 1.2),       # inside: unchanged
    (0.1, 1.0, 2.0, 0.5, 0.5),       # floored at phi- = 1 - 0.5
    (3.0, 2.0, 1.0, 0.5, 2.5),       # capped at phi+ = 2 + 0.5
    (0.0, 2.0, 1.0, 0.5, 1.25),      # floored at phibar - delta2 = 1.5 - 0.25
    (0.5, -0.1, -0.5, 0.5, -0.1 / 3.0),  # phi+ = -0.1/(1 + 0.2/0.1)
    (7.0, 4.0, 4.0, 0.5, 4.0),       # equal neighbours: phiL
]


@pytest.mark.parametrize("case", LIMITER_CASES)
def test_limiter_hand_computed(case):
    phimid0, phiL, phiR, ratio, want = case
    d = np.array([0.3, 0.4])  # |d| = 0.5
    got = O.limit(phimid0, phiL, phiR, d, 0.5 / ratio, 2)
    assert abs(got - want) <= 1e-15 * max(1.0, abs(want))


@pytest.mark.parametrize("name", ["mir10", "m3d"])
def test_zero_gradients_reduce_to_solve_for_flux(name):
    """With grad = 0 the predictor and limiter leave the boosted states
    unchanged, so dQL = (dt A) solve_for_flux(WL - v, WR - v, n, v)."""
    m, _ = golden_mesh(name)
    nd = m["ndim"]
    for kind in (O.EXACT, O.HLLC):
        b = bundle(m, kind)
        b["grad"] = np.zeros((nd, nd + 2, m["C"]))
        dQ, dQL, st, _ = O.face_flux_pass(kind, b)
        v, n = O.face_velocities(b)
        L, R = b["left"], b["right"]
        WL = b["W"][:, L].copy()
        WR = np.where(R >= 0, b["W"][:, np.maximum(R, 0)], 0.0)
        WL[1:1 + nd] -= v
        WR[1:1 + nd] -= v
        F, sts, _ = O.solve_for_flux(kind, nd, WL, WR, n, v=v, refl=(R < 0).astype(np.uint8))
        dt = np.minimum(b["dt"][L], np.where(R >= 0, b["dt"][np.maximum(R, 0)], np.inf))
        want = (dt * b["area"]) * F
        live = b["area"] != 0
        assert np.array_equal(dQL[:, live], want[:, live])


@pytest.mark.parametrize("name", ["per12", "p3d"])
def test_uniform_comoving_flow_gives_no_net_change(name):
    m, _ = golden_mesh(name)
    nd = m["ndim"]
    W = np.zeros((nd + 2, m["C"]))
    W[0], W[1], W[2], W[-1] = 1.3, 0.4, -0.2, 2.0
    if nd == 3:
        W[3] = 0.1
    m = dict(m)
    m["W"], m["vel"] = W, W[1:1 + nd].copy()
    for kind in (O.EXACT, O.HLLC):
        b = bundle(m, kind)
        b["grad"] = O.gradients(kind, b)
        assert np.abs(b["grad"]).max() == 0.0
        dQ, dQL, st, _ = O.face_flux_pass(kind, b)
        assert np.abs(dQ).max() <= 1e-13 * np.abs(dQL).max()


@pytest.mark.parametrize("nd", [2, 3])
@pytest.mark.parametrize("kind", [O.EXACT, O.HLLC])
def test_periodic_pass_conserves(kind, nd):
    if nd == 2:
        m = M.voronoi_periodic(M.jittered_grid(20, 0.25, seed=2))
        W = M.kh_ics(m)
    else:
        m = M.voronoi_periodic(M.poisson_points(800, seed=2))
        W = M.sedov_ics(m, ncells=8)
    b = M.prepare(m, W, G, kind=kind)
    b["grad"] = O.gradients(kind, b)
    dQ, dQL, st, _ = O.face_flux_pass(kind, b)
    assert st.max() == 0
    # periodic faces are evaluated once from each side: the two copies agree
    # to round-off, so the totals vanish to round-off
    assert np.abs(dQ.sum(1)).max() <= 1e-13 * np.abs(dQ).max()
    tw = b["twin"]
    g = np.where((b["ghost"] == 1) & (tw >= 0))[0]
    assert np.allclose(dQL[:, g], -dQL[:, tw[g]], rtol=1e-10, atol=1e-16)


def test_scatter_is_ordered_sum_of_face_fluxes():
    m, _ = golden_mesh("per12")
    b = bundle(m, 1)
    b["grad"] = O.gradients(1, b)
    dQ, dQL, st, _ = O.face_flux_pass(1, b)
    acc = np.zeros_like(dQ)
    for f in range(b["F"]):  # the reference's forward face loop
        acc[:, b["left"][f]] += dQL[:, f]
        if b["right"][f] >= 0 and not b["ghost"][f]:
            acc[:, b["right"][f]] += -dQL[:, f]
    assert np.array_equal(acc, dQ)
    # the gather index lists exactly these contributions, in face order
    off, ent = b["gather_off"], b["gather_entry"]
    acc2 = np.zeros_like(dQ)
    for i in range(b["C"]):
        for e in ent[off[i]:off[i + 1]]:
            x = dQL[:, e >> 1]
            acc2[:, i] += -x if (e & 1) else x
    assert np.array_equal(acc2, dQ)


def test_mass_flux_limiter_caps_outflow():
    m, _ = golden_mesh("per12")
    b = bundle(m, 1, dt=50.0)  # absurd time step: limiter must act
    b["grad"] = np.zeros((2, 4, m["C"]))
    dQ, dQL, st, _ = O.face_flux_pass(1, b)
    L, R = b["left"], b["right"]
    capL = b["area"] / b["total_area"][L] * b["mass"][L]
    capR = b["area"] / b["total_area"][np.maximum(R, 0)] * b["mass"][np.maximum(R, 0)]
    assert np.all(dQL[0] <= capL * (1 + 1e-14))
    real = R >= 0
    assert np.all(-dQL[0][real] <= capR[real] * (1 + 1e-14))
    assert (np.abs(dQL[0] - capL) <= 1e-14 * capL).sum() > 10


def test_update_is_update_Q_then_get_W():
    m, _ = golden_mesh("per12")
    b = bundle(m, 1)
    b["grad"] = O.gradients(1, b)
    dQ, dQL, st, _ = O.face_flux_pass(1, b)
    Q, W, cst = O.update(1, 2, b["Q"], dQ, b["volume"])
    Wr, Qr = O.get_W(1, 2, b["volume"], b["Q"] - dQ)
    assert np.array_equal(W, Wr) and np.array_equal(Q, Qr)
    assert cst.max() == 0


def test_linear_field_gradient_is_exact():
    """The Voronoi gradient estimate is exact for linear fields; the slope
    limiter leaves it unchanged away from extrema."""
    m = M.voronoi_periodic(M.jittered_grid(24, 0.25, seed=4))
    W = np.zeros((4, m["C"]))
    x, y = m["pos"]
    W[0] = 2.0 + 0.3 * x - 0.2 * y
    W[3] = 3.0 + 0.1 * x + 0.4 * y
    W[1] = 0.5 * x
    W[2] = 0.25 * y
    mm = dict(m)
    mm["W"], mm["vel"] = W, np.zeros((2, m["C"]))
    g = O.gradients(0, mm)
    # periodic images break linearity at the box edge: interior cells only
    inner = (x > 0.15) & (x < 0.85) & (y > 0.15) & (y < 0.85)
    want = {(0, 0): 0.3, (1, 0): -0.2, (0, 3): 0.1, (1, 3): 0.4, (0, 1): 0.5, (1, 2): 0.25}
    for (d, k), val in want.items():
        gi = g[d, k, inner]
        exact = np.isclose(gi, val, rtol=1e-9, atol=1e-12)
        assert exact.mean() > 0.98
        # where the limiter acts it only scales the slope by alpha in [0, 1]
        a = gi[~exact] / val
        assert np.all((a >= 0.0) & (a <= 1.0 + 1e-9))


def test_linear_field_gradient_is_exact_3d():
    m = M.voronoi_periodic(M.poisson_points(2000, seed=4))
    W = np.zeros((5, m["C"]))
    x, y, z = m["pos"]
    W[0] = 2.0 + 0.3 * x - 0.2 * y + 0.1 * z
    W[4] = 3.0 + 0.1 * x + 0.4 * y - 0.3 * z
    W[1], W[2], W[3] = 0.5 * x, 0.25 * y, -0.5 * z
    mm = dict(m)
    mm["W"], mm["vel"] = W, np.zeros((3, m["C"]))
    g = O.gradients(0, mm)
    inner = np.all((m["pos"] > 0.2) & (m["pos"] < 0.8), axis=0)
    want = {(0, 0): 0.3, (1, 0): -0.2, (2, 0): 0.1, (0, 4): 0.1, (1, 4): 0.4, (2, 4): -0.3,
            (0, 1): 0.5, (1, 2): 0.25, (2, 3): -0.5}
    for (d, k), val in want.items():
        gi = g[d, k, inner]
        exact = np.isclose(gi, val, rtol=1e-9, atol=1e-12)
        assert exact.mean() > 0.95
        a = gi[~exact] / val
        assert np.all((a >= 0.0) & (a <= 1.0 + 1e-9))


@pytest.mark.parametrize("kind", [O.EXACT, O.HLLC])
def test_threaded_pass_equals_sequential(kind):
    """The CPU baseline's threaded pass (contiguous cell and face ranges, then
    the ordered gather) gives the sequential face loop's results bit for
    bit, with and without an activity mask."""
    m = M.voronoi_periodic(M.jittered_grid(32, 0.25, seed=6))
    b = M.prepare(m, M.kh_ics(m), G, kind=kind)
    for act in (None, (np.arange(m["C"]) % 3 != 0).astype(np.uint8)):
        bb = dict(b)
        if act is not None:
            bb["active"] = act
        g = O.gradients(kind, bb)
        bb["grad"] = g
        dQ, dQL, st, _ = O.face_flux_pass(kind, bb)
        g2, dQ2, dQL2, st2 = O.threaded_pass(kind, bb, 5)
        assert np.array_equal(g2[:, :, act != 0] if act is not None else g2,
                              g[:, :, act != 0] if act is not None else g)
        assert np.array_equal(dQL2, dQL) and np.array_equal(dQ2, dQ)
        assert np.array_equal(st2, st)


def test_flop_counter_counts_operations():
    """The counting build (oracle/libflops.so) computes what the oracle
    computes, bit for bit, and counts the operations of a hand-counted
    function: get_Q takes 1 + ndim products for mass and momenta, the squared
    speed (3 operations in 2D; in 3D the reference forms the 2D sum and then
    the 3D one, 3 + 5), and 6 for the energy (plus the solver's constants,
    counted once per call)."""
    rng = np.random.default_rng(2)
    for nd, per in ((2, 12), (3, 18)):
        for kind in (O.EXACT, O.HLLC):
            W = rng.uniform(0.5, 2.0, (nd + 2, 50))
            V = rng.uniform(0.1, 1.0, 50)
            _, init, _ = O.count_flops(O.params, kind, nd)
            Q, n, t = O.count_flops(O.get_Q, kind, nd, V, W)
            assert np.array_equal(Q, O.get_Q(kind, nd, V, W))
            assert init < 30
            assert n == per * 50 + init and t == 0, (nd, kind, n, init)
    m, _ = golden_mesh("p3d")
    b = bundle(m, 0)
    b["grad"] = O.gradients(0, b)
    r0 = O.face_flux_pass(0, b)
    r1, n, t = O.count_flops(O.face_flux_pass, 0, b)
    assert np.array_equal(r0[1], r1[1]) and np.array_equal(r0[0], r1[0])
    # the exact solver takes 2 pow per pressure function evaluation
    assert n > 500 * b["F"] and t > 0
