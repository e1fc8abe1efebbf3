"""CPU tests pinning the oracle's face-flux composition: the predictor chain of
VorFace::calculate_flux (src/VorFace.cpp:714-814: the half-step time
derivative, VorFace::reconstruct src/VorFace.hpp:174-196, the pairwise limiter,
the positivity fallbacks and the vacuum keep) and its periodic correction.

The reference's calculate_flux cannot run here (it takes a TimeLine, whose
constructors need Boost), so the chain is pinned by what the mathematics
fixes, independently of the oracle's formulas:

  * linear fields, sampled at the generators, have exact gradients (checked),
    so on every face the limiters leave alone the predicted states are known
    in closed form from the Euler equations in primitive form,
        drho/dt = -w.grad rho - rho div v,
        dv/dt   = -(w.grad) v - grad p / rho,
        dp/dt   = -w.grad p - gamma p div v,
    (w the cell's velocity in the face frame; the 2D reference evaluates the
    y-velocity's advection with d(v_y)/dy in place of d(v_y)/dx, which is
    kept) and the face state is
        W_face = W_cell + grad W . d + (dt/2) dW/dt,   d = midpoint - centroid;
    the face flux must then be dt A times the (golden-pinned) solver's flux of
    those two states;
  * a vacuum cell (rho = 0) keeps its raw state (the time derivative would
    divide by zero);
  * the pass is invariant under a periodic translation of all generators,
    which moves faces across the box boundary and so exercises the periodic
    correction of d.

A sign error in d, a dropped 1/2 or dt, or a sign error in the time
derivative changes the predicted states by O(|grad W| |d|) or
O(dt |grad W|) and fails these tests.
"""
import numpy as np
import pytest

import oracle as O
from shadowfax_amd import mesh as M

G = 5.0 / 3.0


def _mesh(nd):
    if nd == 2:
        return M.voronoi_periodic(M.jittered_grid(24, 0.25, seed=4))
    return M.voronoi_periodic(M.poisson_points(3000, seed=4))


def exact_gradient_cells(b, want, tol=1e-10):
    """Cells whose (oracle) gradient equals the analytic one: away from the
    periodic seam and where the cell-wise slope limiter does not act."""
    g = O.gradients(0, b)
    ok = np.ones(b["C"], bool)
    for (d, k), val in want.items():
        ok &= np.abs(g[d, k] - val) <= tol * max(1.0, abs(val))
    # every other gradient component is zero
    nz = np.ones_like(g, bool)
    for (d, k) in want:
        nz[d, k] = False
    ok &= np.all(np.where(nz, np.abs(g) <= tol, True), axis=(0, 1))
    return g, ok


def limiter_noop(phimid, phiL, phiR, dnorm, r, margin=1e-9):
    """True where VorFace::limit (src/VorFace.hpp:92-137) returns phimid
    unchanged: phimid inside the limiter's bounds with a margin."""
    d1 = 0.5 * np.abs(phiL - phiR)
    d2 = 0.25 * np.abs(phiL - phiR)
    lo, hi = np.minimum(phiL, phiR), np.maximum(phiL, phiR)
    with np.errstate(divide="ignore", invalid="ignore"):
        phiplus = np.where((hi + d1) * hi > 0, hi + d1, hi / (1 + d1 / np.abs(hi)))
        phiminus = np.where((lo - d1) * lo > 0, lo - d1, lo / (1 + d1 / np.abs(lo)))
    phibar = phiL + dnorm / r * (phiR - phiL)
    tolr = margin * (np.abs(phiL) + np.abs(phiR) + 1e-300)
    up = phiL < phiR
    ok_up = (phimid >= phiminus + tolr) & (phimid <= phibar + d2 - tolr)
    ok_dn = (phimid <= phiplus - tolr) & (phimid >= phibar - d2 + tolr)
    return np.where(phiL == phiR, phimid == phiL, np.where(up, ok_up, ok_dn))


def face_setup(b):
    """Per face: left/right cells, d vectors (interior faces: no periodic
    correction), unshifted centroid distance r."""
    L, R = b["left"], b["right"]
    interior = (R >= 0) & (b["ghost"] == 0)
    Rs = np.where(R >= 0, R, L)
    mid, cen = b["midpoint"], b["centroid"]
    dL = mid - cen[:, L]
    dR = mid - cen[:, Rs]
    r = np.sqrt(((cen[:, L] - cen[:, Rs]) ** 2).sum(0))
    return L, Rs, interior, dL, dR, r


def check_faces(b, kind, WLp, WRp, sel, tol=1e-11):
    """The oracle's dQL on faces sel equals dt A solve_for_flux(WLp, WRp)."""
    nd = b["ndim"]
    dQ, dQL, st, _ = O.face_flux_pass(kind, b)
    v, n = O.face_velocities(b)
    F, fst, _ = O.solve_for_flux(kind, nd, WLp[:, sel], WRp[:, sel], n[:, sel], v=v[:, sel])
    L, R = b["left"][sel], b["right"][sel]
    dt = np.minimum(b["dt"][L], b["dt"][R])
    want = dt * b["area"][sel] * F
    # the mass-flux limiter must not act on the compared faces (its cap is
    # the area share of the cell's mass times dt / max(dtL, dtR))
    tf = dt / np.maximum(b["dt"][L], b["dt"][R])
    capL = b["area"][sel] / b["total_area"][L] * b["mass"][L] * tf
    capR = b["area"][sel] / b["total_area"][R] * b["mass"][R] * tf
    free = (want[0] < 0.9 * capL) & (-want[0] < 0.9 * capR) & (fst == 0)
    assert free.mean() > 0.9
    got = dQL[:, sel][:, free]
    want = want[:, free]
    scale = np.abs(want).max(0) + 1e-300
    err = np.abs(got - want).max(0) / scale
    return err, int(free.sum())


@pytest.mark.parametrize("nd", [2, 3])
@pytest.mark.parametrize("kind", [O.EXACT, O.HLLC])
@pytest.mark.parametrize("steps", ["uniform", "hierarchy"])
def test_linear_pressure_predictor(nd, kind, steps):
    """Uniform density, uniform flow v0 with generators moving at v0 plus a
    perturbation (so faces move relative to the gas), p = p0 + g.x.  With
    individual time steps (a power-of-two hierarchy) a face predicts and
    integrates over the smaller of its two cells' steps
    (src/VorFace.cpp:695-699)."""
    m = _mesh(nd)
    C = m["C"]
    rng = np.random.default_rng(5)
    gvec = np.array([0.8, -0.5, 0.3][:nd])
    v0 = np.array([0.3, -0.2, 0.1][:nd])
    W = np.zeros((nd + 2, C))
    W[0] = 1.7
    W[1:1 + nd] = v0[:, None]
    W[-1] = 2.0 + gvec @ m["pos"]
    vel = v0[:, None] + 0.05 * rng.normal(size=(nd, C))
    b = M.prepare(m, W, G, kind=kind, vel=vel, dt=0.004)
    if steps == "hierarchy":
        b["dt"] = 0.004 * 0.5 ** rng.integers(0, 4, C)
    want = {(d, nd + 1): gvec[d] for d in range(nd)}
    b["grad"], good = exact_gradient_cells(b, want)
    L, R, interior, dL, dR, r = face_setup(b)
    v, n = O.face_velocities(b)
    w = v0[:, None] - v                               # cell velocity in the face frame
    rho0 = W[0, 0]
    hdt = 0.5 * np.minimum(b["dt"][L], b["dt"][R])
    pL = W[-1, L] + gvec @ dL                          # reconstructed pressures
    pR = W[-1, R] + gvec @ dR
    sel = interior & good[L] & good[R] & (b["area"] > 0)
    sel &= limiter_noop(pL, W[-1, L], W[-1, R], np.sqrt((dL ** 2).sum(0)), r)
    sel &= limiter_noop(pR, W[-1, R], W[-1, L], np.sqrt((dR ** 2).sum(0)), r)
    assert sel.sum() > 0.2 * interior.sum()
    WLp = np.empty((nd + 2, b["F"]))
    WRp = np.empty_like(WLp)
    for Wp, p in ((WLp, pL), (WRp, pR)):
        Wp[0] = rho0                                   # drho/dt = 0
        Wp[1:1 + nd] = w - hdt * gvec[:, None] / rho0  # dv/dt = -grad p/rho
        Wp[-1] = p - hdt * (w * gvec[:, None]).sum(0)  # dp/dt = -w.grad p
    err, nf = check_faces(b, kind, WLp, WRp, sel)
    print("linear pressure nd %d kind %d: %d faces, max err %.3e" % (nd, kind, nf, err.max()))
    assert nf > 100
    assert err.max() <= 1e-11


@pytest.mark.parametrize("nd", [2, 3])
@pytest.mark.parametrize("kind", [O.EXACT, O.HLLC])
def test_linear_velocity_predictor(nd, kind):
    """Uniform density and pressure, v = v0 + A x on a static mesh: the
    predictor carries -rho div v, -(w.grad) v and -gamma p div v."""
    m = _mesh(nd)
    C = m["C"]
    A = np.array([[0.30, -0.20, 0.10], [0.25, -0.15, 0.05], [-0.10, 0.20, 0.12]])[:nd, :nd]
    v0 = np.array([0.2, -0.1, 0.05][:nd])
    W = np.zeros((nd + 2, C))
    W[0] = 1.3
    W[1:1 + nd] = v0[:, None] + A @ m["pos"]
    W[-1] = 2.5
    dt = 0.01
    b = M.prepare(m, W, G, kind=kind, vel=np.zeros((nd, C)), dt=dt)
    # grad[d][1 + k] = dv_k/dx_d = A[k, d]
    want = {(d, 1 + k): A[k, d] for d in range(nd) for k in range(nd)}
    b["grad"], good = exact_gradient_cells(b, want)
    L, R, interior, dL, dR, r = face_setup(b)
    v, n = O.face_velocities(b)
    assert np.abs(v).max() == 0.0                      # static mesh
    hdt = 0.5 * dt
    div = np.trace(A)
    sel = interior & good[L] & good[R] & (b["area"] > 0)
    WLp = np.empty((nd + 2, b["F"]))
    WRp = np.empty_like(WLp)
    for Wp, cell, d, other in ((WLp, L, dL, R), (WRp, R, dR, L)):
        wv = W[1:1 + nd, cell]                         # face frame = lab frame
        vrec = wv + A @ d
        for k in range(nd):
            sel &= limiter_noop(vrec[k], W[1 + k, cell], W[1 + k, other],
                                np.sqrt((d ** 2).sum(0)), r)
        adv = A @ wv                                   # (w.grad) v
        if nd == 2:
            # the 2D reference uses d(v_y)/dy for both terms of v_y
            adv[1] = (wv[0] + wv[1]) * A[1, 1]
        Wp[0] = W[0, cell] * (1.0 - hdt * div)
        Wp[1:1 + nd] = vrec - hdt * adv
        Wp[-1] = W[-1, cell] - hdt * G * W[-1, cell] * div
    assert sel.sum() > 0.1 * interior.sum()
    err, nf = check_faces(b, kind, WLp, WRp, sel)
    print("linear velocity nd %d kind %d: %d faces, max err %.3e" % (nd, kind, nf, err.max()))
    assert nf > 100
    assert err.max() <= 1e-11


@pytest.mark.parametrize("name", ["per", "mir"])
def test_vacuum_cells_keep_their_raw_state(name):
    """With rho = 0 in some cells the predictor's 1/rho gives NaN; the
    reference then keeps the raw (boosted) state of that side
    (src/VorFace.cpp:809-814).  With zero gradients in the other cells every
    face's flux is then dt A solve_for_flux of the raw boosted states."""
    import test_oracle_fluxpass as TF
    m, _ = TF.golden_mesh("per12" if name == "per" else "mir10")
    m = dict(m)
    W = m["W"].copy()
    vac = np.arange(m["C"]) % 7 == 3
    W[:, vac] = 0.0
    m["W"] = W
    nd = m["ndim"]
    rng = np.random.default_rng(3)
    for kind in (O.EXACT, O.HLLC):
        b = M.prepare(m, W, G, kind=kind, vel=m["vel"], dt=1e-3)
        # zero gradients, except in the vacuum cells: without the keep their
        # reconstructed density would leave zero and their predicted
        # velocity divide by it
        b["grad"] = np.zeros((nd, nd + 2, m["C"]))
        b["grad"][:, :, vac] = rng.normal(size=(nd, nd + 2, vac.sum()))
        dQ, dQL, st, _ = O.face_flux_pass(kind, b)
        v, n = O.face_velocities(b)
        L, R = b["left"], b["right"]
        WL = b["W"][:, L].copy()
        WR = np.where(R >= 0, b["W"][:, np.maximum(R, 0)], 0.0)
        WL[1:1 + nd] -= v
        WR[1:1 + nd] -= v
        F, sts, _ = O.solve_for_flux(kind, nd, WL, WR, n, v=v, refl=(R < 0).astype(np.uint8))
        want = (1e-3 * b["area"]) * F
        touch = vac[L] | (vac[np.maximum(R, 0)] & (R >= 0))
        assert touch.sum() > 10
        # faces where the mass-flux limiter acts (zero-mass cells) are left out
        capL = b["area"] / b["total_area"][L] * b["mass"][L]
        Rs = np.maximum(R, 0)
        capR = np.where(R >= 0, b["area"] / b["total_area"][Rs] * b["mass"][Rs], np.inf)
        free = (want[0] <= capL) & (-want[0] <= capR) & (b["area"] != 0)
        assert (touch & free).sum() > 5
        assert np.all(np.isfinite(dQL))
        assert np.array_equal(dQL[:, free], want[:, free])


@pytest.mark.parametrize("nd", [2, 3])
@pytest.mark.parametrize("kind", [O.EXACT, O.HLLC])
def test_periodic_correction_by_translation(nd, kind):
    """Translating every generator by the same vector (modulo the box) gives
    the same tessellation with other faces across the boundary, each cell's
    state moving with its generator.  A face that is interior in the first
    mesh and crosses the boundary in the second reaches its right cell there
    through a periodic copy, whose d carries the box shift
    (src/VorFace.cpp:729-734): d is the same vector in both meshes, taken
    here from the first mesh's interior geometry.  The limiter's distance r
    uses the unshifted centroids (src/VorFace.cpp:740-742), so it spans the
    box in the second mesh.  Only the density varies (gas at rest, uniform
    pressure, moving generators), so each face's predicted states are
        rho = limit(rho_c + grad rho . d, rho_c, rho_o, d, r)
              - dt/2 w . grad rho,   v = w,   p = p0
    (w = -v_face, the gas velocity in the face frame; the limiter's own
    values are pinned by hand-computed cases), and every face's flux must be
    dt A times the solver's flux of those states, in both meshes."""
    if nd == 2:
        pos = M.jittered_grid(20, 0.25, seed=8)
    else:
        pos = M.poisson_points(1500, seed=8)
    C = pos.shape[1]
    rng = np.random.default_rng(12)
    W = np.zeros((nd + 2, C))
    W[0] = 1.0 + 0.3 * np.sin(2 * np.pi * pos[0]) + 0.3 * rng.random(C)
    W[-1] = 1.5
    vel = 0.2 * rng.normal(size=(nd, C))
    dt = 2e-3
    out = []
    for s in (np.zeros(nd), np.array([0.37, 0.61, 0.23][:nd])):
        m = M.voronoi_periodic(np.mod(pos + s[:, None], 1.0))
        b = M.prepare(m, W, G, kind=kind, vel=vel, dt=dt)
        b["grad"] = O.gradients(kind, b)
        dQ, dQL, st, _ = O.face_flux_pass(kind, b)
        assert st.max() == 0
        out.append((b, dQL))
    (b0, q0), (b1, q1) = out
    gscale = np.abs(b0["grad"][:, 0]).max()
    assert np.abs(b0["grad"][:, 0] - b1["grad"][:, 0]).max() <= 1e-12 * gscale
    # faces interior in mesh 0 and across the boundary in mesh 1, by cell pair
    key0 = {}
    for f in np.where((b0["right"] >= 0) & (b0["ghost"] == 0))[0]:
        key0.setdefault((b0["left"][f], b0["right"][f]), []).append(f)
    pairs = []
    for f in np.where(b1["ghost"] == 1)[0]:
        c = key0.get((b1["left"][f], b1["right"][f]))
        if c is not None and len(c) == 1:
            pairs.append((c[0], f))
    f0, f1 = np.array(pairs).T
    assert len(f0) > 20
    L, R = b0["left"][f0], b0["right"][f0]
    cen0, cen1 = b0["centroid"], b1["centroid"]
    mid = b0["midpoint"][:, f0]
    d = {"L": mid - cen0[:, L], "R": mid - cen0[:, R]}
    r0 = np.sqrt(((cen0[:, L] - cen0[:, R]) ** 2).sum(0))
    r1 = np.sqrt(((cen1[:, L] - cen1[:, R]) ** 2).sum(0))
    assert np.all(r1 > 2 * r0)              # the unshifted distance spans the box
    gr = b0["grad"][:, 0]                   # density gradients (invariant)
    clamped_differently = 0
    for b, q, fs, r in ((b0, q0, f0, r0), (b1, q1, f1, r1)):
        vf, n = O.face_velocities(b)
        vf, n = vf[:, fs], n[:, fs]
        states = {}
        for side, cell, other in (("L", L, R), ("R", R, L)):
            dd = d[side]
            rec = W[0, cell] + (gr[:, cell] * dd).sum(0)
            lim = np.array([O.limit(rec[j], W[0, cell[j]], W[0, other[j]], dd[:, j], r[j], nd)
                            for j in range(len(fs))])
            if r is r1:
                lim0 = np.array([O.limit(rec[j], W[0, cell[j]], W[0, other[j]], dd[:, j],
                                         r0[j], nd) for j in range(len(fs))])
                clamped_differently += int((lim != lim0).sum())
            Wp = np.zeros((nd + 2, len(fs)))
            Wp[0] = lim + 0.5 * dt * (vf * gr[:, cell]).sum(0)   # -dt/2 w.grad rho, w = -vf
            Wp[1:1 + nd] = -vf
            Wp[-1] = W[-1, 0]
            states[side] = Wp
        F, fst, _ = O.solve_for_flux(kind, nd, states["L"], states["R"], n, v=vf)
        want = dt * b["area"][fs] * F
        err = np.abs(q[:, fs] - want).max(0) / np.abs(want).max(0)
        print("translation nd %d kind %d: %d boundary faces, max err %.3e"
              % (nd, kind, len(fs), err.max()))
        assert err.max() <= 1e-12
    # the unshifted distance changes the limiter's result on some faces
    assert clamped_differently > 5
