"""Host tessellator (the mesh exporter the flux pass reads), CPU only.

  * face area/midpoint and cell volume/centroid are bit-identical to what the
    reference classes (VorFace::calculate_midpoint/calculate_area,
    VorCell::calculate_volume/calculate_centroid) compute from the same
    stored vertex lists (golden vectors from the compiled reference);
  * the tessellation itself is pinned independently of the reference: cell
    volumes and face areas against Qhull (scipy.spatial) on a 3x3(x3) tiled
    copy of the periodic box, volumes summing to the box, each cell closed
    (sum_f A_f n_f = 0), and the face-list contract (each real-real face in
    both cells' lists, each periodic face once per side with a twin).
"""
import os

import numpy as np
import pytest

from shadowfax_amd import mesh as M

HERE = os.path.dirname(os.path.abspath(__file__))


def golden_points(nd, name):
    if nd == 2:
        n = {"per12": 12, "mir10": 10}[name]
        return M.jittered_grid(n, 0.25, seed=11 + n)
    n = {"p3d": 160, "m3d": 120}[name]
    return M.poisson_points(n, seed=31 + n, ndim=3)


@pytest.mark.parametrize("nd,name", [(2, "per12"), (2, "mir10"), (3, "p3d"), (3, "m3d")])
def test_geometry_matches_reference_classes_bitwise(nd, name):
    z = np.load(os.path.join(HERE, "golden", "mesh_%dd.npz" % nd))
    m = M.voronoi_periodic(golden_points(nd, name))
    for k, ref in (("area", "ref_area"), ("midpoint", "ref_midpoint"),
                   ("volume", "ref_volume"), ("centroid", "ref_centroid")):
        assert np.array_equal(m[k], z[name + "_" + ref]), k


def closure(m):
    pos, L, R, A = m["pos"], m["left"], m["right"], m["area"]
    n = pos[:, R] + m["shift"] - pos[:, L]
    n = n / np.linalg.norm(n, axis=0)
    acc = np.zeros_like(pos)
    np.add.at(acc.T, L, (A * n).T)
    real = m["ghost"] == 0
    np.add.at(acc.T, R[real], -(A * n)[:, real].T)
    return np.abs(acc).max()


@pytest.mark.parametrize("nd", [2, 3])
def test_volumes_closure_and_face_contract(nd):
    pos = M.jittered_grid(40, 0.25, seed=5) if nd == 2 else M.poisson_points(3000, seed=5)
    m = M.voronoi_periodic(pos)
    C, F = m["C"], m["F"]
    assert abs(m["volume"].sum() - 1.0) <= 1e-13
    assert closure(m) <= 1e-14
    L, R, gh, tw = m["left"], m["right"], m["ghost"], m["twin"]
    # each face listed by its left cell; a real-real face also by its right
    # cell, a periodic face only once (its twin is the other side's copy)
    cnt = np.zeros(F, np.int64)
    off, cf = m["cf_off"], m["cf_face"]
    owner = np.repeat(np.arange(C), np.diff(off))
    np.add.at(cnt, cf, 1)
    assert np.all(cnt == np.where(gh == 1, 1, 2))
    assert np.all((owner == L[cf]) | ((owner == R[cf]) & (gh[cf] == 0)))
    assert np.all(L[gh == 0] < R[gh == 0])
    g = np.where(gh == 1)[0]
    assert np.all(tw[g] >= 0) and np.all(tw[gh == 0] == -1)
    assert np.array_equal(tw[tw[g]], g)
    assert np.array_equal(L[tw[g]], R[g]) and np.array_equal(R[tw[g]], L[g])
    assert np.array_equal(m["shift"][:, tw[g]], -m["shift"][:, g])
    # the two copies of a periodic face agree to round-off
    assert np.abs(m["area"][tw[g]] - m["area"][g]).max() <= 1e-15
    # centroids inside the box around their generator
    d = m["centroid"] - m["pos"]
    assert np.abs(d).max() < 0.1


@pytest.mark.parametrize("nd", [2, 3])
def test_cell_closure_vanishes_on_voronoi_cells(nd):
    """sfx_cell_closure on the host tessellation: every cell closes to
    round-off, so prepare() leaves the mixed pass's correction out."""
    pos = M.jittered_grid(40, 0.25, seed=5) if nd == 2 else M.poisson_points(3000, seed=5)
    m = M.voronoi_periodic(pos)
    b = M.prepare(m, M.sod_ics(m), 5.0 / 3.0)
    S = M.cell_closure(b)
    assert S.shape == (nd, b["C"])
    assert (np.abs(S).max(0) / b["total_area"]).max() <= 1e-13
    assert b["closure"] is None


def _golden_mesh(nd, name):
    z = np.load(os.path.join(HERE, "golden", "mesh_%dd.npz" % nd))
    keys = ("pos", "volume", "centroid", "left", "right", "area", "midpoint", "shift", "ghost",
            "twin", "cf_off", "cf_face", "W", "vel")
    m = {k: z[name + "_" + k] for k in keys}
    m.update(ndim=nd, C=m["volume"].shape[0], F=m["area"].shape[0], box=np.ones(nd))
    return m


@pytest.mark.parametrize("nd,name", [(2, "mir10"), (3, "m3d")])
def test_cell_closure_of_open_cells_and_the_gradient_identity(nd, name):
    """The golden mirror meshes turn periodic faces into mirror faces without
    moving them, so their wall cells do not close.  sfx_cell_closure equals a
    per-cell sum over each cell's faces of A (x_i - x_j)/|x_i - x_j| (x_j the
    neighbour's generator: the periodic image, or the mirror image of x_i),
    and with it the reference's gradient sum (src/VorCell.cpp:493-494)
    equals the difference form minus W_i * closure_i for arbitrary states
    (a sign or factor slip in either fails)."""
    m = _golden_mesh(nd, name)
    b = M.prepare(m, m["W"], 5.0 / 3.0, vel=m["vel"])
    S = M.cell_closure(b)
    C = b["C"]
    pos, L, R, sh = m["pos"], m["left"], m["right"], m["shift"]
    rng = np.random.default_rng(7)
    worst_rel, worst_id = 0.0, 0.0
    for i in range(C):
        Si = np.zeros(nd)
        Wi = rng.normal(size=nd + 2)
        ref = np.zeros((nd, nd + 2))
        diff = np.zeros((nd, nd + 2))
        for f in m["cf_face"][m["cf_off"][i]:m["cf_off"][i + 1]]:
            A = m["area"][f]
            if A == 0:
                continue
            if L[f] == i:
                xj = pos[:, R[f] if R[f] >= 0 else i] + sh[:, f]
            else:
                xj = pos[:, L[f]]
            rij = pos[:, i] - xj
            rn = np.linalg.norm(rij)
            c = m["midpoint"][:, f] - 0.5 * (pos[:, i] + xj)
            Si += A * rij / rn
            Wj = rng.normal(size=nd + 2)
            for l in range(nd):
                ref[l] += A * ((Wj - Wi) * c[l] / rn - 0.5 * (Wi + Wj) * rij[l] / rn)
                diff[l] += (Wj - Wi) * (A * (c[l] - 0.5 * rij[l]) / rn)
        worst_rel = max(worst_rel, np.abs(S[:, i] - Si).max() / b["total_area"][i])
        corr = diff - np.outer(S[:, i], Wi)
        worst_id = max(worst_id, np.abs(ref - corr).max() / max(np.abs(ref).max(), 1e-300))
    assert worst_rel <= 1e-14
    assert worst_id <= 1e-12
    # the wall cells do not close, so prepare() hands the correction on
    rel = np.abs(S).max(0) / b["total_area"]
    assert rel.max() > 1e-3 and b["closure"] is not None


@pytest.mark.parametrize("nd,C", [(2, 300), (3, 400)])
def test_tessellation_matches_qhull(nd, C):
    spatial = pytest.importorskip("scipy.spatial")
    rng = np.random.default_rng(17 + nd)
    pos = rng.random((nd, C))
    m = M.voronoi_periodic(pos)
    tiles = np.array(np.meshgrid(*[(-1, 0, 1)] * nd, indexing="ij")).reshape(nd, -1).T
    center = int(np.where((tiles == 0).all(1))[0][0])
    P = np.concatenate([pos.T + t for t in tiles])
    vor = spatial.Voronoi(P)
    vol = np.array([spatial.ConvexHull(vor.vertices[vor.regions[vor.point_region[center * C + i]]])
                    .volume for i in range(C)])
    assert np.abs(vol - m["volume"]).max() <= 1e-12 * m["volume"].min()
    # faces: (cell, neighbour, shift) -> area
    ref = {}
    for (p, q), rv in zip(vor.ridge_points, vor.ridge_vertices):
        for a, b in ((p, q), (q, p)):
            if a // C != center:
                continue
            V = vor.vertices[rv]
            if nd == 2:
                area = np.linalg.norm(V[0] - V[1])
            else:
                c = V.mean(0)
                area = sum(0.5 * np.linalg.norm(np.cross(V[k] - c, V[(k + 1) % len(V)] - c))
                           for k in range(len(V)))
            ref[(a % C, b % C, tuple(tiles[b // C]))] = area
    mine = {}
    for f in range(m["F"]):
        sh = tuple(np.round(m["shift"][:, f]).astype(int))
        mine[(m["left"][f], m["right"][f], sh)] = m["area"][f]
        if not m["ghost"][f]:
            mine[(m["right"][f], m["left"][f], sh)] = m["area"][f]
    assert set(mine) == set(ref)
    assert max(abs(mine[k] - ref[k]) for k in ref) <= 1e-13


@pytest.mark.parametrize("nd,n", [(2, 16), (3, 8)])
def test_hilbert_order_is_a_hilbert_curve(nd, n):
    """Points at the centres of an n^nd grid, in Hilbert order, form a path
    of unit steps visiting every grid cell once (the defining property of the
    curve); random points come out as a permutation with short steps."""
    g = np.stack(np.meshgrid(*[np.arange(n)] * nd, indexing="ij")).reshape(nd, -1)
    rng = np.random.default_rng(3)
    g = g[:, rng.permutation(g.shape[1])]
    pos = (g + 0.5) / n
    order = M.hilbert_order(pos)
    assert np.array_equal(np.sort(order), np.arange(g.shape[1]))
    steps = np.abs(np.diff(g[:, order], axis=1)).sum(0)
    assert np.all(steps == 1)
    p = M.poisson_points(20000, seed=1, ndim=nd)
    o = M.hilbert_order(p)
    assert np.array_equal(np.sort(o), np.arange(20000))
    d = np.linalg.norm(np.diff(p[:, o], axis=1), axis=0)
    assert np.median(d) < 2.0 * 20000 ** (-1.0 / nd)


def test_entry_groups_by_hand():
    """sfx_mesh_entry_groups: groups of 64 consecutive cells, each as many
    record rows as its cell with the most entries; the last group short."""
    import ctypes as ct
    from shadowfax_amd import _capi as C
    counts = np.array([3] * 64 + [5] + [2] * 63 + [7, 1], np.int64)  # 130 cells
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    grp = np.zeros(4, np.int64)
    p = lambda a: a.ctypes.data_as(ct.POINTER(ct.c_int64))
    rows = C.lib().sfx_mesh_entry_groups(130, p(off), p(grp))
    assert rows == 3 + 5 + 7 and list(grp) == [0, 3, 8, 15]
    assert C.lib().sfx_mesh_entry_groups(-1, p(off), p(grp)) == -1


def _sorted_groups(counts, window):
    import ctypes as ct
    from shadowfax_amd import _capi as C
    C_ = len(counts)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    slot = np.zeros(C_, np.int64)
    grp = np.zeros((C_ + 63) // 64 + 1, np.int64)
    p = lambda a: a.ctypes.data_as(ct.POINTER(ct.c_int64))
    rows = C.lib().sfx_mesh_entry_groups_sorted(C_, p(off), window, p(slot), p(grp))
    return rows, slot, grp


def test_sorted_entry_groups_by_hand():
    """sfx_mesh_entry_groups_sorted: within each window the cells take slots
    in descending order of entry count, ties in cell order; a group's rows
    are its first slot's count; bad windows are refused."""
    counts = np.array([3] * 64 + [5] + [2] * 63 + [7, 1], np.int64)  # 130 cells
    rows, slot, grp = _sorted_groups(counts, 128)
    cells, n = slot & 0xFFFFFFFF, slot >> 32
    # window 0 (cells 0..127): cell 64 (5 entries), then the 3s, then the 2s
    assert list(cells[:66]) == [64] + list(range(64)) + [65]
    assert list(n[:3]) == [5, 3, 3] and n[64] == 3 and n[65] == 2
    # window 1 (cells 128, 129): 7 then 1
    assert list(cells[128:]) == [128, 129] and list(n[128:]) == [7, 1]
    # group 1 starts with cell 63 (3 entries): sorting moves the padding, it
    # cannot remove what the counts force
    assert rows == 5 + 3 + 7 and list(grp) == [0, 5, 8, 15]
    for bad in (0, 63, 100, -64):
        assert _sorted_groups(counts, bad)[0] == -1


@pytest.mark.parametrize("window", [64, 1024])
def test_sorted_entry_groups_on_a_poisson_mesh(window):
    """On a 3D Poisson-Voronoi mesh the sorted slots are a permutation that
    keeps every cell in its window, counts descend within each window, the
    groups' rows are their maxima, and sorting over windows of 1024 cells
    leaves far fewer padding rows than cell order (windows of 64 cells sort
    within each group, which leaves the rows as they were)."""
    import ctypes as ct
    from shadowfax_amd import _capi as C
    m = M.voronoi_periodic(M.hilbert_sorted(M.poisson_points(6000, seed=5)))
    counts = np.diff(np.asarray(m["cf_off"], np.int64))
    rows, slot, grp = _sorted_groups(counts, window)
    cells, n = slot & 0xFFFFFFFF, slot >> 32
    assert np.array_equal(np.sort(cells), np.arange(len(counts)))
    assert np.array_equal(n, counts[cells])
    assert np.array_equal(cells // window, np.arange(len(counts)) // window)
    for w0 in range(0, len(counts), window):
        nw = n[w0:w0 + window]
        assert np.all(np.diff(nw) <= 0)
        cw = cells[w0:w0 + window]
        tie = np.diff(nw) == 0
        assert np.all(np.diff(cw)[tie] > 0)
    ng = (len(counts) + 63) // 64
    pad = np.zeros(ng * 64, np.int64)
    pad[:len(n)] = n
    assert np.array_equal(np.diff(grp), pad.reshape(ng, 64).max(1))
    grp0 = np.zeros(ng + 1, np.int64)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    p = lambda a: a.ctypes.data_as(ct.POINTER(ct.c_int64))
    rows0 = C.lib().sfx_mesh_entry_groups(len(counts), p(off), p(grp0))
    if window == 64:
        assert rows == rows0
    else:
        assert rows < 0.8 * rows0
