"""The oracle on stored inputs (tests/stored_oracle.py), CPU only.

  * given the host's FP64 per-face geometry (sfx_face_geometry), the face
    pass is bit for bit the one that forms the geometry from positions,
    centroids and midpoints (the same expressions);
  * what FP32 storage of the gradients and face geometry costs, with FP64
    arithmetic on the stored values: on Kelvin-Helmholtz and Sedov states
    (the benchmark workloads' states) no face moves by more than 1e-6; on the
    golden 3D meshes, whose cold gas (sound speed 4e-3) moves at 0.3 in
    random directions, single faces move by a few 1e-6 and every cell stays
    within 1e-6 of its gross flux.  This storage bound is what the mixed
    pass's per-face error is held to on such meshes (tests/test_gpu_fluxpass.py).
"""
import os

import numpy as np
import pytest

import oracle as O
import stored_oracle as SO
from shadowfax_amd import mesh as M
from shadowfax_amd.metrics import face_flux_scale, fluxpass_errors

HERE = os.path.dirname(os.path.abspath(__file__))
G = 5.0 / 3.0
KEYS = ("pos", "volume", "centroid", "left", "right", "area", "midpoint", "shift", "ghost",
        "twin", "cf_off", "cf_face", "W", "vel")


def golden_bundle(name, kind):
    nd = 3 if name.endswith("3d") else 2
    z = np.load(os.path.join(HERE, "golden", "mesh_%dd.npz" % nd))
    m = {k: z[name + "_" + k] for k in KEYS}
    m.update(ndim=nd, C=m["volume"].shape[0], F=m["area"].shape[0], box=np.ones(nd))
    return M.prepare(m, m["W"], G, kind=kind, vel=m["vel"])


@pytest.mark.parametrize("name", ["per12", "mir10", "p3d", "m3d"])
@pytest.mark.parametrize("kind", [O.EXACT, O.HLLC])
def test_given_fp64_geometry_is_bit_identical(name, kind):
    b = golden_bundle(name, kind)
    ob = dict(b)
    ob["grad"] = O.gradients(kind, b)
    ref = O.face_flux_pass(kind, ob)
    for k in SO.GEO:
        ob["geo_" + k] = b[k]
    got = O.face_flux_pass(kind, ob)
    for a, r in zip(got[:3], ref[:3]):
        assert np.array_equal(a, r)
    # and a planted slip in the stored geometry shows
    ob["geo_mc"] = b["mc"] * (1 + 1e-9)
    assert not np.array_equal(O.face_flux_pass(kind, ob)[1], ref[1])


def storage_cost(b, kind):
    ob = dict(b)
    ob["grad"] = O.gradients(kind, b)
    dQ, dQL, _, _ = O.face_flux_pass(kind, ob)
    sdQ, sdQL, _ = SO.stored_pass(kind, b, ob["grad"])
    return fluxpass_errors(b, G, dQL, dQ, sdQL, sdQ)


@pytest.mark.parametrize("case", ["kh48", "sedov4000"])
def test_storage_cost_on_benchmark_states_below_1e6(case):
    if case == "kh48":
        m = M.voronoi_periodic(M.jittered_grid(48, 0.25, seed=9))
        kind, b = O.HLLC, None
        b = M.prepare(m, M.kh_ics(m), G, kind=kind)
    else:
        m = M.voronoi_periodic(M.hilbert_sorted(M.poisson_points(4000, seed=9)))
        kind = O.EXACT
        b = M.prepare(m, M.sedov_ics(m, ncells=32), G, kind=kind)
    e = storage_cost(b, kind)
    print(case, e)
    assert e["face_max"] <= 1e-6 and e["cell_gross_max"] <= 1e-6


@pytest.mark.parametrize("name", ["p3d", "m3d"])
def test_storage_cost_on_hypersonic_golden_meshes(name):
    """Single faces move by more than 1e-6; the cells stay within 1e-6; the
    sensitivity estimate covers the storage cost and is itself bounded."""
    b = golden_bundle(name, O.EXACT)
    e = storage_cost(b, O.EXACT)
    print(name, e)
    assert 1e-6 < e["face_max"] <= 1e-5
    assert e["cell_gross_max"] <= 1e-6
    g = O.gradients(O.EXACT, b)
    ob = dict(b)
    ob["grad"] = g
    _, dQL, _, _ = O.face_flux_pass(O.EXACT, ob)
    _, sdQL, _ = SO.stored_pass(O.EXACT, b, g)
    sto = SO.face_err(b, sdQL, dQL, face_flux_scale(b, G))
    sens = SO.sensitivity(O.EXACT, b, g)
    assert np.all(sto <= sens) and sens.max() <= 2e-5
    assert np.any(sto > 1e-6)
