"""GPU parity of the face-flux pass (gradients, face fluxes, ordered gather,
update) through the C ABI, against the oracle on the same meshes.

  FP64 pass: gradients bit-exact (no transcendental on that path); face
  fluxes within 1e-12 of the oracle (normwise per face); the per-cell sums
  are the ordered sums of the GPU's own face fluxes, bit for bit (the
  scatter order of the reference's face loop); update bit-exact given dQ.
  Mixed pass: face fluxes within 1e-6 of the FP64 pass (normwise per face,
  with the same flux floor as the solver tests); per-cell dQ within 1e-6 of
  the FP64 pass relative to the cell's gross flux sum_f |dQL_f| (dQ is a sum
  of face fluxes that largely cancel, so |dQ| itself is no scale); exactly one
  evaluation per periodic face pair, so the totals cancel to FP64 round-off.
"""
import os

import numpy as np
import pytest

import oracle as O
import stored_oracle as SO
from shadowfax_amd import mesh as M
from shadowfax_amd.fluxpass import FluxPass
from shadowfax_amd.metrics import face_flux_scale, fluxpass_errors
from shadowfax_amd.riemann import RiemannSolver

pytestmark = pytest.mark.gpu
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


def sedov_bundle(n, kind, seed=9):
    m = M.voronoi_periodic(M.hilbert_sorted(M.poisson_points(n, seed=seed)))
    return M.prepare(m, M.sedov_ics(m, ncells=min(32, n // 50)), G, kind=kind)


def kh_bundle(n, kind, seed=9):
    m = M.voronoi_periodic(M.jittered_grid(n, 0.25, seed=seed))
    return M.prepare(m, M.kh_ics(m), G, kind=kind)


def run_gpu(b, kind, precision, dev):
    import torch
    s = RiemannSolver(kind, b["ndim"], G, 1e-8, -5.0, precision)
    fp = FluxPass(s, M.DeviceMesh(b, precision, dev))
    fp.run()
    torch.cuda.synchronize()
    out = {"grad": fp.m.grad_soa(), "dQL": fp.m.dQL_soa(), "dQ": fp.m.dQ_soa()}
    out["fstatus"] = fp.m.fstatus.cpu().numpy()
    out["fp"] = fp
    return out


def oracle_pass(b, kind):
    ob = dict(b)
    ob["grad"] = O.gradients(kind, b)
    dQ, dQL, st, _ = O.face_flux_pass(kind, ob)
    return ob["grad"], dQ, dQL, st


def face_scale(b, kind):
    return face_flux_scale(b, G)


@pytest.mark.parametrize("name", ["per12", "mir10", "p3d", "m3d"])
@pytest.mark.parametrize("kind", [0, 1])
def test_fp64_pass_matches_oracle(device, name, kind):
    print("case", name, kind, flush=True)
    b = golden_bundle(name, kind)
    g = run_gpu(b, kind, "fp64", device)
    og, odQ, odQL, ost = oracle_pass(b, kind)
    assert np.array_equal(g["grad"], og), "gradients must be bit-exact"
    assert np.array_equal(g["fstatus"] != 0, ost != 0)
    den = np.maximum(np.abs(odQL).max(0), face_scale(b, kind))
    den = np.where(den > 0, den, 1.0)
    err = (np.abs(g["dQL"] - odQL).max(0) / den).max()
    bit = np.mean(np.all(g["dQL"] == odQL, axis=0))
    print("fp64 %s kind %d: face err %.3e bitwise faces %.4f" % (name, kind, err, bit))
    assert err <= 1e-12
    errc = np.abs(g["dQ"] - odQ).max() / np.abs(odQ).max()
    assert errc <= 1e-12


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("kind", [0, 1])
def test_gather_is_ordered_sum_of_gpu_face_fluxes(device, kind, prec):
    """dQ on the GPU equals, bit for bit, the sum of the GPU's own dQL in the
    reference's face-loop order (left += dQL, right += -dQL), summed in FP64.
    The mixed pass evaluates each periodic pair once and writes the other
    copy's row as the exact negation."""
    b = kh_bundle(32, kind)
    g = run_gpu(b, kind, prec, device)
    dQL = np.asarray(g["dQL"], np.float64)
    if prec == "mixed":
        gh = np.where(b["ghost"] == 1)[0]
        assert len(gh) > 0
        assert np.array_equal(dQL[:, gh], -dQL[:, b["twin"][gh]])
    acc = np.zeros_like(g["dQ"])
    for f in range(b["F"]):
        acc[:, b["left"][f]] += dQL[:, f]
        if b["right"][f] >= 0 and not b["ghost"][f]:
            acc[:, b["right"][f]] += -dQL[:, f]
    assert np.array_equal(acc, g["dQ"])
    # the vectorised restatement over the gather entries agrees
    assert np.array_equal(ordered_sums(b, dQL), g["dQ"])


@pytest.mark.parametrize("nd", [2, 3])
@pytest.mark.parametrize("kind", [0, 1])
def test_mixed_pass_within_1e6_and_conservative(device, kind, nd):
    print("case kind %d nd %d" % (kind, nd), flush=True)
    b = kh_bundle(48, kind) if nd == 2 else sedov_bundle(4000, kind)
    g64 = run_gpu(b, kind, "fp64", device)
    gmx = run_gpu(b, kind, "mixed", device)
    errs = mixed_errors(b, kind, g64, gmx)
    print("mixed vs fp64 kind %d nd %d: %s" % (kind, nd, errs))
    assert errs["face_max"] <= 1e-6
    assert errs["cell_gross_max"] <= 1e-6
    # one evaluation per periodic pair: totals cancel to FP64 round-off
    # (relative to the largest face flux)
    assert errs["conservation"] <= 1e-13


@pytest.mark.parametrize("name", ["per12", "mir10", "p3d", "m3d"])
@pytest.mark.parametrize("kind", [0, 1])
def test_mixed_pass_on_golden_meshes(device, name, kind):
    """The mixed pass against the FP64 pass on the golden periodic and mirror
    meshes.  The mirror meshes' wall cells do not close (their mirror faces
    are periodic faces turned into walls), so the mixed gradients take the
    closure correction there; the periodic ones need none.

    2D (Kelvin-Helmholtz states): every face within 1e-6.  3D: cold gas
    (sound speed 4e-3) moving at 0.3 in random directions, where FP32
    storage of the gradients and face geometry alone moves single faces by a
    few 1e-6 (the oracle on the stored values, tests/stored_oracle.py): every
    face within max(1e-6, 4 x its sensitivity to one-ulp changes of its
    stored inputs); every cell within 1e-6 of its gross flux."""
    print("case %s kind %d" % (name, kind), flush=True)
    b = golden_bundle(name, kind)
    assert (b["closure"] is not None) == name.startswith("m")
    g64 = run_gpu(b, kind, "fp64", device)
    gmx = run_gpu(b, kind, "mixed", device)
    ge = np.abs(gmx["grad"] - g64["grad"]).max((0, 1)) / np.maximum(
        np.abs(g64["grad"]).max((0, 1)), 1e-300)
    errs = mixed_errors(b, kind, g64, gmx)
    tot, sto, ari = SO.split_errors(kind, b, gmx["grad"], g64["dQL"], gmx["dQL"])
    sens = SO.sensitivity(kind, b, g64["grad"])
    print("mixed vs fp64 %s kind %d: grad %.2e %s; storage %.2e arithmetic %.2e; sensitivity "
          "max %.2e, error / sensitivity max %.2f"
          % (name, kind, ge.max(), errs, sto.max(), ari.max(), sens.max(),
             (tot / np.maximum(sens, 1e-300)).max()), flush=True)
    assert ge.max() <= 1e-6
    assert errs["cell_gross_max"] <= 1e-6
    assert np.all(tot <= SO.sensitivity_bound(sens))
    if b["ndim"] == 2:
        assert errs["face_max"] <= 1e-6
    assert np.array_equal(gmx["fstatus"] != 0, g64["fstatus"] != 0)


def mixed_errors(b, kind, g64, gmx):
    """Per-face normwise error of dQL (the twin copy of a periodic pair is
    compared through its partner), per-cell error relative to the gross flux,
    and relative to max |dQ|."""
    return fluxpass_errors(b, G, g64["dQL"], g64["dQ"], gmx["dQL"], gmx["dQ"],
                           once_per_pair=True)


def test_update_step_matches_oracle(device):
    import torch
    b = kh_bundle(24, 1)
    g = run_gpu(b, 1, "fp64", device)
    fp = g["fp"]
    fp.update()
    torch.cuda.synchronize()
    Qo, Wo, cst = O.update(1, 2, b["Q"], g["dQ"], b["volume"])
    assert np.array_equal(fp.m.Q_soa(), Qo)
    assert np.array_equal(fp.m.W_soa(), Wo)
    assert np.abs(fp.m.dQ_soa()).max() == 0.0
    assert int(fp.m.cstatus.max()) == 0
    # the masses the next pass's mass-flux limiter reads are the new Q.m
    assert np.array_equal(fp.m.arr["mass"].cpu().numpy(), Qo[0])
    assert not np.array_equal(Qo[0], b["mass"])


def test_kh512_full_mesh_fp64_vs_oracle(device):
    """512^2 jittered periodic KH mesh (the 2D benchmark mesh), HLLC: the
    whole FP64 pass against the oracle."""
    b = kh_bundle(512, 1, seed=20240603)
    print("faces", b["F"], flush=True)
    g = run_gpu(b, 1, "fp64", device)
    og, odQ, odQL, ost = oracle_pass(b, 1)
    assert np.array_equal(g["grad"], og)
    den = np.maximum(np.abs(odQL).max(0), face_scale(b, 1))
    err = (np.abs(g["dQL"] - odQL).max(0) / np.where(den > 0, den, 1.0)).max()
    print("kh512 fp64 face err %.3e" % err)
    assert err <= 1e-12
    assert np.abs(g["dQ"] - odQ).max() <= 1e-12 * np.abs(odQ).max()


def test_kh512_benchmark_mesh_mixed_vs_fp64(device):
    """The 2D benchmark workload as bench.py times it (512^2 jittered
    generators in Hilbert order, Kelvin-Helmholtz, HLLC): the mixed pass
    within 1e-6 of the FP64 pass on every face and cell, conservative to FP64
    round-off."""
    m = M.voronoi_periodic(M.hilbert_sorted(M.jittered_grid(512, 0.25, seed=20240603)))
    b = M.prepare(m, M.kh_ics(m), G, kind=1)
    g64 = run_gpu(b, 1, "fp64", device)
    gmx = run_gpu(b, 1, "mixed", device)
    errs = mixed_errors(b, 1, g64, gmx)
    print("kh512 mixed vs fp64: %s" % errs, flush=True)
    assert errs["face_max"] <= 1e-6 and errs["cell_gross_max"] <= 1e-6
    assert errs["conservation"] <= 1e-13


def test_sedov3d_fp64_vs_oracle(device):
    """3D periodic Poisson mesh with a Sedov blast (the 3D benchmark workload at
    20,000 cells), exact solver: the whole FP64 pass against the oracle."""
    b = sedov_bundle(20000, 0, seed=20240604)
    print("cells", b["C"], "faces", b["F"], flush=True)
    g = run_gpu(b, 0, "fp64", device)
    og, odQ, odQL, ost = oracle_pass(b, 0)
    assert np.array_equal(g["grad"], og)
    assert np.array_equal(g["fstatus"] != 0, ost != 0)
    den = np.maximum(np.abs(odQL).max(0), face_scale(b, 0))
    err = (np.abs(g["dQL"] - odQL).max(0) / np.where(den > 0, den, 1.0)).max()
    print("sedov3d fp64 face err %.3e" % err)
    assert err <= 1e-12
    assert np.abs(g["dQ"] - odQ).max() <= 1e-12 * np.abs(odQ).max()
    # conservation of the FP64 pass: periodic pairs are evaluated once per
    # side, so the totals cancel to the pair asymmetry (FP64 round-off)
    assert np.abs(g["dQ"].sum(1)).max() <= 1e-13 * np.abs(odQL).max()


def poisson2d_bundle(n, kind, seed=9):
    m = M.voronoi_periodic(M.hilbert_sorted(M.poisson_points(n, seed=seed, ndim=2)))
    return M.prepare(m, M.kh_ics(m), G, kind=kind)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("mesh", ["kh2d", "poisson2d", "sedov3d"])
def test_grouped_records_bit_identical(device, mesh, kind, prec):
    """The gradient records in the grouped layout (64 cells' k-th entries
    side by side), in cell order or over cells ordered by entry count, give
    the entry-ordered layout's gradients, face fluxes and dQ bit for bit.
    The mixed grouped pass keeps the limiter's d vectors of
    a cell's first 8 (2D) or 16 (3D) positions in LDS and reloads the rest:
    the Poisson meshes have cells with more entries than that.  Mixed
    grouped records are also read with plain loads instead of through a
    buffer descriptor (sfx_mesh.egeo_rows = 0, the records' size unknown:
    the path of records of 2 GB or more)."""
    import torch
    print("case %s kind %d %s" % (mesh, kind, prec), flush=True)
    b = {"kh2d": lambda: kh_bundle(20, kind), "poisson2d": lambda: poisson2d_bundle(600, kind),
         "sedov3d": lambda: sedov_bundle(1500, kind)}[mesh]()
    per_cell = np.diff(b["cf_off"])
    if mesh != "kh2d":
        assert per_cell.max() > (8 if b["ndim"] == 2 else 16)
    res = []
    # entry order; groups in cell order; groups over cells sorted by entry
    # count (windows of 64, 256, 512 and 1024 cells)
    for grouped, window in ((False, 0), (True, 0), (True, 64), (True, 256), (True, 512),
                            (True, 1024)):
        s = RiemannSolver(kind, b["ndim"], G, 1e-8, -5.0, prec)
        fp = FluxPass(s, M.DeviceMesh(b, prec, device, grouped=grouped, sort_window=window))
        fp.run()
        torch.cuda.synchronize()
        res.append({"grad": fp.m.grad_soa(), "dQL": fp.m.dQL_soa(), "dQ": fp.m.dQ_soa()})
        if grouped and prec == "mixed":
            fp.m.grad.zero_()
            fp.m.desc.egeo_rows = 0
            fp.run()
            torch.cuda.synchronize()
            res.append({"grad": fp.m.grad_soa(), "dQL": fp.m.dQL_soa(), "dQ": fp.m.dQ_soa()})
    for r in res[1:]:
        for k in ("grad", "dQL", "dQ"):
            assert np.array_equal(res[0][k], r[k]), k


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("mesh", ["per12", "mir10", "p3d", "m3d", "poisson2d", "sedov3d"])
def test_face_records_bit_identical(device, mesh, kind):
    """The mixed face kernel reading each face's own values from the packed
    records of sfx_mesh_face_records gives the face fluxes, status bits and
    dQ of the kernel reading the face arrays, bit for bit (periodic and
    mirror faces, 2D and 3D)."""
    import torch
    print("case %s kind %d" % (mesh, kind), flush=True)
    if mesh in ("per12", "mir10", "p3d", "m3d"):
        b = golden_bundle(mesh, kind)
    else:
        b = poisson2d_bundle(600, kind) if mesh == "poisson2d" else sedov_bundle(1500, kind)
    res = []
    for frec in (False, True):
        s = RiemannSolver(kind, b["ndim"], G, 1e-8, -5.0, "mixed")
        dm = M.DeviceMesh(b, "mixed", device, face_records=frec)
        fp = FluxPass(s, dm)
        assert (dm.desc.frec is not None) == frec
        fp.run()
        torch.cuda.synchronize()
        res.append({"dQL": fp.m.dQL_soa(), "dQ": fp.m.dQ_soa(),
                    "st": fp.m.fstatus.cpu().numpy()})
    for k in ("dQL", "dQ", "st"):
        assert np.array_equal(res[0][k], res[1][k]), k


def test_face_records_refuse_fp64_handles(device):
    """sfx_mesh_face_records builds the mixed pass's records only: an FP64
    handle gets SFX_ERR_ARG and a message, and nothing is launched."""
    import ctypes as ct
    import torch
    from shadowfax_amd import _capi as C
    b = kh_bundle(8, 1)
    s = RiemannSolver(1, 2, G, 1e-8, -5.0, "fp64")
    dm = M.DeviceMesh(b, "fp64", device)
    out = torch.zeros((3, b["F"], 4), dtype=torch.float32, device=device)
    rc = C.lib().sfx_mesh_face_records(s.handle, ct.byref(dm.desc), ct.c_void_p(out.data_ptr()),
                                       None)
    assert rc == C.SFX_ERR_ARG
    assert b"mixed" in C.lib().sfx_last_error()
    torch.cuda.synchronize()
    assert out.abs().max().item() == 0.0


@pytest.mark.parametrize("shape", [(2, 2), (2, 3), (2, 5), (3, 20), (3, 40)])
@pytest.mark.parametrize("kind", [0, 1])
def test_tiny_meshes(device, shape, kind):
    """Meshes of 4 to 40 cells (one partial group of 64, cells that meet
    several periodic images of one neighbour): the FP64 pass against the
    oracle, the mixed pass within 1e-6 of it and conservative."""
    nd, n = shape
    print("case nd %d n %d kind %d" % (nd, n, kind), flush=True)
    if nd == 2:
        b = kh_bundle(n, kind)
    else:
        m = M.voronoi_periodic(M.hilbert_sorted(M.poisson_points(n, seed=4)))
        b = M.prepare(m, M.sedov_ics(m, ncells=2), G, kind=kind)
    g = run_gpu(b, kind, "fp64", device)
    og, odQ, odQL, ost = oracle_pass(b, kind)
    assert np.array_equal(g["grad"], og), "gradients must be bit-exact"
    assert np.array_equal(g["fstatus"] != 0, ost != 0)
    den = np.maximum(np.abs(odQL).max(0), face_scale(b, kind))
    den = np.where(den > 0, den, 1.0)
    assert (np.abs(g["dQL"] - odQL).max(0) / den).max() <= 1e-12
    assert np.abs(g["dQ"] - odQ).max() <= 1e-12 * np.abs(odQ).max()
    gm = run_gpu(b, kind, "mixed", device)
    errs = mixed_errors(b, kind, g, gm)
    assert errs["face_max"] <= 1e-6 and errs["cell_gross_max"] <= 1e-6
    assert errs["conservation"] <= 1e-13


def ordered_sums(b, dQL):
    """Every cell's sum of its gather entries' rows of dQL (the right cell's
    negated) in ascending face order, added one entry at a time in FP64 (the
    order the reference's face loop reaches the cell), vectorised over cells."""
    dQL = np.asarray(dQL, np.float64)
    off = np.asarray(b["gather_off"], np.int64)
    ent = np.asarray(b["gather_entry"], np.int64)
    deg = np.diff(off)
    acc = np.zeros((dQL.shape[0], len(deg)))
    for p in range(int(deg.max()) if len(deg) else 0):
        cells = np.where(deg > p)[0]
        e = ent[off[cells] + p]
        x = dQL[:, e >> 1]
        acc[:, cells] += np.where((e & 1) == 1, -x, x)
    return acc


def test_kh2d_1024_large_mesh_kernels(device):
    """A 2D mesh of 2^20 cells (1024 x 1024 jittered KH, HLLC), where the
    gathers and the record build take their large-mesh variants (bounded
    occupancy, more rows in flight): the FP64 pass equals the oracle
    (gradients bit for bit, fluxes within 1e-12), both gathers are the
    face-ordered sums of their own fluxes bit for bit, and the mixed pass is
    within 1e-6 of the FP64 pass and conservative."""
    m = M.voronoi_periodic(M.hilbert_sorted(M.jittered_grid(1024, 0.25, seed=20240607)))
    b = M.prepare(m, M.kh_ics(m), G, kind=1)
    del m
    assert b["C"] >= 1 << 20
    print("cells", b["C"], "faces", b["F"], flush=True)
    g64 = run_gpu(b, 1, "fp64", device)
    gmx = run_gpu(b, 1, "mixed", device)
    assert np.array_equal(ordered_sums(b, g64["dQL"]), g64["dQ"])
    assert np.array_equal(ordered_sums(b, gmx["dQL"]), gmx["dQ"])
    og, odQ, odQL, ost = oracle_pass(b, 1)
    assert np.array_equal(g64["grad"], og)
    den = np.maximum(np.abs(odQL).max(0), face_scale(b, 1))
    err = (np.abs(g64["dQL"] - odQL).max(0) / np.where(den > 0, den, 1.0)).max()
    errs = mixed_errors(b, 1, g64, gmx)
    print("fp64 vs oracle %.2e; mixed vs fp64 %s" % (err, errs), flush=True)
    assert err <= 1e-12
    assert np.array_equal(g64["fstatus"] != 0, ost != 0)
    assert errs["face_max"] <= 1e-6 and errs["cell_gross_max"] <= 1e-6
    assert errs["conservation"] <= 1e-13


def test_sedov3d_full_size_mixed_errors(device):
    """The 3D benchmark workload at full size (2,097,152 Poisson generators,
    Sedov blast, exact solver), mixed pass against the FP64 pass: every cell
    within 1e-6 of its gross flux, 99% of faces within 1e-8, conservation to
    FP64 round-off.  Faces above 1e-6, and a sample of 2,000 faces, are
    recomputed by the oracle: from the FP64 pass's inputs (it must equal the
    FP64 pass) and from the mixed pass's stored inputs (tests/stored_oracle.py).
    Storage alone moves the worst faces (in the blast's hot core, where
    fluxes are driven by pressure differences a thousandth of the pressure)
    by a few 1e-6; every checked face is within max(1e-6, 4 x its
    sensitivity to one-ulp changes of its stored inputs)."""
    nthreads = min(os.cpu_count() or 1, 16)
    m = M.voronoi_periodic(M.hilbert_sorted(M.poisson_points(2097152, seed=20240604)),
                           nthreads=nthreads)
    b = M.prepare(m, M.sedov_ics(m), G, kind=0)
    del m
    print("cells", b["C"], "faces", b["F"], flush=True)
    g64 = run_gpu(b, 0, "fp64", device)
    gmx = run_gpu(b, 0, "mixed", device)
    # the gathers of meshes this large (their own kernel variants) sum each
    # cell's rows in face order, bit for bit
    assert np.array_equal(ordered_sums(b, g64["dQL"]), g64["dQ"])
    assert np.array_equal(ordered_sums(b, gmx["dQL"]), gmx["dQ"])
    errs = mixed_errors(b, 0, g64, gmx)
    den = np.maximum(np.abs(g64["dQL"]).max(0), face_scale(b, 0))
    ef = np.abs(np.asarray(gmx["dQL"], np.float64) - g64["dQL"]).max(0) / np.where(den > 0, den, 1.0)
    nbad = int((ef > 1e-6).sum())
    print("sedov3d full size mixed vs fp64: %s; faces above 1e-6: %d of %d" % (errs, nbad, b["F"]),
          flush=True)
    assert errs["cell_gross_max"] <= 1e-6
    assert errs["face_p99"] <= 1e-8
    assert errs["conservation"] <= 1e-13
    rng = np.random.default_rng(5)
    faces = np.unique(np.concatenate([np.where(ef > 1e-6)[0],
                                      rng.integers(0, b["F"], 2000)]))
    tot, sto, ari, o64 = SO.split_errors_on(0, b, faces, g64["grad"], gmx["grad"], g64["dQL"],
                                            gmx["dQL"])
    print("checked %d faces: oracle vs fp64 %.2e, storage %.2e, arithmetic %.2e"
          % (len(faces), o64.max(), sto.max(), ari.max()), flush=True)
    assert o64.max() <= 1e-12
    sens = SO.sensitivity(0, SO.face_subset(b, faces), g64["grad"])
    bad = tot > 1e-6
    print("faces above 1e-6: error %s storage %s arithmetic %s sensitivity %s"
          % (tot[bad], sto[bad], ari[bad], sens[bad]), flush=True)
    assert np.all(tot <= SO.sensitivity_bound(sens))
