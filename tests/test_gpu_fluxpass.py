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
from shadowfax_amd import mesh as M
from shadowfax_amd.fluxpass import FluxPass
from shadowfax_amd.riemann import RiemannSolver

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
G = 5.0 / 3.0
KEYS = ("pos", "volume", "centroid", "left", "right", "area", "midpoint", "shift", "ghost",
        "twin", "cf_off", "cf_face", "W", "vel")


def golden_bundle(name, kind):
    z = np.load(os.path.join(HERE, "golden", "mesh_2d.npz"))
    m = {k: z[name + "_" + k] for k in KEYS}
    m.update(ndim=2, C=m["volume"].shape[0], F=m["area"].shape[0], box=np.ones(2))
    return M.prepare(m, m["W"], G, kind=kind, vel=m["vel"])


def kh_bundle(n, kind, seed=9):
    m = M.voronoi_periodic(M.jittered_grid(n, 0.25, seed=seed))
    return M.prepare(m, M.kh_ics(m), G, kind=kind)


def run_gpu(b, kind, precision, dev):
    import torch
    s = RiemannSolver(kind, b["ndim"], G, 1e-8, -5.0, precision)
    fp = FluxPass(s, M.DeviceMesh(b, precision, dev))
    fp.run()
    torch.cuda.synchronize()
    out = {k: getattr(fp.m, k).double().cpu().numpy() for k in ("grad", "dQL", "dQ")}
    out["fstatus"] = fp.m.fstatus.cpu().numpy()
    out["fp"] = fp
    return out


def oracle_pass(b, kind):
    ob = dict(b)
    ob["grad"] = O.gradients(kind, b)
    dQ, dQL, st, _ = O.face_flux_pass(kind, ob)
    return ob["grad"], dQ, dQL, st


def face_scale(b, kind):
    """dt A rho_bar a_bar max(1, a_bar) per face: the flux floor of the
    normwise metric, carried to dQL units."""
    W, L, R = b["W"], b["left"], b["right"]
    Rr = np.where(R >= 0, R, L)
    a = np.sqrt(G * W[-1] / W[0])
    rho = 0.5 * (W[0, L] + W[0, Rr])
    ab = 0.5 * (a[L] + a[Rr])
    return b["dt"][L] * b["area"] * rho * ab * np.maximum(ab, 1.0)


@pytest.mark.parametrize("name", ["per12", "mir10"])
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


@pytest.mark.parametrize("kind", [0, 1])
def test_gather_is_ordered_sum_of_gpu_face_fluxes(device, kind):
    """dQ on the GPU equals, bit for bit, the sum of the GPU's own dQL in the
    reference's face-loop order (left += dQL, right += -dQL)."""
    b = kh_bundle(32, kind)
    g = run_gpu(b, kind, "fp64", device)
    dQL = g["dQL"]
    acc = np.zeros_like(g["dQ"])
    for f in range(b["F"]):
        acc[:, b["left"][f]] += dQL[:, f]
        if b["right"][f] >= 0 and not b["ghost"][f]:
            acc[:, b["right"][f]] += -dQL[:, f]
    assert np.array_equal(acc, g["dQ"])


@pytest.mark.parametrize("kind", [0, 1])
def test_mixed_pass_within_1e6_and_conservative(device, kind):
    b = kh_bundle(48, kind)
    g64 = run_gpu(b, kind, "fp64", device)
    gmx = run_gpu(b, kind, "mixed", device)
    errs = mixed_errors(b, kind, g64, gmx)
    print("mixed vs fp64 kind %d: %s" % (kind, errs))
    assert errs["face_max"] <= 1e-6
    assert errs["cell_gross_max"] <= 1e-6
    scale = np.abs(g64["dQ"]).max()
    # one evaluation per periodic pair: totals cancel to FP64 round-off
    tot = np.abs(gmx["dQ"].sum(1)).max() / scale
    print("mixed conservation residual %.3e" % tot)
    assert tot <= 1e-13


def mixed_errors(b, kind, g64, gmx):
    """Per-face normwise error of dQL (the twin copy of a periodic pair is
    compared through its partner), per-cell error relative to the gross flux,
    and relative to max |dQ|."""
    dQL64, dQLmx = g64["dQL"], gmx["dQL"].copy()
    tw = b["twin"]
    skip = (b["ghost"] == 1) & (tw >= 0) & (tw < np.arange(b["F"]))
    dQLmx[:, skip] = -dQLmx[:, tw[skip]]
    den = np.maximum(np.abs(dQL64).max(0), face_scale(b, kind))
    den = np.where(den > 0, den, 1.0)
    ef = np.abs(dQLmx - dQL64).max(0) / den
    gross = np.zeros(b["C"])
    np.add.at(gross, b["left"], np.abs(dQL64).max(0))
    real = (b["right"] >= 0) & (b["ghost"] == 0)
    np.add.at(gross, b["right"][real], np.abs(dQL64[:, real]).max(0))
    ec = np.abs(gmx["dQ"] - g64["dQ"]).max(0) / gross
    return {"face_max": float(ef.max()), "face_p99": float(np.quantile(ef, 0.99)),
            "cell_gross_max": float(ec.max()), "cell_gross_p99": float(np.quantile(ec, 0.99)),
            "cell_vs_maxdQ": float(np.abs(gmx["dQ"] - g64["dQ"]).max() / np.abs(g64["dQ"]).max())}


def test_update_step_matches_oracle(device):
    import torch
    b = kh_bundle(24, 1)
    g = run_gpu(b, 1, "fp64", device)
    fp = g["fp"]
    fp.update()
    torch.cuda.synchronize()
    Qo, Wo, cst = O.update(1, 2, b["Q"], g["dQ"], b["volume"])
    assert np.array_equal(fp.m.Q.cpu().numpy(), Qo)
    assert np.array_equal(fp.m.arr["W"].cpu().numpy(), Wo)
    assert np.abs(fp.m.dQ.cpu().numpy()).max() == 0.0
    assert int(fp.m.cstatus.max()) == 0


def test_config2_full_mesh_fp64_vs_oracle(device):
    """512^2 jittered periodic KH mesh (the config-2 workload), HLLC: the
    whole FP64 pass against the oracle."""
    b = kh_bundle(512, 1, seed=20240603)
    print("faces", b["F"], flush=True)
    g = run_gpu(b, 1, "fp64", device)
    og, odQ, odQL, ost = oracle_pass(b, 1)
    assert np.array_equal(g["grad"], og)
    den = np.maximum(np.abs(odQL).max(0), face_scale(b, 1))
    err = (np.abs(g["dQL"] - odQL).max(0) / np.where(den > 0, den, 1.0)).max()
    print("config2 fp64 face err %.3e" % err)
    assert err <= 1e-12
    assert np.abs(g["dQ"] - odQ).max() <= 1e-12 * np.abs(odQ).max()
