"""Serving on the GPU: an ensemble step is the flux pass + update of each
admitted mesh, bit for bit the standalone FluxPass.step of that mesh (same
per-item code), through repeated replays of the captured graph with other
work in between; the plain per-mesh path likewise; serving a trace applies
every request exactly once, in arrival order per mesh; the autotuner leaves
the states as they were."""
import numpy as np
import pytest

from shadowfax_amd import mesh as M
from shadowfax_amd import serving as S
from shadowfax_amd.fluxpass import FluxPass
from shadowfax_amd.riemann import RiemannSolver

pytestmark = pytest.mark.gpu
G = 5.0 / 3.0


def bundles(kind):
    out = []
    for k, (n, ic) in enumerate(((12, 1), (20, 0), (9, 2), (16, 1))):
        m = M.voronoi_periodic(M.hilbert_sorted(M.jittered_grid(n, 0.25, seed=40 + k)))
        W = (M.kh_ics(m) if ic == 1 else M.sod_ics(m) if ic == 0 else
             M.sedov_ics(m, gamma=G, ncells=4))
        out.append(M.prepare(m, W, G, kind=kind))
    return out


def make_ensemble(kind, prec, dev, bs, arena_cells=None):
    s = RiemannSolver(kind, 2, G, 1e-8, -5.0, prec)
    C = sum(b["C"] for b in bs)
    F = sum(b["F"] for b in bs)
    E = sum(int(b["cf_off"][-1]) for b in bs)
    ac = arena_cells or C
    af = F if arena_cells is None else int(3.5 * ac)
    ens = S.Ensemble(s, len(bs), C, F, E, ac, af)
    ids = [ens.add(M.DeviceMesh(b, prec, dev)) for b in bs]
    return ens, ids


def standalone(b, kind, prec, dev, steps):
    import torch
    fp = FluxPass(RiemannSolver(kind, 2, G, 1e-8, -5.0, prec), M.DeviceMesh(b, prec, dev))
    for _ in range(steps):
        fp.step()
    torch.cuda.synchronize()
    return fp.m.W_soa(), fp.m.Q_soa()


def check_meshes(ens, bs, kind, prec, dev, counts):
    for mid, k in enumerate(counts):
        W, Q = standalone(bs[mid], kind, prec, dev, k)
        assert np.array_equal(ens.read(mid, S.ENS_W), W), (mid, k)
        assert np.array_equal(ens.read(mid, S.ENS_Q), Q), (mid, k)
        assert np.abs(ens.read(mid, S.ENS_DQ)).max() == 0.0


@pytest.mark.parametrize("per_cu", [4, 0])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("kind", [0, 1])
def test_ensemble_steps_match_standalone(device, kind, prec, per_cu):
    print("case kind %d %s per_cu %d" % (kind, prec, per_cu), flush=True)
    bs = bundles(kind)
    ens, ids = make_ensemble(kind, prec, device, bs)
    ens.set_launch(128, per_cu)  # persistent grid / one block per tile
    ens.step([0, 2, 3])
    ens.sync()
    ens.step([3, 1])          # a second replay of the graph, another batch
    ens.step_plain(0)         # plain launches in between
    ens.sync()
    ens.step([0, 1, 2, 3])
    ens.sync()
    check_meshes(ens, bs, kind, prec, device, [3, 2, 2, 3])


@pytest.mark.parametrize("per_cu", [4, 0])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_serve_trace_applies_every_request_once(device, prec, per_cu):
    kind = 1
    bs = bundles(kind)
    # an arena that holds only some meshes at a time, so batches split
    ens, ids = make_ensemble(kind, prec, device, bs, arena_cells=500)
    ens.set_launch(64, per_cu)
    t = np.linspace(0.0, 2e-3, 24)
    mesh = np.array([0, 1, 0, 2, 3, 3, 1, 0, 2, 2, 1, 3, 0, 1, 2, 3, 0, 0, 1, 2, 3, 1, 1, 0],
                    np.int32)
    # a burst: everything queued at once, so steps take several meshes
    lat, bat, st = ens.serve(np.zeros(len(t)), mesh, S.SERVE_BATCHED)
    print("stats", st, flush=True)
    assert st["nreq"] == len(t) and np.all(lat > 0)
    assert st["max_batch"] >= 2 and st["nbatch"] < len(t)
    for mid in range(4):
        b = bat[mesh == mid]
        assert np.all(np.diff(b) > 0)  # a mesh's steps in arrival order, one per step
    lat2, bat2, st2 = ens.serve(t, mesh, S.SERVE_PLAIN)
    assert st2["nbatch"] == len(t) and st2["max_batch"] == 1
    counts = [2 * int((mesh == k).sum()) for k in range(4)]
    check_meshes(ens, bs, kind, prec, device, counts)


def test_autotune_keeps_states(device):
    kind, prec = 1, "mixed"
    bs = bundles(kind)
    ens, ids = make_ensemble(kind, prec, device, bs)
    before = [ens.read(k, S.ENS_W) for k in range(4)]
    block, per_cu, ms = ens.autotune([0, 1, 2, 3], reps=2)
    print("autotune", block, per_cu, ms, flush=True)
    assert block in (64, 128, 256) and per_cu in (0, 1, 2, 4, 8) and ms > 0
    for k in range(4):
        assert np.array_equal(ens.read(k, S.ENS_W), before[k])
    ens.step([0, 1, 2, 3])
    ens.sync()
    check_meshes(ens, bs, kind, prec, device, [1, 1, 1, 1])


@pytest.mark.parametrize("block", [64, 256])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_device_queue_applies_every_request_once(device, prec, block):
    """SFX_SERVE_QUEUE: the persistent kernel drains the posted requests,
    admits them (a small arena, so batches split; a second request for a
    mesh waits for the next batch), and steps them; every request once, each
    mesh's steps in arrival order, states bit for bit the standalone steps;
    then a trace arriving over 2 ms, and the queue mode again."""
    kind = 1
    bs = bundles(kind)
    ens, ids = make_ensemble(kind, prec, device, bs, arena_cells=500)
    ens.set_launch(block, 0)
    mesh = np.array([0, 1, 0, 2, 3, 3, 1, 0, 2, 2, 1, 3, 0, 1, 2, 3, 0, 0, 1, 2, 3, 1, 1, 0],
                    np.int32)
    print("case %s block %d: burst" % (prec, block), flush=True)
    lat, bat, st = ens.serve(np.zeros(len(mesh)), mesh, S.SERVE_QUEUE)
    print("stats", st, flush=True)
    assert st["nreq"] == len(mesh) and np.all(lat > 0)
    assert st["max_batch"] >= 2 and st["nbatch"] < len(mesh)
    for mid in range(4):
        b = bat[mesh == mid]
        assert np.all(np.diff(b) > 0)  # one step per batch, in arrival order
    print("case %s block %d: spread arrivals" % (prec, block), flush=True)
    lat2, bat2, st2 = ens.serve(np.linspace(0.0, 2e-3, len(mesh)), mesh, S.SERVE_QUEUE)
    assert st2["nreq"] == len(mesh) and np.all(lat2 > 0)
    counts = [2 * int((mesh == k).sum()) for k in range(4)]
    if prec == "fp64":
        check_meshes(ens, bs, kind, prec, device, counts)
        return
    # mixed: the persistent kernel inlines the same per-item code into one
    # function, where the compiler may fuse a multiply-add the standalone
    # kernels leave separate (the mixed code is compiled with contraction;
    # the FP64 code without): equal to rounding, not bit for bit
    for mid, k in enumerate(counts):
        W, Q = standalone(bs[mid], kind, prec, device, k)
        dW = np.abs(ens.read(mid, S.ENS_W) - W).max() / np.abs(W).max()
        dQ = np.abs(ens.read(mid, S.ENS_Q) - Q).max() / np.abs(Q).max()
        print("mesh %d after %d steps: W %.2e Q %.2e" % (mid, k, dW, dQ), flush=True)
        assert dW <= 1e-12 and dQ <= 1e-12


def workload_mesh(w, classes, i):
    """Mesh i of the serving workload as its own prepared bundle: its size
    class's tessellation with its own states, as build_ensemble copies it."""
    b = dict(classes[int(w["cls"][i])])
    pos = np.mod(b["pos"] + w["shift"][:, i:i + 1], 1.0)
    W, vel, Q, dt = S.mesh_state(b, pos, int(w["ic"][i]), float(w["icu"][i]), w["gamma"],
                                 w["kind"])
    b.update(W=W, vel=vel, Q=Q, mass=Q[0].copy(), dt=np.full(b["C"], dt), dt_value=dt)
    return b


def test_serving_workload_2000_meshes(device):
    """The serving workload at full size (the bench's): 2,000 periodic 2D
    meshes, cell counts log-uniform in [1e3, 2.5e5] (88M cells), Sod /
    Kelvin-Helmholtz / Sedov states, HLLC; the 4,000 requests of its Poisson
    trace served at once, batched mixed and plain FP64.  Each path applies
    every request exactly once: meshes without a request keep their states;
    the smallest, the largest and 12 random meshes equal that many standalone
    FluxPass steps bit for bit; on those meshes the mixed pass is within
    1e-6 of the FP64 pass per face and per cell and conserves to FP64
    round-off."""
    import torch
    from shadowfax_amd.metrics import fluxpass_errors
    w = S.ensemble_workload(2000)
    classes = S.ensemble_classes(w)
    nreq = 4000
    _, mesh = S.poisson_trace(nreq, 1.0, 2000, w["seed"] + 1)
    counts = np.bincount(mesh, minlength=2000)
    sizes = np.array([classes[int(c)]["C"] for c in w["cls"]])
    rng = np.random.default_rng(11)
    sample = sorted({int(np.argmin(sizes)), int(np.argmax(sizes))} |
                    set(rng.choice(2000, 12, replace=False).tolist()))
    untouched = np.where(counts == 0)[0][:20]
    assert len(untouched) > 0 and counts[sample].max() >= 2
    got = {}
    for prec, mode in (("mixed", S.SERVE_BATCHED), ("fp64", S.SERVE_PLAIN)):
        print("serving %s: building the ensemble" % prec, flush=True)
        solver = RiemannSolver(1, 2, G, 1e-8, -5.0, prec)
        ens, faces = S.build_ensemble(solver, device, w, classes)
        assert sum(ens.cells) > 8e7
        Q0 = {int(i): ens.read(int(i), S.ENS_Q) for i in untouched}
        lat, bat, st = ens.serve(np.zeros(nreq), mesh, mode)
        print("  served %d requests in %d steps, %.3f s" % (st["nreq"], st["nbatch"], st["wall_s"]),
              flush=True)
        assert st["nreq"] == nreq and np.all(np.isfinite(lat)) and np.all(lat > 0)
        for i in untouched:
            assert np.array_equal(ens.read(int(i), S.ENS_Q), Q0[int(i)]), i
        got[prec] = {i: (ens.read(i, S.ENS_W), ens.read(i, S.ENS_Q)) for i in sample}
        ens.close()
        del ens
        torch.cuda.empty_cache()
    for i in sample:
        b = workload_mesh(w, classes, i)
        print("mesh %d: %d cells, %d requests" % (i, b["C"], counts[i]), flush=True)
        for prec in ("mixed", "fp64"):
            W, Q = standalone(b, 1, prec, device, int(counts[i]))
            assert np.array_equal(got[prec][i][0], W), (i, prec)
            assert np.array_equal(got[prec][i][1], Q), (i, prec)
        runs = {}
        for prec in ("fp64", "mixed"):
            fp = FluxPass(RiemannSolver(1, 2, G, 1e-8, -5.0, prec), M.DeviceMesh(b, prec, device))
            fp.run()
            torch.cuda.synchronize()
            runs[prec] = (fp.m.dQL_soa(), fp.m.dQ_soa())
        e = fluxpass_errors(b, G, runs["fp64"][0], runs["fp64"][1], runs["mixed"][0],
                            runs["mixed"][1], once_per_pair=True)
        assert e["face_max"] <= 1e-6 and e["cell_gross_max"] <= 1e-6, (i, e)
        assert e["conservation"] <= 1e-13, (i, e)
