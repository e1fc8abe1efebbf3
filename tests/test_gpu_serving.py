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
