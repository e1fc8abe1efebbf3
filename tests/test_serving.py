"""Serving, host side (CPU): FIFO admission by scratch footprint and the
generator of the 2,000-mesh serving workload."""
import numpy as np

from shadowfax_amd import serving as S


CELLS = np.array([100, 200, 50, 400, 10], np.int64)
FACES = 3 * CELLS


def admit(pending, arena_cells, arena_faces=10**9, max_batch=1024):
    return list(S.serve_admit(pending, CELLS, FACES, arena_cells, arena_faces, max_batch))


def test_admit_takes_the_fifo_prefix_that_fits():
    assert admit([0, 1, 2], 1000) == [0, 1, 2]
    # 100 + 200 fit in 320, 50 more would not: stop there (no overtaking)
    assert admit([0, 1, 2, 4], 320) == [0, 1]
    assert admit([3, 0], 300) == []  # the head does not fit: nothing goes
    assert admit([0, 1, 2], 1000, 1000, max_batch=2) == [0, 1]
    # face budget binds as well
    assert admit([0, 1, 2], 1000, arena_faces=3 * 250) == [0]


def test_admit_defers_a_second_step_of_the_same_mesh():
    # mesh 0 twice: the second request waits for the next batch, later
    # requests for other meshes still go
    assert admit([0, 2, 0, 4], 1000) == [0, 1, 3]
    assert admit([], 1000) == []


def test_ensemble_workload():
    w = S.ensemble_workload(n_meshes=2000)
    sizes = S.mesh_sizes(2000, w["seed"])
    assert sizes.min() >= 1e3 and sizes.max() <= 2.5e5
    # log-uniform: the mean of log10 C is the midpoint of [3, log10 2.5e5]
    assert abs(np.log10(sizes).mean() - 0.5 * (3 + np.log10(2.5e5))) < 0.05
    sides = w["sides"]
    assert sides.min() ** 2 >= 900 and sides.max() ** 2 <= 2.6e5
    # every mesh is assigned the class nearest to its size in log space
    c = sides[w["cls"]] ** 2
    assert np.all(np.abs(np.log(c) - np.log(sizes)) <= np.log(sides[1] ** 2 / sides[0] ** 2))
    assert set(np.unique(w["ic"])) == {0, 1, 2}
    assert np.all((w["shift"] >= 0) & (w["shift"] < 1))
    t, m = S.poisson_trace(20000, 1000.0, 2000, 7)
    assert np.all(np.diff(t) > 0) and m.min() >= 0 and m.max() < 2000
    assert abs(t[-1] / 20000 - 1e-3) < 5e-5  # mean gap 1/rate
