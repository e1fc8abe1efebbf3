"""The device time step (sfx_hydro_timestep, TimeLine::calculate_timestep for
gas cells) against the oracle, bit for bit: new start/end times, dt,
active flags, status, system step and active count, in 2D and 3D, for both
solvers' sound speeds (the exact one guards rho = 0), in individual and
global mode.  (The 3D characteristic length goes through cbrt, which the
device and glibc may round differently in the last bit; that moves a
step only when the CFL time lands on a power-of-two boundary to within an
ulp, which random inputs do not.)"""
import numpy as np
import pytest

import oracle as O
from shadowfax_amd.riemann import RiemannSolver

pytestmark = pytest.mark.gpu
G = 5.0 / 3.0


def case(rng, C, nd, kind):
    W = np.empty((C, nd + 2))
    W[:, 0] = 10 ** rng.uniform(-2, 2, C)
    W[:, 1:1 + nd] = rng.uniform(-2, 2, (C, nd))
    W[:, -1] = 10 ** rng.uniform(-3, 1, C)
    if kind == 0:
        W[::31, 0] = 0.0  # get_soundspeed's rho = 0 guard
    vel = W[:, 1:1 + nd] + rng.uniform(-0.5, 0.5, (C, nd))
    V = 10 ** rng.uniform(-7, -3, C)
    return W, vel, V


@pytest.mark.parametrize("glob", [0, 1])
@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("nd", [2, 3])
def test_timestep_matches_oracle(device, nd, kind, glob):
    import torch
    print("case nd %d kind %d global %d" % (nd, kind, glob), flush=True)
    rng = np.random.default_rng(11 + nd + 3 * kind + 5 * glob)
    C = 100003
    W, vel, V = case(rng, C, nd, kind)
    M = 1 << 60
    d = dict(integer_maxtime=M, current_time=M >> 4, timestep=M >> 10, max_timestep=M >> 3,
             min_timestep=1 << 10, tottime=2e-3, mintime=0.125, cfl=0.4,
             global_timestep=glob)
    now = d["current_time"] + d["timestep"]
    end = np.where(rng.uniform(size=C) < 0.6, now, now + (M >> 9)).astype(np.uint64)
    if glob:
        end[:] = now
    start = end - np.uint64(M >> 12)
    s0, e0, dt0, a0, ss0, na0, st0 = O.hydro_timestep(kind, nd, W, vel, V, O.timeline(**d),
                                                      start, end, gamma=G)
    s = RiemannSolver(kind, nd, G, 1e-8, -5.0, "fp64")
    t = lambda a, dt: torch.tensor(a, dtype=dt, device=device)
    st_d, en_d = t(start.view(np.int64), torch.int64), t(end.view(np.int64), torch.int64)
    dt, act, status, res = s.hydro_timestep(t(W, torch.float64), t(vel, torch.float64),
                                            t(V, torch.float64), d, st_d, en_d)
    torch.cuda.synchronize()
    assert np.array_equal(st_d.cpu().numpy().view(np.uint64), s0)
    assert np.array_equal(en_d.cpu().numpy().view(np.uint64), e0)
    assert np.array_equal(dt.cpu().numpy(), dt0)
    assert np.array_equal(act.cpu().numpy(), a0)
    assert np.array_equal(status.cpu().numpy(), st0)
    r = res.cpu().numpy().view(np.uint64)
    assert r[0] == ss0 and r[1] == na0


def test_timestep_rejects_bad_timeline(device):
    import torch
    from shadowfax_amd import _capi as C
    s = RiemannSolver(0, 2, G, 1e-8, -5.0, "fp64")
    z = lambda n, dt: torch.zeros(n, dtype=dt, device=device)
    bad = dict(integer_maxtime=0, current_time=0, timestep=0, max_timestep=1, min_timestep=0,
               tottime=1.0, mintime=0.0, cfl=0.4, global_timestep=0)
    with pytest.raises(C.SfxError):
        s.hydro_timestep(z((4, 4), torch.float64), z((4, 2), torch.float64),
                         z(4, torch.float64), bad, z(4, torch.int64), z(4, torch.int64))
