"""Seeded synthetic inputs for the batched Riemann workloads.

Counter-based SplitMix64: value(stream, i) depends only on (seed, stream, i),
so any slice of a batch can be regenerated independently and the CPU and GPU
legs see the same numbers.  By default every generated value is rounded to
float32 and stored as float64, so the FP64 and the mixed (FP32-storage)
paths read identical inputs and differences measure arithmetic only;
round32=False keeps the FP64 draws (then the mixed path's FP32 storage
rounds them, and the difference includes that).

Problem families (ndim = 3 unless stated):
  uniform   rho, p = 10**U(-3, 3) per side; velocity components U(-2, 2);
            n uniform on the sphere (z = U(-1, 1), phi = U(0, 2 pi)); v = 0.
  vacuum    uniform, then one side (chosen by a coin flip) replaced by
            rho = 1e-8, p = 1e-10.
  shock     uniform, then p_R = p_L * 1e6**(+-1) (sign by a coin flip).
  mixed     per problem: vacuum with probability 0.05, shock with 0.05,
            otherwise uniform.
"""
import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)

# stream ids
S_RHOL, S_RHOR, S_PL, S_PR, S_VL, S_VR, S_NZ, S_NPHI, S_CAT, S_SIDE, \
    S_SHOCK = range(11)


def splitmix64(seed, stream, idx):
    """idx: uint64 array. Returns uint64 array."""
    with np.errstate(over="ignore"):
        base = np.uint64(seed) * GOLDEN + np.uint64(stream + 1) * M1
        z = base + (idx.astype(np.uint64) + np.uint64(1)) * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, stream, idx):
    """U[0, 1) with 53 random bits."""
    return (splitmix64(seed, stream, idx) >> np.uint64(11)).astype(
        np.float64) * (1.0 / 9007199254740992.0)


def f32(x):
    return np.asarray(x, np.float64).astype(np.float32).astype(np.float64)


def riemann_batch(N, seed=1, family="uniform", ndim=3, start=0, round32=True):
    """Returns dict with WL, WR (ndim+2, N), n (ndim, N), float64 holding
    float32-representable values (round32=False: the draws unrounded), plus
    'category' (0 uniform, 1 vacuum, 2 shock)."""
    idx = np.arange(start, start + N, dtype=np.uint64)
    NC = ndim + 2
    WL = np.empty((NC, N))
    WR = np.empty((NC, N))
    WL[0] = 10.0 ** (6.0 * uniform(seed, S_RHOL, idx) - 3.0)
    WR[0] = 10.0 ** (6.0 * uniform(seed, S_RHOR, idx) - 3.0)
    WL[NC - 1] = 10.0 ** (6.0 * uniform(seed, S_PL, idx) - 3.0)
    WR[NC - 1] = 10.0 ** (6.0 * uniform(seed, S_PR, idx) - 3.0)
    for d in range(ndim):
        # sub-streams per component: offset the counter by d*2^40
        off = np.uint64(d) << np.uint64(40)
        WL[1 + d] = 4.0 * uniform(seed, S_VL, idx + off) - 2.0
        WR[1 + d] = 4.0 * uniform(seed, S_VR, idx + off) - 2.0
    if ndim == 3:
        z = 2.0 * uniform(seed, S_NZ, idx) - 1.0
        phi = 2.0 * np.pi * uniform(seed, S_NPHI, idx)
        s = np.sqrt(np.maximum(0.0, 1.0 - z * z))
        n = np.stack([s * np.cos(phi), s * np.sin(phi), z])
    else:
        phi = 2.0 * np.pi * uniform(seed, S_NPHI, idx)
        n = np.stack([np.cos(phi), np.sin(phi)])
    cat = np.zeros(N, np.uint8)
    if family == "vacuum":
        cat[:] = 1
    elif family == "shock":
        cat[:] = 2
    elif family == "mixed":
        u = uniform(seed, S_CAT, idx)
        cat[u < 0.05] = 1
        cat[(u >= 0.05) & (u < 0.10)] = 2
    elif family != "uniform":
        raise ValueError(family)
    side = uniform(seed, S_SIDE, idx) < 0.5
    m = cat == 1
    ml = m & side
    mr = m & ~side
    WL[0, ml] = 1e-8
    WL[NC - 1, ml] = 1e-10
    WR[0, mr] = 1e-8
    WR[NC - 1, mr] = 1e-10
    m = cat == 2
    up = uniform(seed, S_SHOCK, idx) < 0.5
    WR[NC - 1, m & up] = WL[NC - 1, m & up] * 1e6
    WR[NC - 1, m & ~up] = WL[NC - 1, m & ~up] * 1e-6
    if not round32:
        return {"WL": WL, "WR": WR, "n": n, "category": cat}
    return {"WL": f32(WL), "WR": f32(WR), "n": f32(n), "category": cat}


# The batches the solver benchmark and tests are quoted on: 262,144 problems
# with states log-uniform over six decades, and 4,194,304 with 5% near-vacuum
# and 5% 1e6-shock problems mixed in.
RIEMANN_262K = dict(N=262144, family="uniform", ndim=3, seed=20240601,
               gamma=5.0 / 3.0, tol=1e-8, cutoff=-5.0)
RIEMANN_4M = dict(N=4194304, family="mixed", ndim=3, seed=20240602,
               gamma=5.0 / 3.0, tol=1e-8, cutoff=-5.0)
