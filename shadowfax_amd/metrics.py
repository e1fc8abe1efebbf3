"""Error metrics for comparing flux batches (numpy, float64).

Per-component relative error is ill-posed where a flux component crosses
zero (e.g. the mass flux at a contact at rest), so errors are normwise per
problem:

    err_i = max_k |F_ik - R_ik| / max(max_k |R_ik|, rho_bar a_bar max(1, a_bar))

with rho_bar, a_bar the mean density and sound speed of the two input states;
the floor keeps problems with a near-zero reference flux from dividing by
round-off.  Arrays are (ncomp, N).
"""
import numpy as np


def _np(a):
    if isinstance(a, np.ndarray):
        return a.astype(np.float64, copy=False)
    return a.detach().double().cpu().numpy()


def flux_scale(WL, WR, gamma):
    WL, WR = _np(WL), _np(WR)
    P = WL.shape[0] - 1
    with np.errstate(divide="ignore", invalid="ignore"):
        aL = np.where(WL[0] > 0, np.sqrt(gamma * np.abs(WL[P]) / WL[0]), 0.0)
        aR = np.where(WR[0] > 0, np.sqrt(gamma * np.abs(WR[P]) / WR[0]), 0.0)
    rho = 0.5 * (WL[0] + WR[0])
    a = 0.5 * (aL + aR)
    return rho * a * np.maximum(a, 1.0)


def normwise_error(F, R, WL, WR, gamma):
    F, R = _np(F), _np(R)
    num = np.abs(F - R).max(0)
    den = np.maximum(np.abs(R).max(0), flux_scale(WL, WR, gamma))
    den = np.where(den > 0, den, 1.0)
    return num / den


def summarize(err):
    """max, p99, p50 and count of a 1-D error array."""
    err = _np(err)
    err = err[np.isfinite(err)]
    if err.size == 0:
        return {"max": 0.0, "p99": 0.0, "p50": 0.0, "n": 0}
    q = np.quantile(err, [0.5, 0.99])
    return {"max": float(err.max()), "p99": float(q[1]), "p50": float(q[0]),
            "n": int(err.size)}
