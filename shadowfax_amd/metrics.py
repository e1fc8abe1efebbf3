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


# ---------------------------------------------------------- flux pass --

def face_flux_scale(b, gamma):
    """dt A rho_bar a_bar max(1, a_bar) per face: the flux floor of the
    normwise metric carried to dQL = dt A F units (b: a prepared mesh)."""
    W, L, R = b["W"], b["left"], b["right"]
    Rr = np.where(R >= 0, R, L)
    a = np.sqrt(gamma * W[-1] / W[0])
    rho = 0.5 * (W[0, L] + W[0, Rr])
    ab = 0.5 * (a[L] + a[Rr])
    return b["dt"][L] * b["area"] * rho * ab * np.maximum(ab, 1.0)


def fluxpass_errors(b, gamma, dQL_ref, dQ_ref, dQL, dQ, once_per_pair=False):
    """Errors of one flux pass (dQL [nc][F], dQ [nc][C]) against a reference
    pass on the same mesh:
      face_*: normwise per face, floor face_flux_scale;
      cell_gross_*: per cell, |dQ - dQ_ref|_inf over the cell's gross flux
        sum_f |dQL_ref,f|_inf (dQ is a sum of face fluxes that largely
        cancel, so |dQ| itself is no scale);
      cell_vs_maxdQ: max |dQ - dQ_ref| / max |dQ_ref|;
      conservation: max_k |sum_i dQ_ki| / max |dQL_ref|.
    once_per_pair: the pass evaluated each periodic face pair once; the
    second copy's dQL is taken as its twin's, negated (what the mixed pass
    writes there)."""
    dQL_ref, dQ_ref, dQL, dQ = map(_np, (dQL_ref, dQ_ref, dQL, dQ))
    F = dQL.shape[1]
    if once_per_pair:
        dQL = dQL.copy()
        tw = b["twin"]
        skip = (b["ghost"] == 1) & (tw >= 0) & (tw < np.arange(F))
        dQL[:, skip] = -dQL[:, tw[skip]]
    den = np.maximum(np.abs(dQL_ref).max(0), face_flux_scale(b, gamma))
    den = np.where(den > 0, den, 1.0)
    ef = np.abs(dQL - dQL_ref).max(0) / den
    gross = np.zeros(dQ.shape[1])
    np.add.at(gross, b["left"], np.abs(dQL_ref).max(0))
    real = (b["right"] >= 0) & (b["ghost"] == 0)
    np.add.at(gross, b["right"][real], np.abs(dQL_ref[:, real]).max(0))
    ec = np.abs(dQ - dQ_ref).max(0) / np.where(gross > 0, gross, 1.0)
    scale = np.abs(dQL_ref).max()
    return {"face_max": float(ef.max()), "face_p99": float(np.quantile(ef, 0.99)),
            "cell_gross_max": float(ec.max()), "cell_gross_p99": float(np.quantile(ec, 0.99)),
            "cell_vs_maxdQ": float(np.abs(dQ - dQ_ref).max() / np.abs(dQ_ref).max()),
            "conservation": float(np.abs(dQ.sum(1)).max() / scale)}
