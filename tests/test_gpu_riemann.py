"""GPU parity of the batched Riemann solvers (through the C ABI).

  * FP64 plain path vs the golden vectors of the compiled reference and vs
    the oracle: normwise relative error <= 1e-12 (the device pow() rounds
    differently from glibc, so bitwise equality is reported, not required);
    status bits must agree exactly.
  * mixed path (FP32 storage and bulk arithmetic, FP64 final Newton step):
    normwise relative error <= 1e-6 against the FP64 results.
  * full-size (4,194,304 problems, the benchmark configuration) properties:
    mixed vs FP64 error, uniform state gives the Euler flux, solver counters.
  * edge fixtures (exact-zero densities, negative pressures, NaNs): the status
    bits of the reference's exceptions and the positions of NaN fluxes must
    agree exactly, the finite fluxes to the tolerance of the path.
  * ExactRiemannSolver::solve (batched and scalar drop-ins): half states,
    Mach numbers and status against the golden vectors, and the reflective
    write-back of WR.
"""
import os

import numpy as np
import pytest

import oracle as O
from shadowfax_amd import workloads as WK
from shadowfax_amd.metrics import normwise_error, summarize
from shadowfax_amd.riemann import RiemannSolver, ExactRiemannSolver, HLLCRiemannSolver

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FP64_TOL = 1e-12
MIXED_TOL = 1e-6
G = 5.0 / 3.0


def _t(a, dtype, dev):
    import torch
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device=dev)


def gpu_flux(kind, ndim, WL, WR, n, v, refl, dev, precision="fp64", f32=False, gamma=G):
    import torch
    s = RiemannSolver(kind, ndim, gamma, 1e-8, -5.0, precision)
    dt = torch.float32 if f32 else torch.float64
    N = WL.shape[1]
    st = torch.zeros(N, dtype=torch.uint8, device=dev)
    F = s.solve_for_flux_batch(_t(WL, dt, dev), _t(WR, dt, dev), _t(n, dt, dev),
                               None if v is None else _t(v, dt, dev),
                               None if refl is None else _t(refl, torch.uint8, dev),
                               status=st)
    torch.cuda.synchronize()
    return F.double().cpu().numpy(), st.cpu().numpy(), s


FAMILIES = ["uniform", "mixed", "vacuum", "shock", "zero", "negp", "nan"]


def finite_err(F, R, WL, WR, ok):
    """NaN fluxes must sit where the reference's are; the normwise error of
    the others (summarized)."""
    nanF, nanR = np.isnan(F[:, ok]).any(0), np.isnan(R[:, ok]).any(0)
    assert np.array_equal(nanF, nanR), "NaN fluxes differ from the reference's"
    fin = np.where(ok)[0][~nanR]
    s = summarize(normwise_error(F[:, fin], R[:, fin], WL[:, fin], WR[:, fin], G))
    s["nan"] = int(nanR.sum())
    return s


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("fam", FAMILIES)
@pytest.mark.parametrize("variant", ["plain", "moving"])
@pytest.mark.parametrize("kind", [0, 1])
def test_fp64_matches_reference_golden(device, ndim, fam, variant, kind):
    print("case", ndim, fam, variant, kind, flush=True)
    g = np.load(os.path.join(HERE, "golden", "riemann_%dd.npz" % ndim))
    key = "%s_%s_" % (fam, variant)
    name = "exact" if kind == 0 else "hllc"
    F, st, _ = gpu_flux(kind, ndim, g[key + "WL"], g[key + "WR"], g[key + "n"], g[key + "v"],
                        g[key + "refl"], device)
    ref_st = g[key + "st_" + name]
    assert np.array_equal(st != 0, ref_st != 0)
    ok = ref_st == 0
    R = g[key + "F_" + name]
    s = finite_err(F, R, g[key + "WL"], g[key + "WR"], ok)
    bitwise = np.mean(np.all((F[:, ok] == R[:, ok]) | np.isnan(R[:, ok]), axis=0))
    print("fp64 %dd %s %s %s: err %s bitwise %.4f" % (ndim, fam, variant, name, s, bitwise))
    assert s["max"] <= FP64_TOL


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("fam", FAMILIES)
def test_exact_solve_batch_matches_reference_golden(device, ndim, fam):
    """sfx_exact_solve_batch (ExactRiemannSolver::solve over a batch): half
    states at x/t = 0, the Mach number of the strongest shock (from 0) and
    the status, against the compiled reference, plain and moving (25%
    reflective) variants."""
    import torch
    g = np.load(os.path.join(HERE, "golden", "riemann_%dd.npz" % ndim))
    for variant in ("plain", "moving"):
        key = "%s_%s_" % (fam, variant)
        print("case", ndim, key, flush=True)
        WL, WR, n, refl = (g[key + k] for k in ("WL", "WR", "n", "refl"))
        N = WL.shape[1]
        s = RiemannSolver(0, ndim, G, 1e-8, -5.0, "fp64")
        mach = torch.zeros(N, dtype=torch.float64, device=device)
        st = torch.zeros(N, dtype=torch.uint8, device=device)
        f64 = torch.float64
        Wh = s.solve_batch(_t(WL, f64, device), _t(WR, f64, device), _t(n, f64, device),
                           _t(refl, torch.uint8, device), mach=mach, status=st)
        torch.cuda.synchronize()
        Wh, mach, st = Wh.cpu().numpy(), mach.cpu().numpy(), st.cpu().numpy()
        ref_st = g[key + "st_solve"]
        assert np.array_equal(st != 0, ref_st != 0)
        ok = ref_st == 0
        R = g[key + "Whalf"]
        scale = np.maximum(np.abs(WL).max(0), np.abs(np.where(refl, WL, WR)).max(0))
        scale = np.where(np.isfinite(scale) & (scale > 0), scale, 1.0)
        assert np.array_equal(np.isnan(Wh[:, ok]), np.isnan(R[:, ok]))
        err = np.nan_to_num(np.abs(Wh - R)).max(0)[ok] / scale[ok]
        me = np.abs(mach - g[key + "mach"])[ok] / np.maximum(1.0, np.abs(g[key + "mach"][ok]))
        print("  half state err %.3e  mach err %.3e" % (err.max(), me.max()), flush=True)
        assert err.max() <= FP64_TOL
        assert me.max() <= FP64_TOL


@pytest.mark.parametrize("fam", FAMILIES)
def test_scalar_solve_matches_reference_golden(device, fam):
    """The scalar drop-in sfx_solve (RiemannSolver::solve, host pointers) on
    48 problems of each family, a third of them reflective: half state,
    Mach number, the RiemannSolverException (raised as in the reference)
    and the write-back of WR = WL on reflective faces."""
    from shadowfax_amd.riemann import RiemannSolverException
    g = np.load(os.path.join(HERE, "golden", "riemann_3d.npz"))
    key = "%s_moving_" % fam
    s = ExactRiemannSolver(G, ndim=3)
    refl_all = g[key + "refl"]
    idx = np.concatenate([np.where(refl_all == 1)[0][:16], np.where(refl_all == 0)[0][:32]])
    for i in idx:
        WL = g[key + "WL"][:, i].copy()
        WR = g[key + "WR"][:, i].copy()
        refl = bool(refl_all[i])
        ref_st = g[key + "st_solve"][i]
        try:
            W, mach = s.solve(WL, WR, g[key + "n"][:, i], 0.0, reflective=refl)
            raised = False
        except RiemannSolverException:
            raised = True
        assert raised == (ref_st != 0), i
        if refl:
            assert np.array_equal(WR, WL, equal_nan=True), i
        else:
            assert np.array_equal(WR, g[key + "WR"][:, i], equal_nan=True), i
        if not raised:
            R = g[key + "Whalf"][:, i]
            scale = max(np.abs(WL).max(), np.abs(WR).max(), 1e-300)
            assert np.abs(W - R).max() <= FP64_TOL * scale, (i, W, R)
            assert abs(mach - g[key + "mach"][i]) <= FP64_TOL * max(1.0, abs(mach)), i


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("fam", FAMILIES)
@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("f32", [True, False])
def test_mixed_within_1e6_of_fp64(device, ndim, fam, kind, f32):
    print("case", ndim, fam, kind, f32, flush=True)
    g = np.load(os.path.join(HERE, "golden", "riemann_%dd.npz" % ndim))
    key = "%s_moving_" % fam
    name = "exact" if kind == 0 else "hllc"
    F, st, _ = gpu_flux(kind, ndim, g[key + "WL"], g[key + "WR"], g[key + "n"], g[key + "v"],
                        g[key + "refl"], device, precision="mixed", f32=f32)
    ref_st = g[key + "st_" + name]
    if fam in ("negp", "nan"):
        # Every problem of these families has a NaN or a negative pressure
        # among its inputs.  The mixed path fails the same problems and puts
        # NaN fluxes where the reference does (finite_err asserts it); the
        # finite fluxes the reference returns for the rest come out of
        # comparisons with NaN sound speeds or pressures (a NaN compares
        # false, so some branch is taken), which the mixed path, ordering its
        # decisions differently, is not held to.  The FP64 path is held to
        # them (test_fp64_matches_reference_golden).
        assert np.array_equal(st != 0, ref_st != 0)
        s = finite_err(F, g[key + "F_" + name], g[key + "WL"], g[key + "WR"],
                       (ref_st == 0) & (st == 0))
        print("mixed %dd %s %s f32=%s (invalid inputs): %s" % (ndim, fam, name, f32, s))
        return
    ok = (ref_st == 0) & (st == 0)
    assert ok.mean() > 0.99
    s = finite_err(F, g[key + "F_" + name], g[key + "WL"], g[key + "WR"], ok)
    print("mixed %dd %s %s f32=%s: err %s" % (ndim, fam, name, f32, s))
    assert s["max"] <= MIXED_TOL


@pytest.mark.parametrize("kind", [0, 1])
def test_gamma_14(device, kind):
    g = np.load(os.path.join(HERE, "golden", "riemann_3d.npz"))
    key = "g14_mixed_moving_"
    name = "exact" if kind == 0 else "hllc"
    for prec in ("fp64", "mixed"):
        F, st, _ = gpu_flux(kind, 3, g[key + "WL"], g[key + "WR"], g[key + "n"], g[key + "v"],
                            g[key + "refl"], device, precision=prec, gamma=1.4)
        ok = (g[key + "st_" + name] == 0) & (st == 0)
        err = summarize(normwise_error(F[:, ok], g[key + "F_" + name][:, ok],
                                       g[key + "WL"][:, ok], g[key + "WR"][:, ok], 1.4))
        print("gamma 1.4", name, prec, err)
        assert err["max"] <= (FP64_TOL if prec == "fp64" else MIXED_TOL)


def test_fp64_vs_oracle_and_counters(device):
    """Counters: every problem counted; Brent/vacuum counts as in the oracle."""
    b = WK.riemann_batch(20000, seed=3, family="mixed", ndim=3)
    F, st, s = gpu_flux(0, 3, b["WL"], b["WR"], b["n"], None, None, device)
    Fo, sto, c = O.solve_for_flux(O.EXACT, 3, b["WL"], b["WR"], b["n"])
    ok = sto == 0
    err = summarize(normwise_error(F[:, ok], Fo[:, ok], b["WL"][:, ok], b["WR"][:, ok], G))
    assert err["max"] <= FP64_TOL
    stats = s.stats()
    assert stats["neval"] == 20000
    assert stats["nvacuum"] == c.nvacuum
    # At convergence f(p) is within an ulp of zero, so a device pow() that
    # rounds differently from glibc can flip `fpguess > 0` and send the last
    # step through Brent; both land on the same root (checked above).
    assert abs(stats["nbrent"] - c.nbrent) <= max(5, c.nbrent // 100)


def test_scalar_entry_points(device):
    """solve / solve_for_flux drop-ins (host pointers) and the reflective
    write-back of WR."""
    s = ExactRiemannSolver(5.0 / 3.0, ndim=3)
    WL = np.array([1., 0., 0., 0., 1.])
    WR = np.array([0.125, 0., 0., 0., 0.1])
    W, mach = s.solve(WL, WR, [1., 0., 0.])
    assert abs(W[0] - 0.47969) < 1e-4 and abs(W[1] - 0.841194) < 1e-4
    assert abs(W[4] - 0.293945) < 1e-4
    F = s.solve_for_flux(WL, WR, [1., 0., 0.], [0.1, 0., 0.])
    Fo, _, _ = O.solve_for_flux(O.EXACT, 3, WL[:, None], WR[:, None], np.array([[1.], [0.], [0.]]),
                                v=np.array([[0.1], [0.], [0.]]))
    assert np.allclose(F, Fo[:, 0], rtol=1e-13, atol=1e-15)
    WR2 = WR.copy()
    s.solve_for_flux(WL, WR2, [1., 0., 0.], None, reflective=True)
    assert np.array_equal(WR2, WL)
    h = HLLCRiemannSolver(5.0 / 3.0, ndim=2)
    F2 = h.solve_for_flux([1., 0., 0., 1.], [0.125, 0., 0., 0.1], [1., 0.])
    Fo2, _, _ = O.solve_for_flux(O.HLLC, 2, np.array([[1.], [0.], [0.], [1.]]),
                                 np.array([[0.125], [0.], [0.], [0.1]]), np.array([[1.], [0.]]))
    assert np.allclose(F2, Fo2[:, 0], rtol=1e-13, atol=1e-15)
    assert s.tag() == "EXAC" and h.tag() == "HLLC"


@pytest.mark.parametrize("kind", [0, 1])
def test_full_size_262k(device, kind):
    """All 262,144 problems of the CPU-runnable batch (rho, p
    log-uniform over six decades, velocities in [-2, 2], normals uniform on
    the sphere): the FP64 path against the oracle (the reference's algorithm,
    bit-exact with the compiled reference on the golden vectors) to 1e-12,
    the mixed path against FP64 to 1e-6, and the vacuum count."""
    cfg = WK.RIEMANN_262K
    b = WK.riemann_batch(cfg["N"], seed=cfg["seed"], family=cfg["family"], ndim=3)
    print("262k kind", kind, flush=True)
    F64, st64, s = gpu_flux(kind, 3, b["WL"], b["WR"], b["n"], None, None, device)
    Fo, sto, c = O.solve_for_flux(kind, 3, b["WL"], b["WR"], b["n"])
    ok = (st64 == 0) & (sto == 0)
    assert ok.mean() > 0.999 and np.array_equal(st64 != 0, sto != 0)
    err = summarize(normwise_error(F64[:, ok], Fo[:, ok], b["WL"][:, ok], b["WR"][:, ok], G))
    print("262k fp64 vs oracle", err, flush=True)
    assert err["max"] <= FP64_TOL
    assert s.stats()["nvacuum"] == c.nvacuum
    F32, st32, _ = gpu_flux(kind, 3, b["WL"], b["WR"], b["n"], None, None, device,
                            precision="mixed", f32=True)
    ok2 = ok & (st32 == 0)
    e2 = summarize(normwise_error(F32[:, ok2], F64[:, ok2], b["WL"][:, ok2], b["WR"][:, ok2], G))
    print("262k mixed vs fp64", e2, flush=True)
    assert e2["max"] <= MIXED_TOL


@pytest.mark.parametrize("kind", [0, 1])
def test_full_size_4m(device, kind):
    """4,194,304 problems of the benchmark configuration: mixed (FP32
    storage) vs FP64 on the GPU, plus a 20,000-problem oracle spot check of
    the FP64 path."""
    import torch
    cfg = WK.RIEMANN_4M
    b = WK.riemann_batch(cfg["N"], seed=cfg["seed"], family=cfg["family"], ndim=3)
    print("generated", flush=True)
    F64, st64, _ = gpu_flux(kind, 3, b["WL"], b["WR"], b["n"], None, None, device)
    F32, st32, _ = gpu_flux(kind, 3, b["WL"], b["WR"], b["n"], None, None, device,
                            precision="mixed", f32=True)
    ok = (st64 == 0) & (st32 == 0)
    assert ok.mean() > 0.999
    err = summarize(normwise_error(F32[:, ok], F64[:, ok], b["WL"][:, ok], b["WR"][:, ok], G))
    print("4m kind", kind, "mixed vs fp64", err)
    assert err["max"] <= MIXED_TOL
    idx = np.arange(0, cfg["N"], cfg["N"] // 20000)[:20000]
    sub = {k: b[k][:, idx] for k in ("WL", "WR", "n")}
    Fo, sto, _ = O.solve_for_flux(kind, 3, sub["WL"], sub["WR"], sub["n"])
    okk = (sto == 0) & ok[idx]
    e2 = summarize(normwise_error(F64[:, idx][:, okk], Fo[:, okk], sub["WL"][:, okk],
                                  sub["WR"][:, okk], G))
    assert e2["max"] <= FP64_TOL
    del F64, F32
    torch.cuda.empty_cache()


@pytest.mark.parametrize("kind", [0, 1])
def test_uniform_state_gives_euler_flux_gpu(device, kind):
    b = WK.riemann_batch(100000, seed=11, ndim=3)
    WL = b["WL"]
    n = b["n"]
    for prec, f32, tol in (("fp64", False, 1e-12), ("mixed", True, 1e-6)):
        F, st, _ = gpu_flux(kind, 3, WL, WL, n, None, None, device, precision=prec, f32=f32)
        rho, p, vel = WL[0], WL[4], WL[1:4]
        un = (vel * n).sum(0)
        E = p / (G - 1) + 0.5 * rho * (vel * vel).sum(0)
        R = np.empty_like(WL)
        R[0] = rho * un
        R[1:4] = rho * vel * un + p * n
        R[4] = (E + p) * un
        err = summarize(normwise_error(F, R, WL, WL, G))
        print(kind, prec, err)
        assert err["max"] <= tol


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("kind", [0, 1])
def test_cell_operations_match_reference_golden(device, ndim, kind):
    """get_time_derivative, get_Q and get_W (with the pressure-floor and
    vacuum edge cases in the fixtures) on the GPU against the compiled
    reference, bit for bit (no transcendental on these paths)."""
    import torch
    g = np.load(os.path.join(HERE, "golden", "riemann_%dd.npz" % ndim))
    name = ("exact", "hllc")[kind]
    s = RiemannSolver(kind, ndim, G, 1e-8, -5.0, "fp64")
    f64 = torch.float64
    d = s.get_time_derivative_batch(_t(g["dwdt_W"], f64, device), _t(g["dwdt_grad"], f64, device))
    Q = s.get_Q_batch(_t(g["q_vol"], f64, device), _t(g["q_W"], f64, device))
    Qin = _t(g["w_Qin_" + name], f64, device)
    W = s.get_W_batch(_t(g["q_vol"], f64, device), Qin)
    torch.cuda.synchronize()
    assert np.array_equal(d.cpu().numpy(), g["dwdt_" + name])
    assert np.array_equal(Q.cpu().numpy(), g["q_Q_" + name])
    assert np.array_equal(W.cpu().numpy(), g["w_W_" + name])
    assert np.array_equal(Qin.cpu().numpy(), g["w_Qout_" + name])


@pytest.mark.parametrize("f32", [True, False])
def test_compacted_exact_equals_one_kernel(device, f32):
    """The two-phase (wave-compacted) mixed exact solve returns bit for bit
    what the one-kernel path returns, with the same counters, for split
    points that queue many, few or no problems."""
    import torch
    b = WK.riemann_batch(200000, seed=5, family="mixed", ndim=3)
    dt = torch.float32 if f32 else torch.float64
    WL, WR, n = (_t(b[k], dt, device) for k in ("WL", "WR", "n"))
    refl = _t((np.arange(200000) % 7 == 0).astype(np.uint8), torch.uint8, device)
    out = {}
    for split in (0, 1, 4, 30):
        s = RiemannSolver(0, 3, G, 1e-8, -5.0, "mixed")
        s.set_split(split)
        st = torch.zeros(200000, dtype=torch.uint8, device=device)
        F = s.solve_for_flux_batch(WL, WR, n, None, refl, status=st)
        torch.cuda.synchronize()
        out[split] = (F.cpu().numpy(), st.cpu().numpy(), s.stats())
    F0, st0, c0 = out[0]
    for split in (1, 4, 30):
        F1, st1, c1 = out[split]
        assert np.array_equal(F1, F0), split
        assert np.array_equal(st1, st0), split
        assert c1 == c0, split


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_empty_and_ragged_batches(device, kind, precision):
    """N = 0 launches nothing and succeeds; batch sizes that leave a partial
    wavefront or block (1, 63, 65, 257, 65,537) give the oracle's fluxes to
    the tolerance of their path, and the partial tail writes nothing past N."""
    import torch
    s = RiemannSolver(kind, 3, G, 1e-8, -5.0, precision)
    dt = torch.float32 if precision == "mixed" else torch.float64
    z = torch.zeros((5, 0), dtype=dt, device=device)
    F = s.solve_for_flux_batch(z, z.clone(), torch.zeros((3, 0), dtype=dt, device=device))
    torch.cuda.synchronize()
    assert tuple(F.shape) == (5, 0)
    b = WK.riemann_batch(65537, seed=11, family="mixed", ndim=3)
    for N in (1, 63, 65, 257, 65537):
        WL, WR, n = b["WL"][:, :N], b["WR"][:, :N], b["n"][:, :N]
        # outputs are the heads of larger buffers whose tails hold a
        # sentinel: the kernels write only the first 5 N fluxes, N flags
        Fbuf = torch.full((5 * N + 64,), 7.25, dtype=dt, device=device)
        sbuf = torch.full((N + 64,), 0xAB, dtype=torch.uint8, device=device)
        F = Fbuf[:5 * N].view(5, N)
        st = sbuf[:N]
        st.zero_()
        s.solve_for_flux_batch(_t(WL, dt, device), _t(WR, dt, device), _t(n, dt, device),
                               F=F, status=st)
        torch.cuda.synchronize()
        assert torch.all(Fbuf[5 * N:] == 7.25) and torch.all(sbuf[N:] == 0xAB)
        Fo, sto, _ = O.solve_for_flux(O.EXACT if kind == 0 else O.HLLC, 3, WL, WR, n)
        ok = (sto == 0) & (st.cpu().numpy() == 0)
        err = summarize(normwise_error(F.double().cpu().numpy()[:, ok], Fo[:, ok], WL[:, ok],
                                       WR[:, ok], G))
        assert err["max"] <= (FP64_TOL if precision == "fp64" else MIXED_TOL), (N, err)
