"""Benchmark of the face-flux path on one MI355X.

Headline workload: one full face-flux pass over the 3D benchmark mesh --
2,097,152 Poisson generators in the 3D periodic unit box (seed 20240604) in
Hilbert order, Voronoi faces from the host tessellator (16.5M faces), Sedov
blast (rho 1, p 1e-5, v 0, E = 1 in the 32 cells nearest the centre), exact
Riemann solver.  One step = gradients (VorCell::estimate_gradient) + face
fluxes (VorFace::set_v + calculate_flux) + the ordered per-cell gather (the
scatter of VorTessManager::hydro's face loop), in the mixed-precision path
(FP32 bulk, FP64 final Newton step for p*, FP64 conserved accumulation),
with the mesh, states and work buffers resident in HBM.  The plain all-FP64
pass is timed the same way; their ratio and the mixed-vs-FP64 errors are
reported with the line.

Also measured (extra objects of the same JSON line): the 2D pass (512 x 512
jittered periodic Kelvin-Helmholtz mesh, HLLC), the batched Riemann solves
(4,194,304 problems with 5% near-vacuum and 5% 1e6 shocks; 262,144
problems), and the 2,000-mesh serving run (p50/p99 request latency).

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--no-cpu]
                    [--no-riemann] [--no-2d] [--no-serve]

roofline: the pass's dominant kernel (the face-flux kernel), its counted
FP operations (profiles/flop_counts.json, the reference algorithm on these
exact inputs) and its algorithmic bytes per launch over its mean launch time
(HIP events on the stream it runs on), against the FP64/FP32 vector peak
(78.6 TF; non-packed FP32 issues at the FP64 rate) and 8 TB/s; bound = the
nearer roof.  The measured HBM traffic comes from profiles/hbm_traffic.json.
cpu_baseline: the oracle's FP64 restatement of the same pass over the same
mesh on the host's cores (threads = the job's CPU share), and one thread on
a contiguous sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# VALU issue: 256 CUs x 4 SIMDs, one wave64 instruction per SIMD per 4 cycles
# (the FP64/FP32 FMA rate), 2.4 GHz peak engine clock
VALU_PEAK_WAVE_INSTS = 256 * 4 * 2.4e9 / 4
FP32_PEAK_TF = 157.3
FP64_PEAK_TF = 78.6
METRIC = "face-fluxes/sec at ≤1e-6 rel err vs FP64 plain GPU; p50/p99 served-step latency"
SEED_3D = 20240604
SEED_2D = 20240603


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_variant(fn, steps, warmup, stream):
    """warmup untimed calls, then steps calls each bracketed by HIP events on
    `stream` (the stream the kernels are launched on), synchronised before and
    after."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    t0 = time.perf_counter()
    for i in range(steps):
        starts[i].record(stream)
        fn()
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    return {"ms_mean": float(np.mean(per)), "ms_min": float(np.min(per)),
            "ms_p50": float(np.median(per)), "wall_s": wall}


def host_threads():
    """The job's CPU share: OMP_NUM_THREADS when the environment sets it (16
    per GPU on the MI355X boxes), else every core in the affinity mask."""
    aff = len(os.sched_getaffinity(0))
    try:
        n = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        n = 0
    return max(1, min(aff, n) if n > 0 else aff), aff


def mesh3d():
    from shadowfax_amd import mesh as M
    t0 = time.perf_counter()
    m = M.voronoi_periodic(M.hilbert_sorted(M.poisson_points(2097152, seed=SEED_3D)),
                           nthreads=host_threads()[0])
    t_mesh = time.perf_counter() - t0
    b = M.prepare(m, M.sedov_ics(m), 5.0 / 3.0, kind=0)
    return b, t_mesh


def mesh2d(n=512):
    from shadowfax_amd import mesh as M
    t0 = time.perf_counter()
    # generators in Hilbert order, as the reference keeps its particles
    m = M.voronoi_periodic(M.hilbert_sorted(M.jittered_grid(n, 0.25, seed=SEED_2D)),
                           nthreads=host_threads()[0])
    t_mesh = time.perf_counter() - t0
    b = M.prepare(m, M.kh_ics(m), 5.0 / 3.0, kind=1)
    return b, t_mesh


def alg_bytes(b, precision, part="pass"):
    """Algorithmic bytes: every unique input read once, every output written
    once, at the variant's storage width.  Per face: the two cell indices (2 x
    4), the area, the midpoint and the periodic shift (1 + 2 ndim values);
    per cell: W (ndim + 2, FP64 in both variants), the gradients (ndim (ndim
    + 2)), centroid, generator position and velocity (3 ndim), mass, total
    area and dt (3, FP64); written: dQ (ndim + 2, FP64) for the pass, dQL per
    face (ndim + 2 values) for the face kernel alone, which reads no
    centroid or position beyond the face's own geometry.  FP64 storage:
    48 F + 200 C (2D) and 64 F + 296 C (3D)."""
    nd, C, F = b["ndim"], b["C"], b["F"]
    NC = nd + 2
    t = 8 if precision == "fp64" else 4
    face = 8 + t * (1 + 2 * nd)
    cell_in = 8 * NC + t * nd * NC + t * 3 * nd + 8 * 3
    if part == "pass":
        return face * F + (cell_in + 8 * NC) * C
    # the face kernel: face geometry in, the two cells' states, gradients,
    # velocities and scalars in (no centroid/position: those enter through
    # the face geometry), dQL out
    cell_face = 8 * NC + t * nd * NC + t * nd + 8 * 3
    return (face + t * NC) * F + cell_face * C


def load_profile(name):
    try:
        with open(os.path.join(ROOT, "profiles", name)) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def measured_traffic(kernel, items=None):
    """HBM bytes per launch of a kernel from the committed hardware-counter
    summary (profiles/hbm_traffic.json), if it was measured on this many
    items (the counters record the launch grid: items rounded up to whole
    blocks of 256)."""
    d = load_profile("hbm_traffic.json")
    k = d["kernels"].get(kernel) if d else None
    if not k or (items is not None and (int(k.get("grid", -1)) + 255) // 256 != (items + 255) // 256):
        return None
    return float(k["traffic_bytes"])


def bench_fluxpass(b, kind, steps, warmup, dev):
    """One full face-flux pass per step (gradients + face fluxes + ordered
    gather) in the FP64 and the mixed path: pass and per-kernel times,
    mixed-vs-FP64 errors, conservation, the per-tessellation record build."""
    import torch
    from shadowfax_amd import mesh as M
    from shadowfax_amd.fluxpass import FluxPass
    from shadowfax_amd.metrics import fluxpass_errors
    from shadowfax_amd.riemann import RiemannSolver
    C, F, nd = b["C"], b["F"], b["ndim"]
    stream = torch.cuda.current_stream()
    out = {"C": C, "F": F, "solver": "hllc" if kind else "exact"}
    res = {}
    for prec in ("fp64", "mixed"):
        s = RiemannSolver(kind, nd, 5.0 / 3.0, 1e-8, -5.0, prec)
        s.enable_stats(False)
        dm = M.DeviceMesh(b, prec, dev)
        # the per-tessellation record build (face records, then
        # sfx_mesh_entry_geometry): a first build loads the kernels' code,
        # the second is timed (a simulation rebuilds them every step)
        dm.build_entries(s.handle)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        torch.cuda.synchronize()
        ev[0].record(stream)
        dm.build_entries(s.handle)
        ev[1].record(stream)
        torch.cuda.synchronize()
        ms_build = ev[0].elapsed_time(ev[1])
        fp = FluxPass(s, dm)
        # the three kernels timed one at a time first (each with its own
        # warm-up), then the pass: the GPU has left its idle clocks by then
        # (timed after its warm-up alone, a pass's first steps here ran up
        # to 10% slower than the rest)
        parts = {"gradients": fp.gradients,
                 "faces": lambda fp=fp: fp.face_fluxes(status=False),
                 "gather": fp.gather}
        ms_parts = {}
        for name, fn in parts.items():
            ms_parts["ms_" + name] = time_variant(fn, steps, 1, stream)["ms_mean"]
        r = time_variant(lambda fp=fp: fp.run(status=False), steps, warmup, stream)
        r.update(ms_parts)
        r["ms_build_records"] = ms_build
        r["face_fluxes_per_s"] = F / (r["ms_mean"] * 1e-3)
        r["fresh_mesh_ms"] = r["ms_mean"] + ms_build
        fp.run()
        torch.cuda.synchronize()
        res[prec] = (fp.m.dQL_soa(), fp.m.dQ_soa(), int((fp.m.fstatus != 0).sum().item()))
        r["failed_faces"] = res[prec][2]
        r["alg_bytes_pass"] = alg_bytes(b, prec)
        r["alg_bytes_faces"] = alg_bytes(b, prec, "faces")
        out[prec] = r
        del fp, dm
        torch.cuda.empty_cache()
    out["mixed_speedup"] = out["fp64"]["ms_mean"] / out["mixed"]["ms_mean"]
    out["mixed_vs_fp64"] = fluxpass_errors(b, 5.0 / 3.0, res["fp64"][0], res["fp64"][1],
                                           res["mixed"][0], res["mixed"][1], once_per_pair=True)
    ef = face_errors(b, res["fp64"][0], res["mixed"][0])
    out["mixed_vs_fp64"]["faces_above_1e-6"] = int((ef > 1e-6).sum())
    out["fp64_conservation"] = float(np.abs(res["fp64"][1].sum(1)).max() /
                                     np.abs(res["fp64"][0]).max())
    return out


def face_errors(b, dQL64, dQLmx):
    from shadowfax_amd.metrics import face_flux_scale
    d = np.asarray(dQLmx, np.float64).copy()
    tw = b["twin"]
    sk = (b["ghost"] == 1) & (tw >= 0) & (tw < np.arange(b["F"]))
    d[:, sk] = -d[:, tw[sk]]
    den = np.maximum(np.abs(dQL64).max(0), face_flux_scale(b, 5.0 / 3.0))
    return np.abs(d - dQL64).max(0) / np.where(den > 0, den, 1.0)


def roofline(b, fp, workload_key, face_kernel, pass_kernels):
    """The face kernel's roofline (counted operations and algorithmic bytes
    over its mean launch time) and the pass's (max(T_hbm, T_valu) / t)."""
    counts = load_profile("flop_counts.json")
    ms_face = fp["mixed"]["ms_faces"]
    t = ms_face * 1e-3
    byt = fp["mixed"]["alg_bytes_faces"]
    gbs = byt / t / 1e9
    out = {}
    cnt = counts.get(workload_key) if counts else None
    if cnt and cnt["faces"] == b["F"]:
        ops = cnt["faces_kernel"]["ops"]
        tf = ops / t / 1e12
        f_hbm, f_valu = gbs / HBM_PEAK_GBS, tf / FP64_PEAK_TF
        if f_valu >= f_hbm:
            out = {"bound": "valu", "achieved": tf, "peak": FP64_PEAK_TF, "unit": "TFLOP/s",
                   "frac": f_valu}
        else:
            out = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                   "frac": f_hbm}
        out["flops_per_launch"] = ops
        out["flops_per_face"] = cnt["faces_kernel"]["per_face"]
        out["transcendentals_per_face"] = cnt["faces_kernel"]["transcendentals_per_face"]
        out["hbm"] = {"achieved_GBs": gbs, "frac": f_hbm}
        out["valu"] = {"achieved_TFs": tf, "frac": f_valu,
                       "peak_note": "FP64 vector peak; non-packed FP32 issues at the same rate"}
        # the whole pass: max(T_hbm, T_valu) / t, T_hbm its algorithmic bytes at the
        # HBM peak, T_valu its counted operations at the vector peak
        pass_ops = cnt["pass_ops"]
        for prec in ("fp64", "mixed"):
            tp = fp[prec]["ms_mean"] * 1e-3
            T_hbm = fp[prec]["alg_bytes_pass"] / (HBM_PEAK_GBS * 1e9)
            T_valu = pass_ops / (FP64_PEAK_TF * 1e12)
            out["pass_" + prec] = {"T_hbm_ms": T_hbm * 1e3, "T_valu_ms": T_valu * 1e3,
                                   "t_ms": tp * 1e3, "frac": max(T_hbm, T_valu) / tp,
                                   "alg_bytes": fp[prec]["alg_bytes_pass"], "ops": pass_ops}
    else:
        out = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "frac": gbs / HBM_PEAK_GBS, "note": "no FP operation count for this mesh"}
    traffic = measured_traffic(face_kernel, b["F"])
    out["traffic"] = traffic
    out["traffic_source"] = ("profiles/hbm_traffic.json (rocprofv3 FETCH_SIZE + WRITE_SIZE per "
                             "dispatch, scaled by a copy kernel of known size)") if traffic else None
    out["kernel"] = face_kernel
    out["kernel_ms"] = ms_face
    out["alg_bytes_per_launch"] = byt
    out["pass_kernels"] = pass_kernels
    return out


def bench_riemann(steps, warmup, dev, N=None):
    """Batched solve_for_flux over the 4,194,304-problem batch (5% near-vacuum,
    5% 1e6 shocks): exact and HLLC, FP64 and mixed, errors and counters; the
    mixed error also for unrounded FP64 inputs (the batch is otherwise drawn
    FP32-representable, so both paths read identical inputs)."""
    import torch
    from shadowfax_amd import workloads as WK
    from shadowfax_amd.metrics import normwise_error, summarize
    from shadowfax_amd.riemann import RiemannSolver
    cfg = dict(WK.RIEMANN_4M)
    if N:
        cfg["N"] = N
    N = cfg["N"]
    g = cfg["gamma"]
    b = WK.riemann_batch(N, seed=cfg["seed"], family=cfg["family"], ndim=3)
    d64 = {k: torch.tensor(b[k], dtype=torch.float64, device=dev) for k in ("WL", "WR", "n")}
    d32 = {k: torch.tensor(b[k], dtype=torch.float32, device=dev) for k in ("WL", "WR", "n")}
    stream = torch.cuda.current_stream()
    variants, outs = {}, {}
    for kname, kind, prec in (("exact", 0, "fp64"), ("exact", 0, "mixed"),
                              ("hllc", 1, "fp64"), ("hllc", 1, "mixed")):
        s = RiemannSolver(kind, 3, g, cfg["tol"], cfg["cutoff"], prec)
        s.enable_stats(False)
        src = d64 if prec == "fp64" else d32
        F = torch.empty_like(src["WL"])
        fn = lambda s=s, src=src, F=F: s.solve_for_flux_batch(src["WL"], src["WR"], src["n"],
                                                              F=F, stream=stream)
        log("  timing", kname, prec)
        r = time_variant(fn, steps, warmup, stream)
        r["face_fluxes_per_s"] = N / (r["ms_mean"] * 1e-3)
        s.enable_stats(True)
        s.reset_stats()
        st = torch.zeros(N, dtype=torch.uint8, device=dev)
        s.solve_for_flux_batch(src["WL"], src["WR"], src["n"], F=F, status=st, stream=stream)
        torch.cuda.synchronize()
        r["stats"] = s.stats()
        outs[(kname, prec)] = (F.double().cpu().numpy(), st.cpu().numpy())
        variants["%s_%s" % (kname, prec)] = r
    errors = {}
    for kname in ("exact", "hllc"):
        F64, st64 = outs[(kname, "fp64")]
        F32, st32 = outs[(kname, "mixed")]
        ok = (st64 == 0) & (st32 == 0)
        errors[kname + "_mixed_vs_fp64"] = summarize(
            normwise_error(F32[:, ok], F64[:, ok], b["WL"][:, ok], b["WR"][:, ok], g))
        variants[kname + "_mixed"]["speedup_vs_fp64"] = (
            variants[kname + "_fp64"]["ms_mean"] / variants[kname + "_mixed"]["ms_mean"])
    # unrounded inputs: the same draws kept in FP64; the FP64 path reads them
    # as they are, the mixed path reads their FP32 rounding (its storage)
    ub = WK.riemann_batch(N, seed=cfg["seed"], family=cfg["family"], ndim=3, round32=False)
    for kname, kind in (("exact", 0), ("hllc", 1)):
        Fs = {}
        for prec, dt in (("fp64", torch.float64), ("mixed", torch.float32)):
            s = RiemannSolver(kind, 3, g, cfg["tol"], cfg["cutoff"], prec)
            st = torch.zeros(N, dtype=torch.uint8, device=dev)
            F = s.solve_for_flux_batch(*(torch.tensor(ub[k], dtype=dt, device=dev)
                                         for k in ("WL", "WR", "n")), status=st)
            torch.cuda.synchronize()
            Fs[prec] = (F.double().cpu().numpy(), st.cpu().numpy())
        ok = (Fs["fp64"][1] == 0) & (Fs["mixed"][1] == 0)
        errors[kname + "_mixed_vs_fp64_unrounded_inputs"] = summarize(
            normwise_error(Fs["mixed"][0][:, ok], Fs["fp64"][0][:, ok], ub["WL"][:, ok],
                           ub["WR"][:, ok], g))
        del Fs
    del d64, d32
    torch.cuda.empty_cache()
    return {"workload": "batched Riemann solve_for_flux, %d 3D problems, gamma 5/3, 5%% "
                        "near-vacuum + 5%% 1e6 shocks (seed %d)" % (N, cfg["seed"]),
            "N": N, "variants": variants, "errors": errors}


def bench_serving(dev, n_meshes=2000, nreq=4000, load=0.7):
    """Serving: an ensemble of n_meshes 2D periodic meshes (cell counts
    log-uniform in [1e3, 2.5e5], Sod / Kelvin-Helmholtz / Sedov states, HLLC)
    resident in HBM; requests (one flux pass + update of one mesh) arrive as
    a Poisson process and are served open loop.  Batched mixed serving (the
    captured graph with persistent kernels, admitted by scratch footprint)
    against the plain per-mesh FP64 launch; each at load x its own capacity
    (capacity: the same trace with every request present at t = 0), and the
    batched server also at the plain one's offered rate."""
    import torch
    from shadowfax_amd import serving as S
    from shadowfax_amd.riemann import RiemannSolver
    w = S.ensemble_workload(n_meshes)
    out = {"workload": "%d periodic 2D meshes, cells log-uniform in [1e3, 2.5e5] (%d "
                       "jittered-grid size classes, each mesh a translated copy with its own "
                       "states), Sod/KH/Sedov, HLLC; Poisson arrivals, uniform mesh choice"
                       % (n_meshes, len(np.unique(w["cls"]))),
           "meshes": n_meshes, "requests": nreq, "load": load}
    t0 = time.perf_counter()
    classes = S.ensemble_classes(w)
    out["tessellation_s"] = time.perf_counter() - t0
    plain_rate = None
    for prec, mode, name in (("fp64", S.SERVE_PLAIN, "plain_fp64"),
                             ("mixed", S.SERVE_BATCHED, "batched_mixed")):
        t0 = time.perf_counter()
        solver = RiemannSolver(1, 2, 5.0 / 3.0, 1e-8, -5.0, prec)
        solver.enable_stats(False)
        ens, faces = S.build_ensemble(solver, dev, w, classes, log=log)
        r = {"build_s": time.perf_counter() - t0}
        tot_c = sum(ens.cells)
        r["cells"], r["faces"] = int(tot_c), int(faces.sum())
        if mode == S.SERVE_BATCHED:
            # autotune the launch on a batch of about a million cells
            ids, c = [], 0
            for k in range(n_meshes):
                if c + ens.cells[k] > min(ens.arena_cells, 1_000_000):
                    break
                ids.append(k)
                c += ens.cells[k]
            blk, pcu, ms = ens.autotune(ids, reps=3)
            r["launch"] = {"block": blk, "blocks_per_cu": pcu, "autotune_batch_cells": c,
                           "autotune_ms": ms}
        arr, mesh = S.poisson_trace(nreq, 1.0, n_meshes, w["seed"] + 1)
        lat, bat, st = ens.serve(np.zeros(nreq), mesh, mode)
        cap_req = nreq / st["wall_s"]
        r["capacity"] = {"requests_per_s": cap_req,
                         "face_fluxes_per_s": float(faces[mesh].sum()) / st["wall_s"],
                         "mean_batch": st["mean_batch"], "steps": st["nbatch"]}
        rate = load * cap_req
        lat, bat, st = ens.serve(arr / rate, mesh, mode)
        r["open_loop"] = {"rate_per_s": rate, "p50_ms": st["p50_s"] * 1e3,
                          "p99_ms": st["p99_s"] * 1e3, "mean_ms": st["mean_s"] * 1e3,
                          "max_ms": st["max_s"] * 1e3, "mean_batch": st["mean_batch"],
                          "steps": st["nbatch"],
                          "face_fluxes_per_s": float(faces[mesh].sum()) / st["wall_s"]}
        if mode == S.SERVE_PLAIN:
            plain_rate = rate
        elif plain_rate:
            # the same requests at the plain path's offered rate: the two
            # servers' latencies at equal load
            lat, bat, st = ens.serve(arr / plain_rate, mesh, mode)
            r["open_loop_at_plain_rate"] = {
                "rate_per_s": plain_rate, "p50_ms": st["p50_s"] * 1e3,
                "p99_ms": st["p99_s"] * 1e3, "mean_ms": st["mean_s"] * 1e3,
                "max_ms": st["max_s"] * 1e3, "mean_batch": st["mean_batch"],
                "steps": st["nbatch"]}
        out[name] = r
        log("  %s: capacity %.0f req/s, p50 %.3f ms p99 %.3f ms" %
            (name, cap_req, r["open_loop"]["p50_ms"], r["open_loop"]["p99_ms"]))
        if mode == S.SERVE_BATCHED:
            # the same ensemble served from the request queue on the device
            # (one persistent kernel: admission, planning, the four phases)
            q = {}
            lat, bat, st = ens.serve(np.zeros(nreq), mesh, S.SERVE_QUEUE)
            cap_q = nreq / st["wall_s"]
            q["capacity"] = {"requests_per_s": cap_q,
                             "face_fluxes_per_s": float(faces[mesh].sum()) / st["wall_s"],
                             "mean_batch": st["mean_batch"], "steps": st["nbatch"]}
            rate = load * cap_q
            lat, bat, st = ens.serve(arr / rate, mesh, S.SERVE_QUEUE)
            q["open_loop"] = {"rate_per_s": rate, "p50_ms": st["p50_s"] * 1e3,
                              "p99_ms": st["p99_s"] * 1e3, "mean_ms": st["mean_s"] * 1e3,
                              "max_ms": st["max_s"] * 1e3, "mean_batch": st["mean_batch"],
                              "steps": st["nbatch"],
                              "face_fluxes_per_s": float(faces[mesh].sum()) / st["wall_s"]}
            out["queue_mixed"] = q
            log("  queue_mixed: capacity %.0f req/s, p50 %.3f ms p99 %.3f ms" %
                (cap_q, q["open_loop"]["p50_ms"], q["open_loop"]["p99_ms"]))
        ens.close()
        del ens
        torch.cuda.empty_cache()
    return out


def copy_probe(mib, dev):
    """A plain elementwise copy of known size (reads and writes mib MiB per
    launch): under a hardware-counter run its FETCH_SIZE / WRITE_SIZE per
    dispatch give the counters' scale."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device=dev)
    y = torch.empty_like(x)
    for _ in range(5):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    log("copy probe: %d MiB read + %d MiB written per launch" % (mib, mib))


def measured_peaks(dev, gib=2, reps=5):
    """The GPU's measured ceilings (libsfxprobe.so): a float4 copy of gib GiB
    (read + written bytes over the launch), and FP64 / FP32 FMA loops (2
    flops an FMA, the packed FP32 forms off as in every kernel here); the
    best of reps launches each, timed with HIP events on the launch stream.
    The copy runs one block of 256 a CU with four non-temporal loads in
    flight a lane: of 1-4-8 loads a lane at 1-32 blocks a CU, the fastest
    (5.6-5.8 TB/s; 4.4-4.8 TB/s at 8 or more blocks a CU, torch's copy_
    5.1)."""
    import ctypes as ct
    import torch
    lib = ct.CDLL(os.path.join(ROOT, "shadowfax_amd", "libsfxprobe.so"))
    lib.sfxp_copy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int64, ct.c_int, ct.c_void_p]
    lib.sfxp_fma.argtypes = [ct.c_int, ct.c_void_p, ct.c_int, ct.c_int, ct.c_void_p]
    st = torch.cuda.current_stream()
    sp = ct.c_void_p(st.cuda_stream)
    props = torch.cuda.get_device_properties(dev)
    blocks = props.multi_processor_count * 8

    def best(launch):
        t = []
        for _ in range(reps + 1):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(st)
            rc = launch()
            e1.record(st)
            torch.cuda.synchronize()
            if rc != 0:
                raise RuntimeError("probe launch failed (%d)" % rc)
            t.append(e0.elapsed_time(e1) * 1e-3)
        return min(t[1:])

    n4 = gib * (1 << 30) // 16
    x = torch.ones(n4 * 4, dtype=torch.float32, device=dev)
    y = torch.empty_like(x)
    tc = best(lambda: lib.sfxp_copy(ct.c_void_p(x.data_ptr()), ct.c_void_p(y.data_ptr()), n4,
                                    props.multi_processor_count, sp))
    assert bool(torch.equal(x[-4:], y[-4:]))
    del x, y
    iters = 4096
    out = torch.empty(blocks * 256, dtype=torch.float64, device=dev)
    flops = blocks * 256 * 16 * iters
    t64 = best(lambda: lib.sfxp_fma(1, ct.c_void_p(out.data_ptr()), blocks, iters, sp))
    t32 = best(lambda: lib.sfxp_fma(0, ct.c_void_p(out.data_ptr()), blocks, iters, sp))
    del out
    torch.cuda.empty_cache()
    return {"copy_GBs": 2 * n4 * 16 / tc / 1e9, "fma_f64_TFs": flops / t64 / 1e12,
            "fma_f32_TFs": flops / t32 / 1e12,
            "how": "float4 copy of %d GiB (read + written bytes, one block of 256 a CU), FMA "
                   "loops of %d x 8 chains on %d blocks of 256; best of %d launches"
                   % (gib, iters, blocks, reps)}


def measured_fracs(roof, mp):
    """A roofline's achieved rates also as fractions of the measured
    ceilings (the published peaks stay the reference)."""
    if not mp:
        return
    if "hbm" in roof:
        roof["hbm"]["frac_of_measured"] = roof["hbm"]["achieved_GBs"] / mp["copy_GBs"]
    if "valu" in roof:
        roof["valu"]["frac_of_measured"] = roof["valu"]["achieved_TFs"] / mp["fma_f64_TFs"]


def cpu_baseline(b, kind, runs=5):
    """The oracle's FP64 restatement of the same pass (gradients, face
    fluxes, ordered gather) over the same mesh: on the job's CPU share of
    threads over the whole mesh (contiguous cell and face ranges per thread,
    then the per-cell gather in face order), and on one thread over a
    contiguous sample (the first 1/16 of the cells and faces: faces are in
    creation order, so about the same cells).  Each: one warm-up run, then
    the median of `runs` timed runs."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.lib()
    T, aff = host_threads()
    C, F = b["C"], b["F"]
    cut = lambda n, k: (0, max(1, n // k))

    def median_run(threads, **kw):
        O.threaded_pass(kind, b, threads, **kw)
        walls, phases = [], []
        for _ in range(runs):
            times = {}
            t0 = time.perf_counter()
            O.threaded_pass(kind, b, threads, timer=lambda k, s: times.__setitem__(k, s), **kw)
            walls.append(time.perf_counter() - t0)
            phases.append(times)
        k = int(np.argsort(walls)[len(walls) // 2])
        return walls[k], phases[k]

    tT, times = median_run(T)
    t1, _ = median_run(1, cells=cut(C, 16), faces=cut(F, 16))
    Fs = cut(F, 16)[1]
    return {"value": F / tT, "unit": "face-fluxes/s", "cores": T, "kind": "port",
            "sample": ("the oracle's FP64 pass over the whole mesh (%d cells, %d faces) on %d "
                       "threads: %.2f s (gradients %.2f, faces %.2f, gather %.2f); one thread "
                       "over the first %d faces and %d cells: %.2f s, %.3g face-fluxes/s; "
                       "medians of %d runs after one warm-up run each"
                       % (C, F, T, tT, times["gradients"], times["faces"], times["gather"],
                          Fs, cut(C, 16)[1], t1, Fs / t1, runs)),
            "single_thread": Fs / t1, "affinity_cores": aff, "nproc": os.cpu_count(),
            "cpu_model": cpu_model()}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-riemann", action="store_true", help="skip the batched Riemann solves")
    ap.add_argument("--no-2d", action="store_true", help="skip the 2D flux pass")
    ap.add_argument("--no-serve", action="store_true", help="skip serving (2,000-mesh ensemble)")
    ap.add_argument("--no-peaks", action="store_true",
                    help="skip measuring the copy and FMA ceilings")
    ap.add_argument("--riemann-N", type=int, default=None)
    ap.add_argument("--serve-meshes", type=int, default=2000)
    ap.add_argument("--serve-requests", type=int, default=4000)
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="also run a plain FP32 copy kernel over this many MiB (5 launches), "
                         "the scale reference for hardware byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the face-flux path runs on a single "
              "MI355X; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")

    log("3D benchmark mesh (2,097,152 Poisson generators)")
    b3, t_mesh = mesh3d()
    log("  %d cells, %d faces, host tessellation %.1f s" % (b3["C"], b3["F"], t_mesh))
    fp3 = bench_fluxpass(b3, 0, args.steps, args.warmup, dev)
    fp3["host_mesh_s"] = t_mesh
    ms = fp3["mixed"]["ms_mean"]
    F = b3["F"]
    face_k = "sfx::flux::face_flux_kernel<3, 0, 1, float, double, true>"
    roof = roofline(b3, fp3, "3d", face_k,
                    ["sfx::flux::gradient_lds_kernel<3>", face_k,
                     "sfx::flux::gather_kernel<3, 1, float, 16>"])
    out = {"metric": METRIC, "value": F / (ms * 1e-3), "unit": "face-fluxes/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32+f64",
           "data": "synthetic (seeded Poisson generators, host Voronoi tessellation, Sedov "
                   "blast initial conditions)",
           "config": {"workload": "one face-flux pass (gradients + face fluxes + ordered "
                                  "gather), 3D periodic unit box, 2,097,152 Poisson generators "
                                  "(seed %d), Sedov blast, exact Riemann solver, mixed "
                                  "precision" % SEED_3D,
                      "cells": b3["C"], "faces": F, "solver": "exact", "precision": "mixed",
                      "seed": SEED_3D},
           "roofline": roof,
           "fluxpass_3d_sedov": fp3}
    out["mixed_vs_fp64_speedup"] = fp3["mixed_speedup"]
    out["errors"] = fp3["mixed_vs_fp64"]
    mp = None
    if not args.no_peaks:
        log("measured ceilings (copy, FMA loops)")
        try:
            mp = measured_peaks(dev)
        except (OSError, RuntimeError) as e:
            # a missing or failing probe library leaves the ceilings out,
            # not the line
            out["measured_peaks"] = {"error": str(e)}
        else:
            out["measured_peaks"] = mp
            measured_fracs(roof, mp)
    if args.copy_probe_mb:
        copy_probe(args.copy_probe_mb, dev)
    if not args.no_cpu:
        log("cpu baseline (the oracle's pass over the same mesh)")
        out["cpu_baseline"] = cpu_baseline(b3, 0)
    else:
        out["cpu_baseline"] = None
    del b3
    if not args.no_2d:
        log("flux pass 2D (512x512 KH mesh)")
        b2, t2 = mesh2d()
        fp2 = bench_fluxpass(b2, 1, args.steps, args.warmup, dev)
        fp2["host_mesh_s"] = t2
        fp2["workload"] = ("2D periodic unit box, 512x512 jittered generators (seed %d) in "
                           "Hilbert order, Kelvin-Helmholtz, HLLC" % SEED_2D)
        face2 = "sfx::flux::face_flux_kernel<2, 1, 1, float, double, true>"
        fp2["roofline"] = roofline(b2, fp2, "2d", face2,
                                   ["sfx::flux::gradient_kernel<2, float, double, true>", face2,
                                    "sfx::flux::gather_kernel<2, 1, float, 0>"])
        measured_fracs(fp2["roofline"], mp)
        out["fluxpass_2d_kh"] = fp2
        del b2
    if not args.no_riemann:
        log("batched Riemann solves (4,194,304 problems)")
        out["riemann_4m"] = bench_riemann(args.steps, args.warmup, dev, args.riemann_N)
    if not args.no_serve:
        log("serving (2,000-mesh ensemble)")
        out["serving"] = bench_serving(dev, args.serve_meshes, args.serve_requests)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
