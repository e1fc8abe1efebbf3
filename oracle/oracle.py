"""ctypes bindings to the CPU oracle (liboracle.so) and, when built, to the
reference harness libraries in oracle/_ref/.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.

Array convention: SoA, component-major, numpy float64 arrays of shape
(ncomp, N) made C-contiguous.
"""
import ctypes as ct
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
EXACT, HLLC = 0, 1

ST_SOLVER, ST_NANFLUX, ST_ITERCAP, ST_UPDATEQ = 1, 2, 4, 8

_dp = ct.POINTER(ct.c_double)
_u8p = ct.POINTER(ct.c_uint8)
_i32p = ct.POINTER(ct.c_int32)
_i64p = ct.POINTER(ct.c_int64)


def build(with_ref=None):
    """Compile liboracle.so (and oracle/_ref when the reference tree exists)."""
    subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])
    if with_ref is None:
        with_ref = os.path.isdir("/root/reference/src")
    if with_ref:
        subprocess.check_call(["make", "-s", "-j4", "-C", HERE, "ref"])


class SolverParams(ct.Structure):
    _fields_ = [("kind", ct.c_int), ("ndim", ct.c_int), ("gamma", ct.c_double),
                ("gp1d2g", ct.c_double), ("gm1d2g", ct.c_double),
                ("gm1dgp1", ct.c_double), ("tdgp1", ct.c_double),
                ("tdgm1", ct.c_double), ("gm1d2", ct.c_double),
                ("tgdgm1", ct.c_double), ("ginv", ct.c_double),
                ("tol", ct.c_double), ("cutoff", ct.c_double)]


class Counters(ct.Structure):
    _fields_ = [("neval", ct.c_uint64), ("nbrent", ct.c_uint64),
                ("nvacuum", ct.c_uint64), ("niter", ct.c_uint64)]


class MeshStruct(ct.Structure):
    _fields_ = [("ndim", ct.c_int), ("C", ct.c_int64), ("F", ct.c_int64),
                ("W", _dp), ("grad", _dp), ("centroid", _dp), ("pos", _dp),
                ("vel", _dp), ("mass", _dp), ("total_area", _dp), ("dt", _dp),
                ("volume", _dp), ("active", _u8p), ("left", _i32p),
                ("right", _i32p), ("area", _dp), ("midpoint", _dp),
                ("shift", _dp), ("ghost", _u8p), ("cf_off", _i64p),
                ("cf_face", _i32p), ("geo_dL", _dp), ("geo_dR", _dp),
                ("geo_rRL", _dp), ("geo_mc", _dp), ("geo_rcent", _dp)]


class Timeline(ct.Structure):
    """sfxo_timeline: the integer timeline of TimeLine."""
    _fields_ = [("integer_maxtime", ct.c_uint64), ("current_time", ct.c_uint64),
                ("timestep", ct.c_uint64), ("max_timestep", ct.c_uint64),
                ("min_timestep", ct.c_uint64), ("tottime", ct.c_double),
                ("mintime", ct.c_double), ("cfl", ct.c_double),
                ("global_timestep", ct.c_int32), ("pad", ct.c_int32)]


def timeline(**kw):
    t = Timeline()
    for k, v in kw.items():
        setattr(t, k, v)
    return t


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run make -C oracle")
        _lib = _bind(ct.CDLL(path))
    return _lib


def _bind(L):
    if True:
        L.sfxo_solver_init.argtypes = [ct.POINTER(SolverParams), ct.c_int,
                                       ct.c_int, ct.c_double, ct.c_double,
                                       ct.c_double]
        L.sfxo_solve_for_flux_batch.restype = ct.c_uint
        L.sfxo_solve_for_flux_batch.argtypes = [
            ct.POINTER(SolverParams), ct.c_int64, _dp, _dp, _dp, _dp, _u8p,
            _dp, _u8p, ct.POINTER(Counters)]
        L.sfxo_exact_solve_batch.restype = ct.c_uint
        L.sfxo_exact_solve_batch.argtypes = [
            ct.POINTER(SolverParams), ct.c_int64, _dp, _dp, _dp, _u8p, _dp,
            _dp, _u8p, ct.POINTER(Counters)]
        L.sfxo_time_derivative.argtypes = [ct.POINTER(SolverParams), _dp, _dp,
                                           _dp]
        L.sfxo_limit.restype = ct.c_double
        L.sfxo_limit.argtypes = [ct.c_double, ct.c_double, ct.c_double, _dp,
                                 ct.c_double, ct.c_int]
        L.sfxo_get_W.restype = ct.c_uint
        L.sfxo_get_W.argtypes = [ct.POINTER(SolverParams), ct.c_double, _dp,
                                 _dp]
        L.sfxo_get_Q.argtypes = [ct.POINTER(SolverParams), ct.c_double, _dp,
                                 _dp]
        L.sfxo_gradients.argtypes = [ct.POINTER(SolverParams),
                                     ct.POINTER(MeshStruct), _dp]
        L.sfxo_face_velocities.argtypes = [ct.POINTER(MeshStruct), _dp, _dp]
        L.sfxo_gradients_range.argtypes = [ct.POINTER(SolverParams),
                                           ct.POINTER(MeshStruct), ct.c_int64,
                                           ct.c_int64, _dp]
        L.sfxo_face_flux_range.restype = ct.c_uint
        L.sfxo_face_flux_range.argtypes = [ct.POINTER(SolverParams),
                                           ct.POINTER(MeshStruct), ct.c_int64,
                                           ct.c_int64, _dp, _u8p,
                                           ct.POINTER(Counters)]
        L.sfxo_gather.argtypes = [ct.POINTER(MeshStruct), _i64p, _i32p, _dp,
                                  ct.c_int64, ct.c_int64, _dp]
        L.sfxo_face_flux_pass.restype = ct.c_uint
        L.sfxo_face_flux_pass.argtypes = [ct.POINTER(SolverParams),
                                          ct.POINTER(MeshStruct), _dp, _dp,
                                          _u8p, ct.POINTER(Counters)]
        L.sfxo_update.restype = ct.c_uint
        L.sfxo_update.argtypes = [ct.POINTER(SolverParams), ct.c_int64, _dp,
                                  _dp, _dp, _dp, ct.POINTER(ct.c_uint64),
                                  ct.c_uint64, _u8p]
        _u64p = ct.POINTER(ct.c_uint64)
        L.sfxo_hydro_timestep.restype = ct.c_uint
        L.sfxo_hydro_timestep.argtypes = [
            ct.POINTER(SolverParams), ct.c_int64, _dp, _dp, _dp,
            ct.POINTER(Timeline), _u64p, _u64p, _dp, _u8p, _u64p, _u64p, _u8p]
    return L


_flops = None


def count_flops(fn, *args, **kw):
    """Run oracle function fn (gradients, face_flux_pass, ...) on the
    counting build (oracle/libflops.so: the same restatement over a scalar
    that counts +, -, *, /, sqrt and transcendentals).  Returns (fn's
    result, operations, transcendentals)."""
    global _lib, _flops
    if _flops is None:
        path = os.path.join(HERE, "libflops.so")
        if not os.path.exists(path):
            subprocess.check_call(["make", "-s", "-C", HERE, "libflops.so"])
        L = _bind(ct.CDLL(path))
        L.sfxf_tally.argtypes = [ct.POINTER(ct.c_ulonglong)]
        _flops = L
    saved = lib()
    _lib = _flops
    try:
        _flops.sfxf_reset()
        out = fn(*args, **kw)
        t = (ct.c_ulonglong * 2)()
        _flops.sfxf_tally(t)
    finally:
        _lib = saved
    return out, int(t[0]), int(t[1])


def _p(a, t=_dp):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


def params(kind, ndim, gamma=5.0 / 3.0, tol=1e-8, cutoff=-5.0):
    s = SolverParams()
    lib().sfxo_solver_init(ct.byref(s), kind, ndim, gamma, tol, cutoff)
    return s


def solve_for_flux(kind, ndim, WL, WR, n, v=None, refl=None, gamma=5.0 / 3.0,
                   tol=1e-8, cutoff=-5.0):
    """Batched solve_for_flux.  Returns (F, status, counters)."""
    s = params(kind, ndim, gamma, tol, cutoff)
    WL, WR, n = _c(WL), _c(WR), _c(n)
    N = WL.shape[1]
    v = None if v is None else _c(v)
    refl = None if refl is None else _c(refl, np.uint8)
    F = np.zeros((ndim + 2, N))
    st = np.zeros(N, np.uint8)
    c = Counters()
    lib().sfxo_solve_for_flux_batch(ct.byref(s), N, _p(WL), _p(WR), _p(n),
                                    _p(v), _p(refl, _u8p), _p(F),
                                    _p(st, _u8p), ct.byref(c))
    return F, st, c


def exact_solve(ndim, WL, WR, n, refl=None, gamma=5.0 / 3.0, tol=1e-8,
                cutoff=-5.0):
    """Batched ExactRiemannSolver::solve.  Returns (Whalf, status, mach),
    mach starting from 0 per problem."""
    s = params(EXACT, ndim, gamma, tol, cutoff)
    WL, WR, n = _c(WL), _c(WR), _c(n)
    N = WL.shape[1]
    refl = None if refl is None else _c(refl, np.uint8)
    out = np.zeros((ndim + 2, N))
    mach = np.zeros(N)
    st = np.zeros(N, np.uint8)
    c = Counters()
    lib().sfxo_exact_solve_batch(ct.byref(s), N, _p(WL), _p(WR), _p(n),
                                 _p(refl, _u8p), _p(out), _p(mach), _p(st, _u8p),
                                 ct.byref(c))
    return out, st, mach


def time_derivative(kind, ndim, W, grad, gamma=5.0 / 3.0):
    """W: (NC, N); grad: (ndim, NC, N) -> (NC, N)."""
    s = params(kind, ndim, gamma)
    NC = ndim + 2
    W = _c(W)
    grad = _c(grad)
    N = W.shape[1]
    out = np.zeros((NC, N))
    w = np.zeros(NC)
    g = np.zeros(ndim * NC)
    o = np.zeros(NC)
    for i in range(N):
        w[:] = W[:, i]
        g[:] = grad[:, :, i].reshape(-1)
        lib().sfxo_time_derivative(ct.byref(s), _p(w), _p(g), _p(o))
        out[:, i] = o
    return out


def limit(phimid0, phiL, phiR, d, r, ndim):
    d = _c(d)
    return lib().sfxo_limit(phimid0, phiL, phiR, _p(d), r, ndim)


def get_W(kind, ndim, volume, Q, gamma=5.0 / 3.0):
    s = params(kind, ndim, gamma)
    Q = _c(Q).copy()
    N = Q.shape[1]
    W = np.zeros_like(Q)
    q = np.zeros(ndim + 2)
    w = np.zeros(ndim + 2)
    for i in range(N):
        q[:] = Q[:, i]
        lib().sfxo_get_W(ct.byref(s), float(volume[i]), _p(q), _p(w))
        Q[:, i] = q
        W[:, i] = w
    return W, Q


def get_Q(kind, ndim, volume, W, gamma=5.0 / 3.0):
    s = params(kind, ndim, gamma)
    W = _c(W)
    N = W.shape[1]
    Q = np.zeros_like(W)
    q = np.zeros(ndim + 2)
    w = np.zeros(ndim + 2)
    for i in range(N):
        w[:] = W[:, i]
        lib().sfxo_get_Q(ct.byref(s), float(volume[i]), _p(w), _p(q))
        Q[:, i] = q
    return Q


def mesh_struct(mesh):
    """mesh: dict of numpy arrays (see shadowfax_amd.mesh) -> (struct, keep)."""
    keep = {}

    def arr(name, dtype=np.float64):
        a = mesh.get(name)
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=dtype)
        keep[name] = a
        return a

    m = MeshStruct()
    m.ndim = int(mesh["ndim"])
    m.C = int(mesh["C"])
    m.F = int(mesh["F"])
    for name in ("W", "grad", "centroid", "pos", "vel", "mass", "total_area",
                 "dt", "volume", "area", "midpoint", "shift"):
        a = arr(name)
        setattr(m, name, _p(a))
    # the stored per-face geometry, when the bundle carries it under geo_*
    for name in ("geo_dL", "geo_dR", "geo_rRL", "geo_mc", "geo_rcent"):
        a = arr(name)
        setattr(m, name, _p(a))
    for name, dt, t in (("active", np.uint8, _u8p), ("ghost", np.uint8, _u8p),
                        ("left", np.int32, _i32p), ("right", np.int32, _i32p),
                        ("cf_off", np.int64, _i64p),
                        ("cf_face", np.int32, _i32p)):
        a = arr(name, dt)
        setattr(m, name, _p(a, t))
    return m, keep


def gradients(kind, mesh, gamma=5.0 / 3.0):
    ndim = int(mesh["ndim"])
    s = params(kind, ndim, gamma)
    m, keep = mesh_struct(mesh)
    out = np.zeros((ndim, ndim + 2, int(mesh["C"])))
    lib().sfxo_gradients(ct.byref(s), ct.byref(m), _p(out))
    return out


def face_velocities(mesh):
    """(vface, normal), each (ndim, F)."""
    m, keep = mesh_struct(mesh)
    nd, F = int(mesh["ndim"]), int(mesh["F"])
    v = np.zeros((nd, F))
    n = np.zeros((nd, F))
    lib().sfxo_face_velocities(ct.byref(m), _p(v), _p(n))
    return v, n


def face_flux_pass(kind, mesh, gamma=5.0 / 3.0, tol=1e-8, cutoff=-5.0):
    """Returns (dQ (NC,C), dQL_face (NC,F), face_status (F,), counters)."""
    ndim = int(mesh["ndim"])
    s = params(kind, ndim, gamma, tol, cutoff)
    m, keep = mesh_struct(mesh)
    C, F = int(mesh["C"]), int(mesh["F"])
    dQ = np.zeros((ndim + 2, C))
    dQL = np.zeros((ndim + 2, F))
    st = np.zeros(F, np.uint8)
    c = Counters()
    lib().sfxo_face_flux_pass(ct.byref(s), ct.byref(m), _p(dQ), _p(dQL),
                              _p(st, _u8p), ct.byref(c))
    return dQ, dQL, st, c


def face_states(kind, mesh, gamma=5.0 / 3.0, tol=1e-8, cutoff=-5.0):
    """(WL, WR) [(ndim+2), F]: the face-frame states every face hands its
    Riemann solver, after the predictor (zeros for skipped faces)."""
    ndim = int(mesh["ndim"])
    s = params(kind, ndim, gamma, tol, cutoff)
    m, keep = mesh_struct(mesh)
    F = int(mesh["F"])
    WL = np.zeros((ndim + 2, F))
    WR = np.zeros((ndim + 2, F))
    lib().sfxo_face_states(ct.byref(s), ct.byref(m), _p(WL), _p(WR))
    return WL, WR


def threaded_pass(kind, mesh, nthreads, gamma=5.0 / 3.0, tol=1e-8, cutoff=-5.0,
                  cells=None, faces=None, timer=None):
    """The flux pass (gradients, face fluxes, ordered gather) on nthreads
    threads, each taking a contiguous range of cells or faces (ctypes
    releases the interpreter lock during the calls).  cells / faces: (lo, hi)
    ranges to restrict the work to (a sample), default all.  Returns (grad,
    dQ, dQL, status); timer(name, seconds) is called per phase."""
    import time
    from concurrent.futures import ThreadPoolExecutor
    ndim = int(mesh["ndim"])
    s = params(kind, ndim, gamma, tol, cutoff)
    C, F = int(mesh["C"]), int(mesh["F"])
    c0, c1 = cells or (0, C)
    f0, f1 = faces or (0, F)
    grad = np.zeros((ndim, ndim + 2, C))
    dQL = np.zeros((ndim + 2, F))
    st = np.zeros(F, np.uint8)
    dQ = np.zeros((ndim + 2, C))
    goff = _c(mesh["gather_off"], np.int64)
    gent = _c(mesh["gather_entry"], np.int32)
    L = lib()

    def cuts(lo, hi):
        e = np.linspace(lo, hi, nthreads + 1).astype(np.int64)
        return list(zip(e[:-1], e[1:]))

    with ThreadPoolExecutor(nthreads) as ex:
        m, keep = mesh_struct(mesh)
        t0 = time.perf_counter()
        list(ex.map(lambda r: L.sfxo_gradients_range(ct.byref(s), ct.byref(m), int(r[0]),
                                                     int(r[1]), _p(grad)), cuts(c0, c1)))
        t1 = time.perf_counter()
        mm = dict(mesh)
        mm["grad"] = grad
        m2, keep2 = mesh_struct(mm)
        list(ex.map(lambda r: L.sfxo_face_flux_range(ct.byref(s), ct.byref(m2), int(r[0]),
                                                     int(r[1]), _p(dQL), _p(st, _u8p), None),
                    cuts(f0, f1)))
        t2 = time.perf_counter()
        list(ex.map(lambda r: L.sfxo_gather(ct.byref(m2), _p(goff, _i64p), _p(gent, _i32p),
                                            _p(dQL), int(r[0]), int(r[1]), _p(dQ)),
                    cuts(c0, c1)))
        t3 = time.perf_counter()
    if timer:
        timer("gradients", t1 - t0)
        timer("faces", t2 - t1)
        timer("gather", t3 - t2)
    return grad, dQ, dQL, st


def update(kind, ndim, Q, dQ, volume, gamma=5.0 / 3.0, W=None, endtime=None, tnow=0):
    """update_Q for every cell; get_W where endtime == tnow (every cell when
    endtime is None).  W: the states kept by the other cells (zeros if None).
    Returns (Q, W, status)."""
    s = params(kind, ndim, gamma)
    Q = _c(Q).copy()
    dQ = _c(dQ).copy()
    volume = _c(volume)
    C = Q.shape[1]
    W = np.zeros_like(Q) if W is None else _c(W).copy()
    st = np.zeros(C, np.uint8)
    et = None if endtime is None else _c(endtime, np.uint64)
    lib().sfxo_update(ct.byref(s), C, _p(Q), _p(dQ), _p(volume), _p(W),
                      _p(et, ct.POINTER(ct.c_uint64)), int(tnow), _p(st, _u8p))
    return Q, W, st


def hydro_timestep(kind, ndim, W, vel, volume, tl, starttime, endtime, gamma=5.0 / 3.0):
    """TimeLine::calculate_timestep (gas, no gravity) over C cells; W [C][ndim+2],
    vel [C][ndim] item-major.  Returns (start, end, dt, active, system_step,
    nactive, status)."""
    s = params(kind, ndim, gamma)
    W = _c(W)
    vel = _c(vel)
    volume = _c(volume)
    C = volume.shape[0]
    start = _c(starttime, np.uint64).copy()
    end = _c(endtime, np.uint64).copy()
    dt = np.zeros(C)
    act = np.zeros(C, np.uint8)
    st = np.zeros(C, np.uint8)
    ss, na = ct.c_uint64(0), ct.c_uint64(0)
    u64 = ct.POINTER(ct.c_uint64)
    lib().sfxo_hydro_timestep(ct.byref(s), C, _p(W), _p(vel), _p(volume), ct.byref(tl),
                              _p(start, u64), _p(end, u64), _p(dt), _p(act, _u8p),
                              ct.byref(ss), ct.byref(na), _p(st, _u8p))
    return start, end, dt, act, ss.value, na.value, st


# ------------------------------------------------------------------ _ref --

_ref = {}


def ref_available(ndim):
    return os.path.exists(os.path.join(HERE, "_ref", "libsfxref%dd.so" % ndim))


def ref(ndim):
    if ndim not in _ref:
        path = os.path.join(HERE, "_ref", "libsfxref%dd.so" % ndim)
        L = ct.CDLL(path)
        L.sfxr_solve_for_flux_batch.argtypes = [
            ct.c_int, ct.c_double, ct.c_double, ct.c_double, ct.c_int64, _dp,
            _dp, _dp, _dp, _u8p, _dp, _dp, _u8p, ct.POINTER(ct.c_uint64)]
        L.sfxr_exact_solve_batch.argtypes = [
            ct.c_double, ct.c_double, ct.c_double, ct.c_int64, _dp, _dp, _dp,
            _u8p, _dp, _dp, _u8p]
        L.sfxr_time_derivative_batch.argtypes = [ct.c_int, ct.c_double,
                                                 ct.c_int64, _dp, _dp, _dp]
        L.sfxr_get_W_batch.argtypes = [ct.c_int, ct.c_double, ct.c_int64, _dp,
                                       _dp, _dp]
        L.sfxr_get_Q_batch.argtypes = [ct.c_int, ct.c_double, ct.c_int64, _dp,
                                       _dp, _dp, _u8p]
        L.sfxr_mesh_gradients.argtypes = [
            ct.c_int64, ct.c_int64, _dp, _dp, _dp, _dp, _dp, _i32p, _i32p,
            _dp, _dp, _dp, _u8p, _i64p, _i32p, _dp, _dp, _dp]
        _ref[ndim] = L
    return _ref[ndim]


def ref_solve_for_flux(kind, ndim, WL, WR, n, v=None, refl=None,
                       gamma=5.0 / 3.0, tol=1e-8, cutoff=-5.0):
    WL, WR, n = _c(WL), _c(WR), _c(n)
    N = WL.shape[1]
    v = None if v is None else _c(v)
    refl = None if refl is None else _c(refl, np.uint8)
    F = np.zeros((ndim + 2, N))
    WRo = np.zeros((ndim + 2, N))
    st = np.zeros(N, np.uint8)
    cnt = (ct.c_uint64 * 3)()
    ref(ndim).sfxr_solve_for_flux_batch(kind, gamma, tol, cutoff, N, _p(WL),
                                        _p(WR), _p(n), _p(v), _p(refl, _u8p),
                                        _p(F), _p(WRo), _p(st, _u8p), cnt)
    return F, st, WRo


def ref_exact_solve(ndim, WL, WR, n, refl=None, gamma=5.0 / 3.0, tol=1e-8,
                    cutoff=-5.0):
    WL, WR, n = _c(WL), _c(WR), _c(n)
    N = WL.shape[1]
    refl = None if refl is None else _c(refl, np.uint8)
    out = np.zeros((ndim + 2, N))
    mach = np.zeros(N)
    st = np.zeros(N, np.uint8)
    ref(ndim).sfxr_exact_solve_batch(gamma, tol, cutoff, N, _p(WL), _p(WR),
                                     _p(n), _p(refl, _u8p), _p(out), _p(mach),
                                     _p(st, _u8p))
    return out, st, mach


def ref_time_derivative(kind, ndim, W, grad, gamma=5.0 / 3.0):
    W = _c(W)
    grad = _c(grad)
    N = W.shape[1]
    out = np.zeros_like(W)
    ref(ndim).sfxr_time_derivative_batch(kind, gamma, N, _p(W), _p(grad),
                                         _p(out))
    return out


def ref_get_W(kind, ndim, volume, Q, gamma=5.0 / 3.0):
    Q = _c(Q).copy()
    volume = _c(volume)
    W = np.zeros_like(Q)
    ref(ndim).sfxr_get_W_batch(kind, gamma, Q.shape[1], _p(volume), _p(Q),
                               _p(W))
    return W, Q


def ref_get_Q(kind, ndim, volume, W, gamma=5.0 / 3.0):
    W = _c(W)
    volume = _c(volume)
    Q = np.zeros_like(W)
    st = np.zeros(W.shape[1], np.uint8)
    ref(ndim).sfxr_get_Q_batch(kind, gamma, W.shape[1], _p(volume), _p(W),
                               _p(Q), _p(st, _u8p))
    return Q, st


def ref_mesh_gradients(mesh):
    ndim = int(mesh["ndim"])
    C, F = int(mesh["C"]), int(mesh["F"])
    g = {k: np.ascontiguousarray(mesh[k]) for k in mesh if k not in
         ("ndim", "C", "F")}
    out = np.zeros((ndim, ndim + 2, C))
    vface = np.zeros((ndim, F))
    normal = np.zeros((ndim, F))
    f64 = lambda k: _p(np.ascontiguousarray(g[k], np.float64))
    keep = [np.ascontiguousarray(g[k], np.float64) for k in
            ("W", "centroid", "pos", "vel", "volume", "area", "midpoint",
             "shift")]
    W, cen, pos, vel, vol, area, mid, shift = keep
    left = np.ascontiguousarray(g["left"], np.int32)
    right = np.ascontiguousarray(g["right"], np.int32)
    ghost = np.ascontiguousarray(g["ghost"], np.uint8)
    cf_off = np.ascontiguousarray(g["cf_off"], np.int64)
    cf_face = np.ascontiguousarray(g["cf_face"], np.int32)
    ref(ndim).sfxr_mesh_gradients(
        C, F, _p(W), _p(cen), _p(pos), _p(vel), _p(vol), _p(left, _i32p),
        _p(right, _i32p), _p(area), _p(mid), _p(shift), _p(ghost, _u8p),
        _p(cf_off, _i64p), _p(cf_face, _i32p), _p(out), _p(vface), _p(normal))
    return out, vface, normal


def ref_mesh_geometry(mesh):
    """Face area/midpoint and cell volume/centroid computed by the reference
    VorFace/VorCell classes from the mesh's stored face vertex lists."""
    ndim = int(mesh["ndim"])
    C, F = int(mesh["C"]), int(mesh["F"])
    pos = np.ascontiguousarray(mesh["pos"], np.float64)
    fv_off = np.ascontiguousarray(mesh["fv_off"], np.int64)
    fverts = np.ascontiguousarray(mesh["fverts"], np.float64)
    cf_off = np.ascontiguousarray(mesh["cf_off"], np.int64)
    cf_face = np.ascontiguousarray(mesh["cf_face"], np.int32)
    area = np.zeros(F)
    mid = np.zeros((ndim, F))
    vol = np.zeros(C)
    cen = np.zeros((ndim, C))
    ref(ndim).sfxr_mesh_geometry(C, F, _p(pos), _p(fv_off, _i64p), _p(fverts),
                                 _p(cf_off, _i64p), _p(cf_face, _i32p), _p(area), _p(mid),
                                 _p(vol), _p(cen))
    return area, mid, vol, cen
