"""Serving an ensemble of independent meshes held in HBM (the 2,000-mesh
serving workload): ctypes wrapper of the sfx_ensemble_* C ABI, the ensemble builder
and the arrival trace.

Each request is one hydro step of one mesh: VorTessManager::hydro (gradients,
face fluxes, ordered gather) then update_Q + get_W (Shadowfax
src/Simulation.cpp:414-483).  A batched step runs every admitted mesh through
one replay of a captured HIP graph; the plain path steps one mesh with
launches sized to it.
"""
import ctypes as ct

import numpy as np

from . import _capi as Cp
from . import mesh as M
from .riemann import RiemannSolver
from .workloads import uniform

_dp = ct.POINTER(ct.c_double)
_i32p = ct.POINTER(ct.c_int32)
_i64p = ct.POINTER(ct.c_int64)

ENS_W, ENS_Q, ENS_DQ = 0, 1, 2
SERVE_BATCHED, SERVE_PLAIN, SERVE_QUEUE = 0, 1, 2


class ServeStats(ct.Structure):
    _fields_ = [("nreq", ct.c_int64), ("nbatch", ct.c_int64), ("max_batch", ct.c_int64),
                ("wall_s", ct.c_double), ("busy_s", ct.c_double), ("p50_s", ct.c_double),
                ("p99_s", ct.c_double), ("mean_s", ct.c_double), ("max_s", ct.c_double),
                ("mean_batch", ct.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _lib():
    L = Cp.lib()
    if not getattr(L, "_ens_sigs_set", False):
        vp, i64, i32 = ct.c_void_p, ct.c_int64, ct.c_int32
        md = ct.POINTER(Cp.MeshDesc)
        L.sfx_ensemble_create.argtypes = [vp, i64, i64, i64, i64, i64, i64, ct.POINTER(vp)]
        L.sfx_ensemble_destroy.argtypes = [vp]
        L.sfx_ensemble_add_mesh.argtypes = [vp, md, vp, i64, i64, _i32p]
        L.sfx_ensemble_info.argtypes = [vp, i32, _i64p, _i64p]
        L.sfx_ensemble_step.argtypes = [vp, _i32p, i32]
        L.sfx_ensemble_step_plain.argtypes = [vp, i32]
        L.sfx_ensemble_sync.argtypes = [vp]
        L.sfx_ensemble_read.argtypes = [vp, i32, ct.c_int, _dp]
        L.sfx_ensemble_set_launch.argtypes = [vp, ct.c_int, ct.c_int]
        L.sfx_ensemble_get_launch.argtypes = [vp, ct.POINTER(ct.c_int), ct.POINTER(ct.c_int)]
        L.sfx_ensemble_autotune.argtypes = [vp, _i32p, i32, ct.c_int, _dp]
        L.sfx_serve_admit.argtypes = [i64, _i32p, _i64p, _i64p, i64, i64, i32, _i32p, _i32p]
        L.sfx_ensemble_serve.argtypes = [vp, ct.c_int, i64, _dp, _i32p, _dp, _i32p,
                                         ct.POINTER(ServeStats)]
        L._ens_sigs_set = True
    return L


def _p(a, t):
    return a.ctypes.data_as(t)


def serve_admit(pending, mesh_cells, mesh_faces, arena_cells, arena_faces, max_batch):
    """FIFO admission by scratch footprint (host only): indices into pending."""
    pending = np.ascontiguousarray(pending, np.int32)
    mc = np.ascontiguousarray(mesh_cells, np.int64)
    mf = np.ascontiguousarray(mesh_faces, np.int64)
    out = np.zeros(max(1, len(pending)), np.int32)
    n = ct.c_int32(0)
    Cp.check(_lib().sfx_serve_admit(len(pending), _p(pending, _i32p), _p(mc, _i64p),
                                    _p(mf, _i64p), arena_cells, arena_faces, max_batch,
                                    _p(out, _i32p), ct.byref(n)), "sfx_serve_admit")
    return out[:n.value].copy()


class Ensemble:
    """Meshes of one dimension and precision resident together on the
    device."""

    def __init__(self, solver: RiemannSolver, max_meshes, max_cells, max_faces, max_entries,
                 arena_cells, arena_faces):
        self.s = solver
        self.ndim = solver.ndim
        self.h = ct.c_void_p()
        L = _lib()
        Cp.check(L.sfx_ensemble_create(solver.handle, max_meshes, max_cells, max_faces,
                                       max_entries, arena_cells, arena_faces, ct.byref(self.h)),
                 "sfx_ensemble_create")
        self.cells, self.faces = [], []
        self.arena_cells, self.arena_faces = arena_cells, arena_faces

    def close(self):
        if self.h:
            _lib().sfx_ensemble_destroy(self.h)
            self.h = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, dmesh, Q=None):
        """Copy a DeviceMesh (its current W, vel, ... and Q) into the
        ensemble; returns the mesh id."""
        mid = ct.c_int32(-1)
        q = dmesh.Q if Q is None else Q
        Cp.check(_lib().sfx_ensemble_add_mesh(self.h, ct.byref(dmesh.desc),
                                              ct.c_void_p(q.data_ptr()), dmesh.E, dmesh.G,
                                              ct.byref(mid)), "sfx_ensemble_add_mesh")
        self.cells.append(dmesh.C)
        self.faces.append(dmesh.F)
        return mid.value

    def step(self, ids):
        ids = np.ascontiguousarray(ids, np.int32)
        Cp.check(_lib().sfx_ensemble_step(self.h, _p(ids, _i32p), len(ids)), "sfx_ensemble_step")

    def step_plain(self, mid):
        Cp.check(_lib().sfx_ensemble_step_plain(self.h, int(mid)), "sfx_ensemble_step_plain")

    def sync(self):
        Cp.check(_lib().sfx_ensemble_sync(self.h), "sfx_ensemble_sync")

    def read(self, mid, what=ENS_W):
        """W, Q or dQ of one mesh, component-major (ndim+2, C)."""
        out = np.zeros((self.cells[mid], self.ndim + 2))
        Cp.check(_lib().sfx_ensemble_read(self.h, int(mid), what, _p(out, _dp)),
                 "sfx_ensemble_read")
        return out.T.copy()

    def set_launch(self, block, blocks_per_cu):
        Cp.check(_lib().sfx_ensemble_set_launch(self.h, block, blocks_per_cu),
                 "sfx_ensemble_set_launch")

    def get_launch(self):
        b, p = ct.c_int(0), ct.c_int(0)
        Cp.check(_lib().sfx_ensemble_get_launch(self.h, ct.byref(b), ct.byref(p)),
                 "sfx_ensemble_get_launch")
        return b.value, p.value

    def autotune(self, ids, reps=5):
        ids = np.ascontiguousarray(ids, np.int32)
        best = ct.c_double(0.0)
        Cp.check(_lib().sfx_ensemble_autotune(self.h, _p(ids, _i32p), len(ids), reps,
                                              ct.byref(best)), "sfx_ensemble_autotune")
        return self.get_launch() + (best.value,)

    def serve(self, arrival_s, mesh_id, mode=SERVE_BATCHED):
        a = np.ascontiguousarray(arrival_s, np.float64)
        m = np.ascontiguousarray(mesh_id, np.int32)
        lat = np.zeros(len(a))
        bat = np.zeros(len(a), np.int32)
        st = ServeStats()
        Cp.check(_lib().sfx_ensemble_serve(self.h, mode, len(a), _p(a, _dp), _p(m, _i32p),
                                           _p(lat, _dp), _p(bat, _i32p), ct.byref(st)),
                 "sfx_ensemble_serve")
        return lat, bat, st.as_dict()


# ------------------------------------------------------------- workload --

def mesh_sizes(n, seed, lo=1e3, hi=2.5e5):
    """Cell counts log-uniform in [lo, hi] (counter-based RNG)."""
    u = uniform(seed, 300, np.arange(n, dtype=np.uint64))
    return np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))


def size_classes(nclasses, lo=1e3, hi=2.5e5):
    """Jittered-grid side lengths of nclasses log-spaced mesh sizes."""
    c = np.exp(np.linspace(np.log(lo), np.log(hi), nclasses))
    return np.unique(np.round(np.sqrt(c)).astype(int))


def ensemble_ics(kind_ic, pos, mesh, gamma, rng_u):
    """Sod, Kelvin-Helmholtz or Sedov states at the (translated) generator
    positions of a mesh."""
    mm = dict(mesh)
    mm["pos"] = pos
    if kind_ic == 0:
        return M.sod_ics(mm, axis=int(rng_u * 2) % 2)
    if kind_ic == 1:
        return M.kh_ics(mm)
    return M.sedov_ics(mm, gamma=gamma, ncells=16)


def ensemble_workload(n_meshes=2000, seed=20240605, nclasses=64, gamma=5.0 / 3.0, kind=1):
    """The serving ensemble: n_meshes periodic 2D meshes with cell counts
    log-uniform in [1e3, 2.5e5], quantised to nclasses log-spaced jittered
    grids (n x n generators in Hilbert order, tessellated once per class);
    each mesh is its class's generator set translated by its own random
    periodic shift (the same tessellation moved, so its own positions), with
    its own initial conditions (Sod, Kelvin-Helmholtz or Sedov, drawn from the
    seed) and its own CFL time step.  Yields (class bundle, per-mesh state)
    lazily through build()."""
    sizes = mesh_sizes(n_meshes, seed)
    sides = size_classes(nclasses)
    # nearest class in log size
    cls = np.abs(np.log(sizes)[:, None] - np.log(sides.astype(float) ** 2)[None, :]).argmin(1)
    ctr = np.arange(n_meshes, dtype=np.uint64)
    shift = np.stack([uniform(seed, 310, ctr), uniform(seed, 311, ctr)])
    icu = uniform(seed, 312, ctr)
    ic = np.minimum((icu * 3).astype(int), 2)
    return {"sides": sides, "cls": cls, "shift": shift, "ic": ic, "icu": icu,
            "seed": seed, "gamma": gamma, "kind": kind, "n": n_meshes}


def class_bundle(side, seed, gamma, kind):
    pos = M.hilbert_sorted(M.jittered_grid(int(side), 0.25, seed=seed + int(side)))
    m = M.voronoi_periodic(pos, nthreads=16)
    return M.prepare(m, M.kh_ics(m), gamma, kind=kind)


def mesh_state(b, pos, ic, icu, gamma, kind):
    """Per-mesh W, vel, Q, mass, dt on a class bundle b."""
    W = ensemble_ics(ic, pos, b, gamma, icu)
    vel = np.ascontiguousarray(W[1:3])
    Q = M.get_Q(W, b["volume"], gamma, kind)
    dt = M.cfl_dt(b, W, vel, gamma)
    return W, vel, Q, dt


def poisson_trace(nreq, rate, n_meshes, seed):
    """Arrival times (Poisson process of the given rate, s^-1) and mesh ids
    (uniform), from the counter-based RNG."""
    ctr = np.arange(nreq, dtype=np.uint64)
    gaps = -np.log1p(-uniform(seed, 320, ctr)) / rate
    mesh = np.minimum((uniform(seed, 321, ctr) * n_meshes).astype(np.int64), n_meshes - 1)
    return np.cumsum(gaps), mesh.astype(np.int32)


def ensemble_classes(w):
    """One tessellated, prepared mesh per size class in use."""
    return {int(k): class_bundle(w["sides"][k], w["seed"], w["gamma"], w["kind"])
            for k in np.unique(w["cls"])}


def build_ensemble(solver, dev, w, classes=None, arena_cells=4_000_000, log=None):
    """The serving ensemble on the device for one solver (precision): every
    mesh's own states on its class's tessellation.  Returns (Ensemble,
    per-mesh faces)."""
    import torch
    g, kind = w["gamma"], w["kind"]
    classes = classes or ensemble_classes(w)
    tmpl = {k: (b, M.DeviceMesh(b, solver.precision, dev)) for k, b in classes.items()}
    sizes = [(tmpl[k][0]["C"], tmpl[k][0]["F"], int(tmpl[k][0]["cf_off"][-1])) for k in w["cls"]]
    C = sum(x[0] for x in sizes)
    F = sum(x[1] for x in sizes)
    E = sum(x[2] for x in sizes)
    cmax = max(x[0] for x in sizes)
    ac = max(arena_cells, cmax)
    af = max(int(3.2 * ac), max(x[1] for x in sizes))
    ens = Ensemble(solver, w["n"], C, F, E, ac, af)
    faces = np.zeros(w["n"], np.int64)
    for i in range(w["n"]):
        b, dm = tmpl[w["cls"][i]]
        pos = np.mod(b["pos"] + w["shift"][:, i:i + 1], 1.0)
        W, vel, Q, dt = mesh_state(b, pos, int(w["ic"][i]), float(w["icu"][i]), g, kind)
        dm.arr["W"].copy_(torch.from_numpy(np.ascontiguousarray(W.T)))
        dm.arr["vel"].copy_(torch.from_numpy(np.ascontiguousarray(vel.T)))
        dm.arr["mass"].copy_(torch.from_numpy(np.ascontiguousarray(Q[0])))
        dm.arr["dt"].fill_(dt)
        dm.Q.copy_(torch.from_numpy(np.ascontiguousarray(Q.T)))
        mid = ens.add(dm)
        faces[mid] = dm.F
        if log and (i + 1) % 500 == 0:
            log("  %d meshes" % (i + 1))
    return ens, faces
