"""Host-side meshes for the face-flux pass.

Tessellation stays on the host (C++ in libshadowfax_amd.so,
sfx_voronoi_periodic); this module turns a tessellation plus initial
conditions into the arrays the flux-pass kernels read, and uploads them.

Array conventions (numpy, component-major):
  cells: W (ndim+2, C), pos/centroid/vel (ndim, C), volume/mass/total_area/dt (C,)
  faces: left/right (F,), area (F,), midpoint/shift (ndim, F), ghost/twin (F,)
  per-cell face list: cf_off (C+1,), cf_face (cf_off[C],)
"""
import ctypes as ct

import numpy as np

from . import _capi as Cp
from .workloads import splitmix64, uniform

_dp = ct.POINTER(ct.c_double)
_i32p = ct.POINTER(ct.c_int32)
_i64p = ct.POINTER(ct.c_int64)
_u8p = ct.POINTER(ct.c_uint8)


class HostMesh(ct.Structure):
    _fields_ = [("ndim", ct.c_int), ("C", ct.c_int64), ("F", ct.c_int64),
                ("pos", _dp), ("volume", _dp), ("centroid", _dp), ("left", _i32p),
                ("right", _i32p), ("area", _dp), ("midpoint", _dp), ("shift", _dp),
                ("ghost", _u8p), ("twin", _i32p), ("cf_off", _i64p), ("cf_face", _i32p),
                ("fv_off", _i64p), ("fverts", _dp)]


def _lib():
    L = Cp.lib()
    if not getattr(L, "_mesh_sigs_set", False):
        L.sfx_voronoi_periodic.argtypes = [ct.c_int, ct.c_int64, _dp, _dp, ct.c_int,
                                           ct.POINTER(HostMesh)]
        L.sfx_host_mesh_free.argtypes = [ct.POINTER(HostMesh)]
        L.sfx_host_mesh_free.restype = None
        L.sfx_face_geometry.argtypes = [ct.c_int, ct.c_int64, ct.c_int64, _dp, _dp, _i32p,
                                        _i32p, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.sfx_mesh_cell_entries.argtypes = [ct.c_int64, ct.c_int64, _i32p, _i64p, _i32p,
                                            _i32p]
        L.sfx_mesh_gather_index.argtypes = [ct.c_int64, ct.c_int64, _i32p, _i32p, _u8p,
                                            _i64p, _i32p]
        L.sfx_hilbert_order.argtypes = [ct.c_int, ct.c_int64, _dp, _dp, _i64p]
        L._mesh_sigs_set = True
    return L


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


# ------------------------------------------------------------ generators --

def jittered_grid(n, jitter=0.25, seed=1, ndim=2, box=1.0):
    """n^ndim generators (i + 1/2 + U(-jitter, jitter)) * h per axis."""
    h = box / n
    idx = np.indices((n,) * ndim).reshape(ndim, -1)[::-1]  # x fastest
    N = idx.shape[1]
    ctr = np.arange(N, dtype=np.uint64)
    pos = np.empty((ndim, N))
    for d in range(ndim):
        u = uniform(seed, 100 + d, ctr)
        pos[d] = (idx[d] + 0.5 + jitter * (2.0 * u - 1.0)) * h
    return pos


def poisson_points(N, seed=1, ndim=3, box=1.0):
    ctr = np.arange(N, dtype=np.uint64)
    return np.stack([uniform(seed, 200 + d, ctr) * box for d in range(ndim)])


def hilbert_order(pos, box=1.0):
    """Permutation putting the generators in Hilbert order (the order the
    reference keeps its particles in, src/utilities/ParticleVector.cpp:222)."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    ndim, C = pos.shape
    boxv = np.ascontiguousarray(np.broadcast_to(np.asarray(box, np.float64), (ndim,)))
    order = np.zeros(C, np.int64)
    Cp.check(_lib().sfx_hilbert_order(ndim, C, _p(pos), _p(boxv), _p(order, _i64p)),
             "sfx_hilbert_order")
    return order


def hilbert_sorted(pos, box=1.0):
    return np.ascontiguousarray(pos[:, hilbert_order(pos, box)])


# --------------------------------------------------------- tessellation --

def voronoi_periodic(pos, box=1.0, nthreads=0):
    """Periodic Voronoi tessellation (host C++).  Returns a dict of arrays."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    ndim, C = pos.shape
    boxv = np.ascontiguousarray(np.broadcast_to(np.asarray(box, np.float64), (ndim,)))
    hm = HostMesh()
    L = _lib()
    Cp.check(L.sfx_voronoi_periodic(ndim, C, _p(pos), _p(boxv), nthreads, ct.byref(hm)),
             "sfx_voronoi_periodic")
    try:
        F = hm.F
        nE = None

        def arr(ptr, n, dt):
            return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True)

        m = {"ndim": ndim, "C": C, "F": F, "box": boxv.copy()}
        m["pos"] = arr(hm.pos, ndim * C, np.float64).reshape(ndim, C)
        m["volume"] = arr(hm.volume, C, np.float64)
        m["centroid"] = arr(hm.centroid, ndim * C, np.float64).reshape(ndim, C)
        m["left"] = arr(hm.left, F, np.int32)
        m["right"] = arr(hm.right, F, np.int32)
        m["area"] = arr(hm.area, F, np.float64)
        m["midpoint"] = arr(hm.midpoint, ndim * F, np.float64).reshape(ndim, F)
        m["shift"] = arr(hm.shift, ndim * F, np.float64).reshape(ndim, F)
        m["ghost"] = arr(hm.ghost, F, np.uint8)
        m["twin"] = arr(hm.twin, F, np.int32)
        m["cf_off"] = arr(hm.cf_off, C + 1, np.int64)
        nE = int(m["cf_off"][-1])
        m["cf_face"] = arr(hm.cf_face, nE, np.int32)
        m["fv_off"] = arr(hm.fv_off, F + 1, np.int64)
        nV = int(m["fv_off"][-1])
        m["fverts"] = arr(hm.fverts, ndim * nV, np.float64).reshape(ndim, nV)
    finally:
        L.sfx_host_mesh_free(ct.byref(hm))
    return m


def total_area(mesh):
    """GasParticle::get_total_area = VorCell::calculate_total_area: the sum of
    the cell's face areas in face-list order (src/VorCell.cpp:407-411)."""
    C = mesh["C"]
    out = np.zeros(C)
    off, cf = mesh["cf_off"], mesh["cf_face"]
    area = mesh["area"]
    # ordered sum per cell (a python loop over the max face count)
    counts = np.diff(off)
    for k in range(int(counts.max()) if C else 0):
        sel = counts > k
        out[sel] = out[sel] + area[cf[off[:-1][sel] + k]]
    return out


def face_geometry(mesh):
    L = _lib()
    ndim, C, F = mesh["ndim"], mesh["C"], mesh["F"]
    g = {k: np.zeros((ndim, F)) for k in ("dL", "dR", "rRL", "mc")}
    g["rcent"] = np.zeros(F)
    c = lambda k, dt=np.float64: np.ascontiguousarray(mesh[k], dtype=dt)
    pos, cen, mid, sh = c("pos"), c("centroid"), c("midpoint"), c("shift")
    left, right = c("left", np.int32), c("right", np.int32)
    Cp.check(L.sfx_face_geometry(ndim, C, F, _p(pos), _p(cen), _p(left, _i32p),
                                 _p(right, _i32p), _p(mid), _p(sh), _p(g["dL"]), _p(g["dR"]),
                                 _p(g["rRL"]), _p(g["mc"]), _p(g["rcent"])),
             "sfx_face_geometry")
    return g


def cell_closure(b):
    """sfx_cell_closure: every cell's closure residual sum A r_ij/|r_ij|
    [ndim][C] (FP64, host), from the face areas, rRL and the cell entries of
    a prepared bundle."""
    L = _lib()
    nd, C, F = b["ndim"], b["C"], b["F"]
    out = np.zeros((nd, C))
    off = np.ascontiguousarray(b["cf_off"], np.int64)
    ent = np.ascontiguousarray(b["cf_entry"], np.int32)
    area, rRL = np.ascontiguousarray(b["area"], np.float64), np.ascontiguousarray(b["rRL"])
    Cp.check(L.sfx_cell_closure(nd, C, F, _p(off, _i64p), _p(ent, _i32p), _p(area), _p(rRL),
                                _p(out)), "sfx_cell_closure")
    return out


# a closure residual below this fraction of the cell's total face area is the
# round-off of a tessellation whose cells close; the mixed pass then leaves
# the correction out
CLOSURE_TOL = 1e-12


def cell_entries(mesh):
    L = _lib()
    off = np.ascontiguousarray(mesh["cf_off"], np.int64)
    cf = np.ascontiguousarray(mesh["cf_face"], np.int32)
    left = np.ascontiguousarray(mesh["left"], np.int32)
    out = np.zeros(len(cf), np.int32)
    Cp.check(L.sfx_mesh_cell_entries(mesh["C"], mesh["F"], _p(left, _i32p), _p(off, _i64p),
                                     _p(cf, _i32p), _p(out, _i32p)), "sfx_mesh_cell_entries")
    return out


def gather_index(mesh):
    L = _lib()
    C, F = mesh["C"], mesh["F"]
    left = np.ascontiguousarray(mesh["left"], np.int32)
    right = np.ascontiguousarray(mesh["right"], np.int32)
    ghost = np.ascontiguousarray(mesh["ghost"], np.uint8)
    off = np.zeros(C + 1, np.int64)
    Cp.check(L.sfx_mesh_gather_index(C, F, _p(left, _i32p), _p(right, _i32p), _p(ghost, _u8p),
                                     _p(off, _i64p), None), "sfx_mesh_gather_index")
    ent = np.zeros(int(off[-1]), np.int32)
    Cp.check(L.sfx_mesh_gather_index(C, F, _p(left, _i32p), _p(right, _i32p), _p(ghost, _u8p),
                                     _p(off, _i64p), _p(ent, _i32p)), "sfx_mesh_gather_index")
    return off, ent


# ------------------------------------------------------ initial conditions --

def get_Q(W, volume, gamma, kind=0):
    """ExactRiemannSolver::get_Q / HLLCRiemannSolver::get_Q (hydro part)."""
    nd = W.shape[0] - 2
    Q = np.empty_like(W)
    Q[0] = W[0] * volume
    for d in range(nd):
        Q[1 + d] = Q[0] * W[1 + d]
    v2 = W[1] * W[1] + W[2] * W[2]
    if nd == 3:
        v2 = W[1] * W[1] + W[2] * W[2] + W[3] * W[3]
    if kind == 1:
        Q[-1] = volume * W[-1] / (gamma - 1.) + 0.5 * Q[0] * v2
    else:
        Q[-1] = (0.5 * W[0] * v2 + W[-1] / (gamma - 1.)) * volume
    return Q


def kh_ics(mesh):
    """Kelvin-Helmholtz: rho 2, vx +0.5 for |y - 0.5| < 0.25, else rho 1,
    vx -0.5; p 2.5; vy = 0.01 sin(4 pi x); at the generator positions."""
    x, y = mesh["pos"][0], mesh["pos"][1]
    inside = np.abs(y - 0.5) < 0.25
    W = np.zeros((4, mesh["C"]))
    W[0] = np.where(inside, 2.0, 1.0)
    W[1] = np.where(inside, 0.5, -0.5)
    W[2] = 0.01 * np.sin(4.0 * np.pi * x)
    W[3] = 2.5
    return W


def sod_ics(mesh, axis=0):
    x = mesh["pos"][axis]
    nd = mesh["ndim"]
    W = np.zeros((nd + 2, mesh["C"]))
    left = x < 0.5
    W[0] = np.where(left, 1.0, 0.125)
    W[-1] = np.where(left, 1.0, 0.1)
    return W


def sedov_ics(mesh, gamma=5.0 / 3.0, E=1.0, ncells=32, rho=1.0, p0=1e-5):
    """rho 1, p 1e-5, v 0; energy E deposited in the ncells cells nearest the
    box centre: p_i = (gamma-1) (E/ncells) / V_i."""
    nd = mesh["ndim"]
    c = mesh["box"][:, None] * 0.5
    r2 = ((mesh["pos"] - c) ** 2).sum(0)
    hot = np.argsort(r2, kind="stable")[:ncells]
    W = np.zeros((nd + 2, mesh["C"]))
    W[0] = rho
    W[-1] = p0
    W[-1, hot] = (gamma - 1.0) * (E / ncells) / mesh["volume"][hot]
    return W


def cfl_dt(mesh, W, vel, gamma, cfl=0.4):
    """Global time step: cfl * min over cells of h / (|v - v_gen| + c),
    h = sqrt(V/pi) (2D) or cbrt(3V/(4 pi)) (3D) (src/TimeLine.cpp:248-262)."""
    nd = mesh["ndim"]
    V = mesh["volume"]
    h = np.sqrt(V / np.pi) if nd == 2 else np.cbrt(3.0 * V / (4.0 * np.pi))
    cs = np.sqrt(gamma * W[-1] / W[0])
    dv = np.sqrt(((W[1:1 + nd] - vel) ** 2).sum(0))
    return cfl * float(np.min(h / (dv + cs)))


def prepare(mesh, W, gamma=5.0 / 3.0, kind=0, vel=None, dt=None, cfl=0.4, active=None):
    """Everything the flux pass reads, as numpy arrays (FP64).

    dt: None (one global CFL step, cfl_dt), a number (the same step for every
    cell) or an array [C] of per-cell steps (individual time steps, e.g.
    sfx_hydro_timestep's dt); active: [C] flags of the cells whose step starts
    now (the face pass skips faces between two inactive cells), None: all."""
    nd, C = mesh["ndim"], mesh["C"]
    b = dict(mesh)
    b["W"] = np.ascontiguousarray(W, np.float64)
    b["vel"] = np.ascontiguousarray(W[1:1 + nd] if vel is None else vel, np.float64)
    Q = get_Q(b["W"], mesh["volume"], gamma, kind)
    b["Q"] = Q
    b["mass"] = Q[0].copy()
    b["total_area"] = total_area(mesh)
    if dt is None:
        dt = cfl_dt(mesh, b["W"], b["vel"], gamma, cfl)
    if np.ndim(dt) == 0:
        b["dt_value"] = float(dt)
        b["dt"] = np.full(C, float(dt))
    else:
        b["dt"] = np.ascontiguousarray(dt, np.float64).reshape(C)
        b["dt_value"] = float(b["dt"].min()) if C else 0.0
    b["active"] = (np.ones(C, np.uint8) if active is None
                   else np.ascontiguousarray(active, np.uint8).reshape(C))
    b.update(face_geometry(mesh))
    b["cf_entry"] = cell_entries(mesh)
    b["gather_off"], b["gather_entry"] = gather_index(mesh)
    S = cell_closure(b)
    ta = np.where(b["total_area"] > 0, b["total_area"], 1.0)
    b["closure"] = S if C and (np.abs(S).max(0) / ta).max() > CLOSURE_TOL else None
    return b


# ------------------------------------------------------------- device --

T_ARRAYS = ("area", "dL", "dR", "rRL", "mc", "rcent")
F64_ARRAYS = ("mass", "total_area", "dt", "volume")


class DeviceMesh:
    """Device copy of a prepared mesh.  storage: 'fp64' or 'mixed' (FP32
    generator velocities, gradients and face geometry; the states W are FP64
    in both).

    Arrays the kernels gather through a neighbour index are item-major:
    W (C, ndim+2), vel (C, ndim), grad (C, ndim, ndim+2), Q and dQ
    (C, ndim+2), dQL (F, ndim+2); per-face geometry is (ndim, F).  The
    *_soa() accessors return host copies in the component-major layout of
    the prepared bundle (and of the oracle)."""

    # cells ordered by entry count within windows of this many cells, per
    # storage (3D Sedov, 2,097,152 cells: mixed gradients 0.389 ms at 64
    # against 0.422 in cell order and 0.455 at 256 -- the FP32 records are
    # read through a buffer descriptor that fetches no padding, so sorting
    # within the group only stops the padding lanes from sharing cache lines
    # with their neighbours', while wider windows scatter a wave's cells and
    # with them its neighbour rows; FP64 gradients, whose records are read
    # padding and all, 1.088 ms at 256 against 1.339 in cell order and 1.207
    # at 1024; 512 moves 1.48x the gradients' unique bytes instead of 1.59x
    # for 3.7% more time)
    SORT_WINDOW = {"mixed": 64, "fp64": 512}

    def __init__(self, b, storage="fp64", device=None, grouped=None, face_records=None,
                 sort_window=None):
        """grouped (default on): the gradient records in the grouped layout,
        over cells ordered by entry count within windows of sort_window
        cells (sfx_mesh_entry_groups_sorted; default SORT_WINDOW[storage];
        0: in cell order, sfx_mesh_entry_groups).  face_records (default
        on, mixed storage only): the face kernel reads each face's own
        values from the packed records of sfx_mesh_face_records, built with
        the entry records."""
        import torch
        self.torch = torch
        self.dev = device or torch.device("cuda:0")
        self.storage = storage
        self.ndim, self.C, self.F = b["ndim"], b["C"], b["F"]
        tdt = torch.float64 if storage == "fp64" else torch.float32
        self.tdt = tdt
        t = lambda a, dt: torch.tensor(np.ascontiguousarray(a), dtype=dt, device=self.dev)
        self.arr = {}
        for k in T_ARRAYS:
            self.arr[k] = t(b[k], tdt)
        # FP64 in both passes: the states, and the generator velocities (the
        # face frame's velocity is subtracted from the states, and cold gas
        # moving at many times its sound speed loses its relative velocity
        # in FP32: 5e-6 face errors on the golden 3D meshes)
        self.arr["W"] = t(b["W"].T, torch.float64)
        self.arr["vel"] = t(b["vel"].T, torch.float64)
        for k in F64_ARRAYS:
            self.arr[k] = t(b[k], torch.float64)
        self.arr["active"] = t(b["active"], torch.uint8)
        self.arr["ghost"] = t(b["ghost"], torch.uint8)
        for k in ("left", "right", "twin", "cf_entry", "gather_entry"):
            self.arr[k] = t(b[k], torch.int32)
        for k in ("cf_off", "gather_off"):
            self.arr[k] = t(b[k], torch.int64)
        self.Q = t(b["Q"].T, torch.float64)
        nd, NC = self.ndim, self.ndim + 2
        self.E = int(b["cf_off"][-1])
        self.G = int(b["gather_off"][-1])
        ER = Cp.lib().sfx_mesh_entry_record(
            nd, Cp.SFX_PREC_FP64 if storage == "fp64" else Cp.SFX_PREC_MIXED)
        self.grouped = True if grouped is None else bool(grouped)
        if sort_window is None:
            sort_window = self.SORT_WINDOW[storage]
        self.sorted = self.grouped and bool(sort_window)
        if self.grouped:
            off = np.ascontiguousarray(b["cf_off"], np.int64)
            grp = np.zeros((self.C + 63) // 64 + 1, np.int64)
            P64 = lambda a: a.ctypes.data_as(ct.POINTER(ct.c_int64))
            if self.sorted:
                slot = np.zeros(max(self.C, 1), np.int64)
                rows = Cp.lib().sfx_mesh_entry_groups_sorted(self.C, P64(off), int(sort_window),
                                                             P64(slot), P64(grp))
                self.arr["egeo_slot"] = t(slot, torch.int64)
            else:
                rows = Cp.lib().sfx_mesh_entry_groups(self.C, P64(off), P64(grp))
            if rows < 0:
                raise Cp.SfxError(Cp.lib().sfx_last_error().decode())
            self.arr["egeo_group"] = t(grp, torch.int64)
            self.egeo_rows = int(rows)
            self.egeo = torch.zeros((rows * 64, ER), dtype=tdt, device=self.dev)
        else:
            self.egeo = torch.zeros((self.E, ER), dtype=tdt, device=self.dev)
        self.egeo_built = False
        self.grad = torch.zeros((self.C, nd, NC), dtype=tdt, device=self.dev)
        self.dQL = torch.zeros((self.F, NC), dtype=tdt, device=self.dev)
        self.dQ = torch.zeros((self.C, NC), dtype=torch.float64, device=self.dev)
        self.fstatus = torch.zeros(self.F, dtype=torch.uint8, device=self.dev)
        self.cstatus = torch.zeros(self.C, dtype=torch.uint8, device=self.dev)
        self.use_twin = storage == "mixed"
        self.face_records = storage == "mixed" and (face_records is None or bool(face_records))
        self.frec = None
        if self.face_records:
            planes = Cp.lib().sfx_mesh_face_record_planes(nd, Cp.SFX_PREC_MIXED)
            self.frec = torch.zeros((planes, self.F, 4), dtype=torch.float32, device=self.dev)
        # the mixed gradients' closure correction, for face lists whose cells
        # do not close (None: they do)
        if storage == "mixed" and b.get("closure") is not None:
            self.arr["closure"] = t(b["closure"], tdt)
        self.desc = self._desc()

    def _desc(self):
        d = Cp.MeshDesc()
        d.ndim, d.C, d.F = self.ndim, self.C, self.F
        a = self.arr
        p = lambda k: ct.c_void_p(a[k].data_ptr())
        for k in ("W", "vel", "mass", "total_area", "dt", "volume", "active", "left", "right",
                  "area", "ghost", "cf_off", "gather_off"):
            setattr(d, k, p(k))
        for k in ("dL", "dR", "rRL", "mc", "rcent"):
            setattr(d, k, p(k))
        d.cf_entry = p("cf_entry")
        d.egeo = ct.c_void_p(self.egeo.data_ptr())
        d.gather_entry = p("gather_entry")
        d.egeo_group = p("egeo_group") if self.grouped else None
        d.egeo_rows = self.egeo_rows if self.grouped else 0
        d.egeo_slot = p("egeo_slot") if self.sorted else None
        d.twin = p("twin") if self.use_twin else None
        d.closure = p("closure") if "closure" in a else None
        d.frec = None  # set once the records are built
        return d

    def build_entries(self, solver_handle, stream=None):
        """Per-entry face records of the gradient pass (once per mesh)."""
        # the face records first: the entry records are then built from them
        if self.frec is not None:
            Cp.check(Cp.lib().sfx_mesh_face_records(solver_handle, ct.byref(self.desc),
                                                    ct.c_void_p(self.frec.data_ptr()), stream),
                     "sfx_mesh_face_records")
            self.desc.frec = ct.c_void_p(self.frec.data_ptr())
        Cp.check(Cp.lib().sfx_mesh_entry_geometry(solver_handle, ct.byref(self.desc), self.E,
                                                  ct.c_void_p(self.egeo.data_ptr()), stream),
                 "sfx_mesh_entry_geometry")
        self.egeo_built = True

    # host copies in the component-major layout of the bundle / oracle
    def grad_soa(self):
        return self.grad.double().cpu().numpy().transpose(1, 2, 0).copy()

    def dQL_soa(self):
        return self.dQL.double().cpu().numpy().T.copy()

    def dQ_soa(self):
        return self.dQ.cpu().numpy().T.copy()

    def Q_soa(self):
        return self.Q.cpu().numpy().T.copy()

    def W_soa(self):
        return self.arr["W"].double().cpu().numpy().T.copy()
