"""HDF5 snapshots and initial conditions in the Gadget and Shadowfax formats.

read_gadget_snapshot mirrors GadgetSnapshotReader::read_snapshot
(src/GadgetSnapshotReader.cpp:79-560, GadgetSnapshotReader.hpp:49-58) and
write_gadget_snapshot the files GadgetSnapshotWriter writes
(src/GadgetSnapshotWriter.cpp:220-500); read_shadowfax_snapshot and
write_shadowfax_snapshot do the same for ShadowfaxSnapshotReader and
ShadowfaxSnapshotWriter (src/ShadowfaxSnapshotReader.cpp,
src/ShadowfaxSnapshotWriter.cpp).  The HDF5 structures are parsed and
written by the native sfx_h5 / sfx_h5w entry points (include/shadowfax_amd.h);
there is no dependency on an HDF5 library.

Values are returned in the file's own units; the file's /Units block (cgs
factors), when present, is returned alongside (the reference converts to its
run's units through UnitConverter, which is outside this package).
"""
import ctypes as ct

import numpy as np

from . import _capi as C

vp = ct.c_void_p
i64 = ct.c_int64
_ip = ct.POINTER(ct.c_int)
_i64p = ct.POINTER(ct.c_int64)

SFX_H5_UINT, SFX_H5_FLOAT, SFX_H5_INT = 0, 1, 2
PAQ_NAMES = ("Dye", "Fe", "Mg")  # src/StateVector.hpp:39-40
UNIT_NAMES = {"U_I": "Unit current in cgs (U_I)", "U_L": "Unit length in cgs (U_L)",
              "U_M": "Unit mass in cgs (U_M)", "U_T": "Unit temperature in cgs (U_T)",
              "U_t": "Unit time in cgs (U_t)"}


def _lib():
    L = C.lib()
    if getattr(L, "_h5_sigs", False):
        return L
    L.sfx_h5_open.argtypes = [ct.c_char_p, ct.POINTER(vp)]
    L.sfx_h5_close.argtypes = [vp]
    L.sfx_h5_exists.argtypes = [vp, ct.c_char_p, ct.c_char_p]
    L.sfx_h5_info.argtypes = [vp, ct.c_char_p, ct.c_char_p, _ip, _ip, _ip, _i64p]
    L.sfx_h5_read_f64.argtypes = [vp, ct.c_char_p, ct.c_char_p, i64, vp]
    L.sfx_h5_read_i64.argtypes = [vp, ct.c_char_p, ct.c_char_p, i64, vp]
    L.sfx_h5_list.argtypes = [vp, ct.c_char_p, ct.c_char_p, i64]
    L.sfx_h5w_create.argtypes = [ct.POINTER(vp)]
    L.sfx_h5w_destroy.argtypes = [vp]
    L.sfx_h5w_group.argtypes = [vp, ct.c_char_p]
    L.sfx_h5w_dataset.argtypes = [vp, ct.c_char_p, ct.c_int, ct.c_int, ct.c_int, _i64p, vp]
    L.sfx_h5w_attr.argtypes = [vp, ct.c_char_p, ct.c_char_p, ct.c_int, ct.c_int, ct.c_int,
                               _i64p, vp]
    L.sfx_h5w_save.argtypes = [vp, ct.c_char_p]
    L._h5_sigs = True
    return L


def _b(s):
    return None if s is None else s.encode()


class H5File:
    """A read-only HDF5 file (superblock version 0, contiguous or compact
    datasets).  Paths are absolute: "/PartType0/Coordinates"."""

    def __init__(self, path):
        self.L = _lib()
        self.h = vp()
        C.check(self.L.sfx_h5_open(str(path).encode(), ct.byref(self.h)), "sfx_h5_open")

    def close(self):
        if self.h:
            self.L.sfx_h5_close(self.h)
            self.h = vp()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def exists(self, path, attr=None):
        rc = self.L.sfx_h5_exists(self.h, _b(path), _b(attr))
        if rc < 0:
            C.check(rc, "sfx_h5_exists")
        return rc == 1

    def info(self, path, attr=None):
        """(type class, type size, shape): class 0 integer, 1 floating point,
        8 enumeration, -1 other."""
        cls, size, rank = ct.c_int(), ct.c_int(), ct.c_int()
        dims = (ct.c_int64 * 8)()
        C.check(self.L.sfx_h5_info(self.h, _b(path), _b(attr), ct.byref(cls), ct.byref(size),
                                   ct.byref(rank), dims), "sfx_h5_info")
        return cls.value, size.value, tuple(dims[d] for d in range(rank.value))

    def read(self, path, attr=None, integer=None):
        """The values as float64 (or int64 for integer types, or when integer
        is True), in the stored shape."""
        cls, _, shape = self.info(path, attr)
        n = int(np.prod(shape)) if shape else 1
        as_int = (cls in (0, 8)) if integer is None else integer
        out = np.empty(n, np.int64 if as_int else np.float64)
        fn = self.L.sfx_h5_read_i64 if as_int else self.L.sfx_h5_read_f64
        C.check(fn(self.h, _b(path), _b(attr), n, out.ctypes.data), "sfx_h5_read")
        return out.reshape(shape)

    def list(self, path="/"):
        """(members, attribute names) of the object at path."""
        cap = 1 << 12
        while True:
            buf = ct.create_string_buffer(cap)
            rc = self.L.sfx_h5_list(self.h, _b(path), buf, cap)
            if rc == 0:
                break
            if rc < 0:
                C.check(rc, "sfx_h5_list")
            cap = rc
        names = [s for s in buf.value.decode().split("\n") if s]
        return ([s for s in names if not s.startswith("@")],
                [s[1:] for s in names if s.startswith("@")])


class H5Writer:
    """Builds an HDF5 file in memory; save() writes it."""

    def __init__(self):
        self.L = _lib()
        self.h = vp()
        C.check(self.L.sfx_h5w_create(ct.byref(self.h)), "sfx_h5w_create")

    def close(self):
        if self.h:
            self.L.sfx_h5w_destroy(self.h)
            self.h = vp()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def group(self, path):
        C.check(self.L.sfx_h5w_group(self.h, _b(path)), "sfx_h5w_group")

    @staticmethod
    def _typed(values, dtype):
        a = np.array(values, dtype=dtype, order="C")  # keeps a 0-d scalar 0-d
        kind = {"f": SFX_H5_FLOAT, "u": SFX_H5_UINT, "i": SFX_H5_INT}[a.dtype.kind]
        return a, kind, a.dtype.itemsize

    def dataset(self, path, values, dtype):
        a, kind, size = self._typed(values, dtype)
        dims = (ct.c_int64 * max(a.ndim, 1))(*a.shape)
        C.check(self.L.sfx_h5w_dataset(self.h, _b(path), kind, size, a.ndim, dims,
                                       a.ctypes.data if a.size else None), "sfx_h5w_dataset")

    def attr(self, path, name, values, dtype):
        a, kind, size = self._typed(values, dtype)
        dims = (ct.c_int64 * max(a.ndim, 1))(*a.shape)
        C.check(self.L.sfx_h5w_attr(self.h, _b(path), _b(name), kind, size, a.ndim, dims,
                                    a.ctypes.data), "sfx_h5w_attr")

    def save(self, path):
        C.check(self.L.sfx_h5w_save(self.h, str(path).encode()), "sfx_h5w_save")


def _vector(f, path, n, ndim=3):
    """An [n][3] (or [n][1]) dataset as float64 [n][3] / [n]."""
    a = f.read(path, integer=False)
    return a.reshape(n, -1)


def _scalar(f, path, n):
    return f.read(path, integer=False).reshape(n)


def keep_inside(pos, box, ndim):
    """GadgetSnapshotReader::keep_inside (GadgetSnapshotReader.hpp:49-58):
    one box length added below the origin, subtracted past the far side."""
    pos = pos.copy()
    for i in range(ndim):
        pos[:, i] = np.where(pos[:, i] < 0.0, pos[:, i] + box[i], pos[:, i])
        pos[:, i] = np.where(pos[:, i] > box[i], pos[:, i] - box[i], pos[:, i])
    return pos


def read_gadget_snapshot(path, ndim=3, read_mass=False):
    """GadgetSnapshotReader::read_snapshot (src/GadgetSnapshotReader.cpp:79-560)
    on one rank: the header (box from BoxSize, padded to 3 values; particle
    counts NumPart_ThisFile[0, 1, 4]; Time as float or double), the run-time
    flags (/RuntimePars: PeriodicBoundariesOn; GravityOn, SofteningLength
    defaulting to false and 0.03; without the group: not periodic, no
    gravity), and the particles:
      gas   (PartType0): positions (kept inside the box), velocities, density
            (or masses with read_mass), pressure = u 2 rho / 3 (gamma 5/3, as
            the reference; with read_mass the internal energy u itself), ids,
            the passive scalars present (Dye, Fe, Mg);
      dm    (PartType1) and stars (PartType4): positions, velocities, masses,
            ids.
    Values in the file's units; "units" holds the /Units cgs factors or None
    (the reference then assumes SI)."""
    out = {}
    with H5File(path) as f:
        box = list(np.atleast_1d(f.read("/Header", "BoxSize", integer=False)))
        while len(box) < 3:
            box.append(box[0])
        out["box"] = np.array(box[:3], np.float64)
        numpart = np.atleast_1d(f.read("/Header", "NumPart_ThisFile", integer=True))
        out["numpart"] = numpart
        out["time"] = float(np.atleast_1d(f.read("/Header", "Time", integer=False))[0])
        if f.exists("/Units"):
            out["units"] = {k: float(np.atleast_1d(f.read("/Units", v, integer=False))[0])
                            for k, v in UNIT_NAMES.items() if f.exists("/Units", v)}
        else:
            out["units"] = None
        if f.exists("/RuntimePars"):
            rp = "/RuntimePars"
            out["periodic"] = (bool(np.atleast_1d(f.read(rp, "PeriodicBoundariesOn",
                                                         integer=True))[0])
                               if f.exists(rp, "PeriodicBoundariesOn") else False)
            out["gravity"] = (bool(np.atleast_1d(f.read(rp, "GravityOn", integer=True))[0])
                              if f.exists(rp, "GravityOn") else False)
            out["hsoft"] = (float(np.atleast_1d(f.read(rp, "SofteningLength",
                                                       integer=False))[0])
                            if f.exists(rp, "SofteningLength") else 0.03)
        else:
            out["periodic"], out["gravity"], out["hsoft"] = False, False, 0.03
        ngas = int(numpart[0])
        if ngas:
            g = "/PartType0/"
            gas = {"positions": keep_inside(_vector(f, g + "Coordinates", ngas), out["box"],
                                            ndim),
                   "velocities": _vector(f, g + "Velocities", ngas)}
            u = _scalar(f, g + "InternalEnergy", ngas)
            if read_mass:
                gas["masses"] = _scalar(f, g + "Masses", ngas)
                gas["pressure"] = u
            else:
                rho = _scalar(f, g + "Density", ngas)
                gas["density"] = rho
                gas["pressure"] = u * 2.0 * rho / 3.0
            gas["ids"] = f.read(g + "ParticleIDs", integer=True).reshape(ngas)
            gas["paqs"] = {p: _scalar(f, g + p, ngas) for p in PAQ_NAMES if f.exists(g + p)}
            out["gas"] = gas
        for key, grp, idx in (("dm", "/PartType1/", 1), ("stars", "/PartType4/", 4)):
            n = int(numpart[idx]) if len(numpart) > idx else 0
            if not n:
                continue
            out[key] = {"positions": keep_inside(_vector(f, grp + "Coordinates", n),
                                                 out["box"], ndim),
                        "velocities": _vector(f, grp + "Velocities", n),
                        "masses": _scalar(f, grp + "Masses", n),
                        "ids": f.read(grp + "ParticleIDs", integer=True).reshape(n)}
    return out


def gas_primitives(snap, ndim):
    """The primitive rows [rho, v (ndim), p] of the gas, in the reference's
    StateVector order (GadgetSnapshotReader.cpp:395-404)."""
    g = snap["gas"]
    return np.column_stack([g["density"], g["velocities"][:, :ndim], g["pressure"]])


def write_gadget_snapshot(path, box, time, positions, density, velocities, pressure, ids=None,
                          masses=None, paqs=None, periodic=True, gravity=False, hsoft=0.03,
                          units=None):
    """The gas part of GadgetSnapshotWriter's output
    (src/GadgetSnapshotWriter.cpp:224-400): /Header (BoxSize, Flag_Entropy_ICs,
    MassTable, NumFilesPerSnapshot, NumPart_ThisFile, NumPart_Total,
    NumPart_Total_HighWord, Time as float), /RuntimePars
    (PeriodicBoundariesOn, GravityOn, SofteningLength with gravity), /Units
    (cgs factors; SI when not given) and /PartType0 with Coordinates
    (double, [n][3]), Density (double), Velocities (float, [n][3]),
    InternalEnergy = 1.5 p / rho (0 where rho = 0), ParticleIDs (uint64),
    every passive scalar (zeros where not given) and, with masses, Masses.
    BoxSize holds three values (z = 0 in 2D), as the reference writes it."""
    pos = np.asarray(positions, np.float64)
    n, nd = pos.shape
    p3 = np.zeros((n, 3))
    p3[:, :nd] = pos
    v3 = np.zeros((n, 3))
    vel = np.asarray(velocities, np.float64)
    v3[:, :vel.shape[1]] = vel
    rho = np.asarray(density, np.float64)
    prs = np.asarray(pressure, np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        u = np.where(rho != 0.0, 1.5 * prs / np.where(rho != 0.0, rho, 1.0), 0.0)
    ids = np.arange(n, dtype=np.uint64) if ids is None else np.asarray(ids, np.uint64)
    units = units or {"U_I": 1.0, "U_L": 100.0, "U_M": 1000.0, "U_T": 1.0, "U_t": 1.0}
    num = np.zeros(6, np.uint32)
    num[0] = n
    with H5Writer() as w:
        w.group("/Header")
        # three values in 2D too, z = 0 (src/GadgetSnapshotWriter.cpp:186-195)
        box3 = np.zeros(3)
        box3[:nd] = np.asarray(box, np.float64)[:nd]
        w.attr("/Header", "BoxSize", box3, np.float64)
        w.attr("/Header", "Flag_Entropy_ICs", np.zeros(6, np.uint32), np.uint32)
        w.attr("/Header", "MassTable", np.zeros(6), np.float64)
        w.attr("/Header", "NumFilesPerSnapshot", np.array([1], np.int32), np.int32)
        w.attr("/Header", "NumPart_ThisFile", num, np.uint32)
        w.attr("/Header", "NumPart_Total", num, np.uint32)
        w.attr("/Header", "NumPart_Total_HighWord", np.zeros(6, np.uint32), np.uint32)
        w.attr("/Header", "Time", np.array([time], np.float32), np.float32)
        w.group("/RuntimePars")
        w.attr("/RuntimePars", "PeriodicBoundariesOn", np.array([int(periodic)], np.uint32),
               np.uint32)
        w.attr("/RuntimePars", "GravityOn", np.array([int(gravity)], np.uint32), np.uint32)
        if gravity:
            w.attr("/RuntimePars", "SofteningLength", np.array([hsoft]), np.float64)
        w.group("/Units")
        for k, name in UNIT_NAMES.items():
            w.attr("/Units", name, np.array([units[k]]), np.float64)
        w.group("/PartType0")
        w.dataset("/PartType0/Coordinates", p3, np.float64)
        w.dataset("/PartType0/Density", rho, np.float64)
        w.dataset("/PartType0/Velocities", v3, np.float32)
        w.dataset("/PartType0/InternalEnergy", u, np.float64)
        w.dataset("/PartType0/ParticleIDs", ids, np.uint64)
        # every passive scalar, zeros where none is given
        # (src/GadgetSnapshotWriter.cpp:376-383)
        paqs = dict(paqs or {})
        unknown = set(paqs) - set(PAQ_NAMES)
        if unknown:
            raise ValueError("unknown passive scalars %s (known: %s)" % (sorted(unknown),
                                                                         PAQ_NAMES))
        for name in PAQ_NAMES:
            vals = paqs.get(name)
            w.dataset("/PartType0/" + name,
                      np.zeros(n) if vals is None else np.asarray(vals, np.float64), np.float64)
        if masses is not None:
            w.dataset("/PartType0/Masses", np.asarray(masses, np.float64), np.float64)
        w.save(path)


# ------------------------------------------------------ Shadowfax format --

def _cells_names(ndim):
    """The datasets of the /cells block in the order the reference writes
    them (src/ShadowfaxSnapshotWriter.cpp:101-132)."""
    ax = "xyz"[:ndim]
    return (["id", "density"] + ["velocity_" + a for a in ax] + ["pressure", "timestep"] +
            ["acceleration_" + a for a in ax] + ["mass"])


DM_NAMES = ("id", "mass", "x", "y", "z", "vx", "vy", "vz", "process")


def write_shadowfax_snapshot(path, ndim, box, time, gas=None, dm=None, origin=None,
                             periodic=True, second_order=True, static=False,
                             global_timestep=False, gamma=5.0 / 3.0, gravity=False, hsoft=0.0):
    """The file ShadowfaxSnapshotWriter::write_snapshot writes
    (src/ShadowfaxSnapshotWriter.cpp:59-277, through FileOutput::write_header
    and FileOutput::write, src/io/FileOutput.cpp:73-254):
      /Header attributes: npart (uint32 [2]: gas, DM), ndim (uint32), box
        (double [2 ndim]: origin, then side lengths), time (double),
        periodic, "second order", static, "global timestep", gravity
        (booleans stored as uint32), "adiabatic index" (double), and hsoft
        (double) when there are DM particles and gravity is on;
      /cells: id, density, velocity_x..., pressure, timestep,
        acceleration_x..., mass; /grid: x, y, z (z = 0 in 2D);
      /DM: id, mass, x, y, z, vx, vy, vz, process (rank, 0 here);
    every dataset a one-dimensional double array.  gas: positions [n][ndim],
    density, velocities [n][ndim], pressure, and optionally ids (default
    0..n-1), timestep, acceleration [n][ndim], masses (default 0).  dm:
    positions, velocities, masses, ids.  Values are written as given: the
    reference converts to its output units and tags each dataset with
    unit_quantity / unit_name (variable-length strings) and unit_SI_value;
    those attributes are not written, and the reference reader, finding no
    unit_quantity, takes the values in its own units
    (src/io/FileInput.cpp:100-113)."""
    if ndim not in (2, 3):
        raise ValueError("ndim must be 2 or 3")
    sides = np.asarray(box, np.float64).reshape(-1)[:ndim]
    org = np.zeros(ndim) if origin is None else np.asarray(origin, np.float64)[:ndim]
    ngas = 0 if gas is None else len(gas["density"])
    ndm = 0 if dm is None else len(dm["masses"])
    b = lambda v: np.array(int(bool(v)), np.uint32)
    with H5Writer() as w:
        w.group("/Header")
        w.attr("/Header", "npart", np.array([ngas, ndm], np.uint32), np.uint32)
        w.attr("/Header", "ndim", np.array(ndim, np.uint32), np.uint32)
        w.attr("/Header", "box", np.concatenate([org, sides]), np.float64)
        w.attr("/Header", "time", np.array(float(time)), np.float64)
        w.attr("/Header", "periodic", b(periodic), np.uint32)
        w.attr("/Header", "second order", b(second_order), np.uint32)
        w.attr("/Header", "static", b(static), np.uint32)
        w.attr("/Header", "global timestep", b(global_timestep), np.uint32)
        w.attr("/Header", "adiabatic index", np.array(float(gamma)), np.float64)
        w.attr("/Header", "gravity", b(gravity), np.uint32)
        if ndm and gravity:
            w.attr("/Header", "hsoft", np.array(float(hsoft)), np.float64)
        if ngas:
            pos = np.asarray(gas["positions"], np.float64).reshape(ngas, -1)
            vel = np.asarray(gas["velocities"], np.float64).reshape(ngas, -1)
            acc = np.asarray(gas.get("acceleration", np.zeros((ngas, ndim))),
                             np.float64).reshape(ngas, -1)
            cols = [np.asarray(gas.get("ids", np.arange(ngas)), np.float64),
                    np.asarray(gas["density"], np.float64)]
            cols += [vel[:, d] for d in range(ndim)]
            cols += [np.asarray(gas["pressure"], np.float64),
                     np.asarray(gas.get("timestep", np.zeros(ngas)), np.float64)]
            cols += [acc[:, d] for d in range(ndim)]
            cols.append(np.asarray(gas.get("masses", np.zeros(ngas)), np.float64))
            w.group("/cells")
            for name, col in zip(_cells_names(ndim), cols):
                w.dataset("/cells/" + name, np.ascontiguousarray(col), np.float64)
            w.group("/grid")
            for d, name in enumerate("xyz"):
                col = pos[:, d] if d < ndim else np.zeros(ngas)
                w.dataset("/grid/" + name, np.ascontiguousarray(col), np.float64)
        if ndm:
            pos = np.asarray(dm["positions"], np.float64).reshape(ndm, -1)
            vel = np.asarray(dm["velocities"], np.float64).reshape(ndm, -1)
            pad = lambda a, d: a[:, d] if d < a.shape[1] else np.zeros(ndm)
            cols = [np.asarray(dm.get("ids", np.arange(ndm)), np.float64),
                    np.asarray(dm["masses"], np.float64)]
            cols += [pad(pos, d) for d in range(3)] + [pad(vel, d) for d in range(3)]
            cols.append(np.zeros(ndm))
            w.group("/DM")
            for name, col in zip(DM_NAMES, cols):
                w.dataset("/DM/" + name, np.ascontiguousarray(col), np.float64)
        w.save(path)


def read_shadowfax_snapshot(path, ndim=3, gamma=None):
    """ShadowfaxSnapshotReader::read_snapshot (src/ShadowfaxSnapshotReader.cpp:
    79-199) with FileInput::read_header (src/io/FileInput.cpp:163-199) on
    one rank.  The header must have ndim equal to the ndim asked for, and,
    when gamma is given, that adiabatic index (Header::check_makeflags,
    src/io/Header.cpp:277-317, stops the run otherwise); hsoft is read only
    with DM particles and gravity (else 0).  Gas: positions from /grid (x, y
    and, in 3D, z), density, velocities, pressure and ids from /cells (ids
    stored as doubles); DM: ids, masses, positions, velocities.  Values as
    stored; "units" maps each dataset that carries unit_SI_value to it (the
    reference converts such datasets to its run's units)."""
    out = {}
    with H5File(path) as f:
        hd = "/Header"
        rd = lambda name, integer: np.atleast_1d(f.read(hd, name, integer=integer))
        fdim = int(rd("ndim", True)[0])
        if fdim != ndim:
            raise ValueError("dimensions of the snapshot (ndim=%d) and of the run (ndim=%d) "
                             "do not match" % (fdim, ndim))
        out["ndim"] = fdim
        out["numpart"] = rd("npart", True)
        out["box"] = rd("box", False)
        out["time"] = float(rd("time", False)[0])
        for key, name in (("periodic", "periodic"), ("second_order", "second order"),
                          ("static", "static"), ("global_timestep", "global timestep"),
                          ("gravity", "gravity")):
            out[key] = bool(rd(name, True)[0])
        out["gamma"] = float(rd("adiabatic index", False)[0])
        if gamma is not None and out["gamma"] != gamma:
            raise ValueError("the snapshot has gamma=%r, the run gamma=%r" % (out["gamma"], gamma))
        ngas = int(out["numpart"][0])
        ndm = int(out["numpart"][1]) if len(out["numpart"]) > 1 else 0
        out["hsoft"] = (float(rd("hsoft", False)[0]) if ndm and out["gravity"] else 0.0)
        units = {}

        def col(group, name, n):
            p = "/%s/%s" % (group, name)
            if f.exists(p, "unit_SI_value"):
                units[p] = float(np.atleast_1d(f.read(p, "unit_SI_value", integer=False))[0])
            return f.read(p, integer=False).reshape(-1)[:n]

        ax = "xyz"[:ndim]
        if ngas:
            out["gas"] = {
                "positions": np.column_stack([col("grid", a, ngas) for a in ax]),
                "density": col("cells", "density", ngas),
                "velocities": np.column_stack([col("cells", "velocity_" + a, ngas) for a in ax]),
                "pressure": col("cells", "pressure", ngas),
                "ids": col("cells", "id", ngas).astype(np.int64)}
        if ndm:
            out["dm"] = {"ids": col("DM", "id", ndm).astype(np.int64),
                         "masses": col("DM", "mass", ndm),
                         "positions": np.column_stack([col("DM", a, ndm) for a in ax]),
                         "velocities": np.column_stack([col("DM", "v" + a, ndm) for a in ax])}
        out["units"] = units
    return out
