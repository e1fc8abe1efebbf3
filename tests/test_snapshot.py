"""Snapshot I/O (GadgetSnapshotReader / GadgetSnapshotWriter): the native
HDF5 reader on the reference test suite's own initial-condition files
(written by the HDF5 library; copied into tests/golden), the writer's files
read back, and the format facts the bytes must show.  CPU only."""
import os

import numpy as np
import pytest

from shadowfax_amd import _capi as C
from shadowfax_amd import snapshot as S

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
COSMO = os.path.join(GOLD, "ic_cosmo3d.hdf5")
NBODY = os.path.join(GOLD, "ic_nbody3d.hdf5")


def test_reference_cosmo_ic_tree():
    with S.H5File(COSMO) as f:
        groups, attrs = f.list("/")
        assert groups == ["Header", "PartType1", "RuntimePars", "Units"] and attrs == []
        members, _ = f.list("/PartType1")
        assert members == ["Coordinates", "Masses", "ParticleIDs", "Velocities"]
        assert f.info("/PartType1/Coordinates") == (1, 4, (4096, 3))
        assert f.info("/PartType1/ParticleIDs") == (0, 4, (4096,))
        assert f.info("/Header", "NumPart_ThisFile") == (0, 4, (6,))
        assert f.exists("/Header", "Redshift") and not f.exists("/PartType0")


def test_reference_cosmo_ic_values():
    s = S.read_gadget_snapshot(COSMO)
    assert list(s["numpart"]) == [0, 4096, 0, 0, 0, 0] and "gas" not in s
    # 4096 = 16^3 particles of the same mass: the float32 Masses dataset is
    # MassTable[1] (a double attribute) rounded to float32
    with S.H5File(COSMO) as f:
        table = f.read("/Header", "MassTable")
        z = float(f.read("/Header", "Redshift"))
    dm = s["dm"]
    assert np.all(dm["masses"] == np.float32(table[1]))
    # Time is the scale factor 1 / (1 + z)
    assert s["time"] == pytest.approx(1.0 / (1.0 + z), rel=1e-12)
    assert np.array_equal(s["box"], [1e5] * 3)
    assert s["periodic"] is True and s["gravity"] is False
    # length unit: one kiloparsec in cm
    assert s["units"]["U_L"] == pytest.approx(3.08568025e21, rel=1e-12)
    assert np.all(dm["positions"] >= 0) and np.all(dm["positions"] <= 1e5)
    assert np.all(np.isfinite(dm["velocities"]))
    assert np.array_equal(np.sort(dm["ids"]), np.arange(dm["ids"].min(), dm["ids"].min() + 4096))


def test_reference_nbody_ic_values():
    s = S.read_gadget_snapshot(NBODY)
    n = 22 ** 3
    assert list(s["numpart"]) == [0, n, 0, 0, 0, 0]
    assert np.array_equal(s["box"], [40.0, 40.0, 40.0])
    # Time stored as a float here, as a double in the other file
    with S.H5File(NBODY) as f:
        assert f.info("/Header", "Time")[:2] == (1, 4)
        assert f.info("/PartType1/ParticleIDs")[:2] == (0, 8)
    assert s["time"] == 0.0
    # RuntimePars without PeriodicBoundariesOn; GravityOn present and 0
    assert s["periodic"] is False and s["gravity"] is False and s["hsoft"] == 0.03
    dm = s["dm"]
    assert dm["positions"].shape == (n, 3) and np.all(dm["positions"] >= 0)
    assert np.all(dm["positions"] <= 40.0) and np.all(dm["masses"] > 0)
    assert len(np.unique(dm["ids"])) == n


def gas_arrays(n, seed=5):
    rng = np.random.default_rng(seed)
    return dict(positions=rng.uniform(0, 2.0, (n, 3)), density=rng.uniform(0.5, 2.0, n),
                velocities=rng.normal(size=(n, 3)), pressure=rng.uniform(0.1, 3.0, n),
                ids=rng.permutation(n).astype(np.uint64) + 7)


def test_writer_round_trip(tmp_path):
    n = 500
    g = gas_arrays(n)
    paqs = {"Dye": np.linspace(0, 1, n), "Fe": np.full(n, 0.25), "Mg": np.zeros(n)}
    masses = np.linspace(1.0, 2.0, n)
    path = str(tmp_path / "gas.hdf5")
    S.write_gadget_snapshot(path, [2.0, 2.0, 2.0], 0.125, paqs=paqs, masses=masses,
                            periodic=True, gravity=True, hsoft=0.05, **g)
    s = S.read_gadget_snapshot(path)
    assert list(s["numpart"]) == [n, 0, 0, 0, 0, 0]
    assert s["time"] == 0.125 and np.array_equal(s["box"], [2.0] * 3)
    assert s["periodic"] is True and s["gravity"] is True and s["hsoft"] == 0.05
    assert s["units"]["U_L"] == 100.0  # SI: one metre in cm
    gas = s["gas"]
    assert np.array_equal(gas["positions"], g["positions"])
    assert np.array_equal(gas["density"], g["density"])
    # Velocities are stored as float (as the reference writes them)
    assert np.array_equal(gas["velocities"], g["velocities"].astype(np.float32))
    # pressure -> u = 1.5 p / rho -> 2 rho u / 3
    assert np.allclose(gas["pressure"], g["pressure"], rtol=4e-16, atol=0)
    assert np.array_equal(gas["ids"], g["ids"].astype(np.int64))
    assert set(gas["paqs"]) == set(paqs)
    for k, v in paqs.items():
        assert np.array_equal(gas["paqs"][k], v)
    sm = S.read_gadget_snapshot(path, read_mass=True)
    assert np.array_equal(sm["gas"]["masses"], masses)
    assert np.allclose(sm["gas"]["pressure"], 1.5 * g["pressure"] / g["density"], rtol=1e-15)


def test_written_bytes_follow_the_format(tmp_path):
    """Superblock version 0 with 8-byte offsets and lengths, the end-of-file
    address equal to the file's size, the root group's cached symbol table
    pointing at a B-tree node and a local heap."""
    path = str(tmp_path / "s.hdf5")
    S.write_gadget_snapshot(path, [1.0, 1.0], 0.0, **{k: v[:, :2] if v.ndim == 2 else v
                                                      for k, v in gas_arrays(10).items()})
    b = open(path, "rb").read()
    assert b[:8] == b"\x89HDF\r\n\x1a\n" and b[8] == 0 and b[13] == 8 and b[14] == 8
    u64 = lambda o: int.from_bytes(b[o:o + 8], "little")
    assert u64(24) == 0 and u64(32) == 2 ** 64 - 1 and u64(40) == len(b)
    assert int.from_bytes(b[72:76], "little") == 1
    assert b[u64(80):u64(80) + 4] == b"TREE" and b[u64(88):u64(88) + 4] == b"HEAP"
    # 2D positions are padded to 3 coordinates; BoxSize holds 3 values with
    # z = 0, as the reference writes it in 2D
    # (src/GadgetSnapshotWriter.cpp:186-195); every passive scalar is written,
    # zeros where none is given (:376-383)
    s = S.read_gadget_snapshot(path, ndim=2)
    assert s["gas"]["positions"].shape == (10, 3) and np.all(s["gas"]["positions"][:, 2] == 0)
    assert np.array_equal(s["box"], [1.0, 1.0, 0.0])
    assert sorted(s["gas"]["paqs"]) == sorted(S.PAQ_NAMES)
    assert all(np.all(v == 0) for v in s["gas"]["paqs"].values())
    with pytest.raises(ValueError):
        S.write_gadget_snapshot(path, [1.0, 1.0], 0.0, paqs={"Zn": np.zeros(10)},
                                **{k: v[:, :2] if v.ndim == 2 else v
                                   for k, v in gas_arrays(10).items()})


def test_group_of_many_members(tmp_path):
    """40 datasets: five symbol table nodes under one B-tree node, found by
    name in sorted order."""
    path = str(tmp_path / "many.hdf5")
    names = ["d%02d" % k for k in range(40)][::-1]
    with S.H5Writer() as w:
        w.group("/G")
        for k, nm in enumerate(names):
            w.dataset("/G/" + nm, np.arange(k + 1, dtype=np.int32) * 3, np.int32)
        w.attr("/G", "scale", np.array(2.5), np.float64)
        w.save(path)
    with S.H5File(path) as f:
        members, attrs = f.list("/G")
        assert members == sorted(names) and attrs == ["scale"]
        for k, nm in enumerate(names):
            assert np.array_equal(f.read("/G/" + nm), np.arange(k + 1) * 3)
        assert f.read("/G", "scale").shape == () and float(f.read("/G", "scale")) == 2.5


def test_reader_errors(tmp_path):
    bad = tmp_path / "bad.hdf5"
    bad.write_bytes(b"not an hdf5 file" * 8)
    with pytest.raises(C.SfxError, match="signature"):
        S.H5File(str(bad))
    b = bytearray(open(COSMO, "rb").read())
    b[8] = 2  # a superblock version this reader does not parse
    v2 = tmp_path / "v2.hdf5"
    v2.write_bytes(bytes(b))
    with pytest.raises(C.SfxError, match="superblock version"):
        S.H5File(str(v2))
    with S.H5File(COSMO) as f:
        with pytest.raises(C.SfxError, match="no object"):
            f.read("/PartType0/Coordinates")
        with pytest.raises(C.SfxError, match="no attribute"):
            f.read("/Header", "Nope")
        with pytest.raises(C.SfxError, match="is a group"):
            f.read("/Header")
    with S.H5Writer() as w:
        with pytest.raises(C.SfxError, match="no group"):
            w.dataset("/missing/x", np.zeros(3), np.float64)


def test_ic_to_mesh_state(tmp_path):
    """A 2D Kelvin-Helmholtz state written as a gas snapshot and read back
    gives the mesh builder the same generators and the flux pass the same
    primitives (velocities through the format's float storage)."""
    from shadowfax_amd import mesh as M
    gen = M.hilbert_sorted(M.jittered_grid(16, 0.25, seed=3))  # [2][C]
    m = M.voronoi_periodic(gen)
    W = M.kh_ics(m)  # [4][C]: rho, vx, vy, p
    path = str(tmp_path / "kh.hdf5")
    S.write_gadget_snapshot(path, [1.0, 1.0], 0.0, positions=gen.T, density=W[0],
                            velocities=W[1:3].T, pressure=W[3])
    s = S.read_gadget_snapshot(path, ndim=2)
    assert np.array_equal(s["gas"]["positions"][:, :2], gen.T)
    Wr = S.gas_primitives(s, 2)
    assert np.array_equal(Wr[:, 0], W[0])
    assert np.array_equal(Wr[:, 1:3], W[1:3].T.astype(np.float32))
    assert np.allclose(Wr[:, 3], W[3], rtol=4e-16, atol=0)
    m2 = M.voronoi_periodic(np.ascontiguousarray(s["gas"]["positions"][:, :2].T))
    assert m2["F"] == m["F"] and np.array_equal(m2["volume"], m["volume"])


# ---------------------------------------------------------- Shadowfax format

def shadowfax_gas(n, ndim, seed=11):
    rng = np.random.default_rng(seed)
    return dict(positions=rng.uniform(0, 3.0, (n, ndim)), density=rng.uniform(0.5, 2.0, n),
                velocities=rng.normal(size=(n, ndim)), pressure=rng.uniform(0.1, 3.0, n),
                ids=rng.permutation(n) + 100, timestep=rng.uniform(0, 1e-3, n),
                acceleration=rng.normal(size=(n, ndim)), masses=rng.uniform(1, 2, n))


def test_shadowfax_round_trip_3d(tmp_path):
    n, m = 300, 40
    gas = shadowfax_gas(n, 3)
    rng = np.random.default_rng(2)
    dm = dict(positions=rng.uniform(0, 3, (m, 3)), velocities=rng.normal(size=(m, 3)),
              masses=rng.uniform(1, 2, m), ids=np.arange(m) + 5000)
    path = str(tmp_path / "sfx3.hdf5")
    S.write_shadowfax_snapshot(path, 3, [3.0, 3.0, 3.0], 0.25, gas=gas, dm=dm, gravity=True,
                               hsoft=0.07, origin=[-1.0, 0.0, 0.5])
    s = S.read_shadowfax_snapshot(path, ndim=3, gamma=5.0 / 3.0)
    assert list(s["numpart"]) == [n, m] and s["ndim"] == 3 and s["time"] == 0.25
    assert np.array_equal(s["box"], [-1.0, 0.0, 0.5, 3.0, 3.0, 3.0])
    assert s["periodic"] and s["second_order"] and not s["static"]
    assert not s["global_timestep"] and s["gravity"] and s["hsoft"] == 0.07
    g = s["gas"]
    for k in ("positions", "density", "velocities", "pressure"):
        assert np.array_equal(g[k], gas[k]), k
    assert np.array_equal(g["ids"], gas["ids"])
    d = s["dm"]
    for k in ("positions", "velocities", "masses", "ids"):
        assert np.array_equal(d[k], dm[k]), k
    assert s["units"] == {}
    # the columns the reader does not use are stored as given
    with S.H5File(path) as f:
        assert np.array_equal(f.read("/cells/timestep"), gas["timestep"])
        assert np.array_equal(f.read("/cells/acceleration_z"), gas["acceleration"][:, 2])
        assert np.array_equal(f.read("/cells/mass"), gas["masses"])
        assert np.array_equal(f.read("/DM/process"), np.zeros(m))


def test_shadowfax_format_facts_2d(tmp_path):
    """Header attribute types and shapes (booleans and counts as uint32,
    scalars with a scalar dataspace, box as origin then sides), the
    datasets of each block, z = 0 in 2D, hsoft only with DM and gravity, and
    the reader's ndim and gamma checks."""
    n = 50
    path = str(tmp_path / "sfx2.hdf5")
    S.write_shadowfax_snapshot(path, 2, [1.0, 2.0], 1.5, gas=shadowfax_gas(n, 2), gravity=True,
                               gamma=1.4)
    with S.H5File(path) as f:
        groups, _ = f.list("/")
        assert groups == ["Header", "cells", "grid"]
        _, attrs = f.list("/Header")
        assert sorted(attrs) == sorted(["npart", "ndim", "box", "time", "periodic",
                                        "second order", "static", "global timestep",
                                        "adiabatic index", "gravity"])
        assert f.info("/Header", "npart") == (0, 4, (2,))
        assert f.info("/Header", "ndim") == (0, 4, ())
        assert f.info("/Header", "periodic") == (0, 4, ())
        assert f.info("/Header", "box") == (1, 8, (4,))
        assert f.info("/Header", "time") == (1, 8, ())
        assert f.list("/cells")[0] == sorted(["id", "density", "velocity_x", "velocity_y",
                                              "pressure", "timestep", "acceleration_x",
                                              "acceleration_y", "mass"])
        assert f.list("/grid")[0] == ["x", "y", "z"]
        assert f.info("/cells/density") == (1, 8, (n,))
        assert np.all(f.read("/grid/z") == 0.0)
    s = S.read_shadowfax_snapshot(path, ndim=2)
    assert np.array_equal(s["box"], [0.0, 0.0, 1.0, 2.0]) and s["gamma"] == 1.4
    assert s["gas"]["positions"].shape == (n, 2) and s["hsoft"] == 0.0 and "dm" not in s
    with pytest.raises(ValueError, match="ndim=2.*ndim=3"):
        S.read_shadowfax_snapshot(path, ndim=3)
    with pytest.raises(ValueError, match="gamma"):
        S.read_shadowfax_snapshot(path, ndim=2, gamma=5.0 / 3.0)


def test_gadget_ic_to_shadowfax_format(tmp_path):
    """The reference's cosmological IC (Gadget format, DM only) rewritten in
    the Shadowfax format and read back: the same particles."""
    g = S.read_gadget_snapshot(COSMO)
    dm = g["dm"]
    path = str(tmp_path / "cosmo_sfx.hdf5")
    S.write_shadowfax_snapshot(path, 3, g["box"], g["time"], dm=dm, periodic=g["periodic"])
    s = S.read_shadowfax_snapshot(path)
    assert list(s["numpart"]) == [0, 4096] and "gas" not in s
    assert s["time"] == g["time"] and np.array_equal(s["box"][3:], g["box"])
    assert s["periodic"] and not s["gravity"] and s["hsoft"] == 0.0
    for k in ("positions", "velocities", "masses", "ids"):
        assert np.array_equal(s["dm"][k], dm[k].astype(s["dm"][k].dtype)), k
