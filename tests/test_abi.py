"""CPU checks of the C ABI library: it loads (no GPU needed to dlopen it),
exports every function declared in include/shadowfax_amd.h, and rejects bad
arguments with error codes instead of crashing.  No compute calls."""
import ctypes as ct
import os
import subprocess

import pytest

from shadowfax_amd import _capi as C


def test_library_exports_every_declared_symbol():
    L = C.lib()
    declared = C.declared_symbols()
    assert len(declared) >= 10
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing


def test_nm_shows_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", C.LIBPATH], capture_output=True,
                         text=True, check=True).stdout
    names = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for s in C.declared_symbols():
        assert s in names, s


def test_version_and_error_string():
    L = C.lib()
    assert b"gfx950" in L.sfx_version()
    assert isinstance(L.sfx_last_error(), bytes)


def test_create_rejects_bad_arguments():
    L = C.lib()
    h = ct.c_void_p()
    assert L.sfx_solver_create(7, 3, 5 / 3, 1e-8, -5.0, 0, ct.byref(h)) == -1
    assert b"kind" in L.sfx_last_error()
    assert L.sfx_solver_create(0, 4, 5 / 3, 1e-8, -5.0, 0, ct.byref(h)) == -1
    assert L.sfx_solver_create(0, 3, 0.9, 1e-8, -5.0, 0, ct.byref(h)) == -1
    assert L.sfx_solver_create(0, 3, 5 / 3, 1e-8, -5.0, 9, ct.byref(h)) == -1
    assert L.sfx_solver_create_by_name(b"TRRS", 3, 5 / 3, 1e-8, -5.0, 0, ct.byref(h)) == -4
    assert L.sfx_solver_create_by_name(b"Bogus", 3, 5 / 3, 1e-8, -5.0, 0, ct.byref(h)) == -1
    # NULL handles are rejected, not dereferenced
    assert L.sfx_solve_for_flux_batch(None, 10, None, None, None, None, None, None, None,
                                      None) == -1
    assert L.sfx_solver_destroy(None) == 0


def test_header_compiles_as_c():
    """include/shadowfax_amd.h is plain C (no C++ or torch types)."""
    root = os.path.dirname(os.path.dirname(C.LIBPATH))
    src = '#include "shadowfax_amd.h"\nint main(void){sfx_stats s; (void)s; return 0;}\n'
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I",
                        os.path.join(root, "include"), "-x", "c", "-", "-o", "/dev/null"],
                       input=src, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_struct_layouts_match_ctypes(tmp_path):
    """The ctypes mirrors of sfx_mesh, sfx_host_mesh and sfx_timeline have
    the C layout (size and every field offset), and the oracle's timeline
    matches the ABI's."""
    import ctypes as ct

    from shadowfax_amd import mesh as M

    root = os.path.dirname(os.path.dirname(C.LIBPATH))
    lines = []
    structs = (("sfx_mesh", C.MeshDesc), ("sfx_host_mesh", M.HostMesh),
               ("sfx_timeline", C.Timeline))
    for cname, py in structs:
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in py._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    src = ('#include <stddef.h>\n#include <stdio.h>\n#include "shadowfax_amd.h"\n'
           "int main(void){%s return 0;}\n" % "\n".join(lines))
    exe = str(tmp_path / "layout")
    r = subprocess.run(["gcc", "-std=c99", "-I", os.path.join(root, "include"), "-x", "c", "-",
                        "-o", exe], input=src, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for cname, py in structs:
        assert got[(cname, "size")] == ct.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)
    import oracle as O
    assert O.Timeline._fields_ == C.Timeline._fields_


def test_entry_record_sizes():
    """Record lengths of the gradient pass's per-entry face records: the
    FP64 records carry c and r_ij, the mixed ones the combined weights."""
    L = C.lib()
    f64, mixed = C.SFX_PREC_FP64, C.SFX_PREC_MIXED
    assert [L.sfx_mesh_entry_record(nd, f64) for nd in (2, 3)] == [8, 12]
    assert [L.sfx_mesh_entry_record(nd, mixed) for nd in (2, 3)] == [8, 8]
    assert L.sfx_mesh_entry_record(1, f64) == -1
    assert L.sfx_mesh_entry_record(3, 7) == -1


def test_face_record_planes():
    """float4 planes per face record of the mixed face pass: 3 in 2D, 4 in
    3D (13 and 17 values); no FP64 records."""
    L = C.lib()
    assert [L.sfx_mesh_face_record_planes(nd, C.SFX_PREC_MIXED) for nd in (2, 3)] == [3, 4]
    assert L.sfx_mesh_face_record_planes(3, C.SFX_PREC_FP64) == -1
    assert L.sfx_mesh_face_record_planes(4, C.SFX_PREC_MIXED) == -1
