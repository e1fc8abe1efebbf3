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
