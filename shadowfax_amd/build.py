"""Build the HIP library libshadowfax_amd.so in-tree for gfx950.

    python -m shadowfax_amd.build        (or shadowfax_amd.build.build())

Per-file flags:
  * FP64 parity kernels: -ffp-contract=off, so expressions keep the
    association of the reference C++ (no FMA contraction);
  * mixed kernels: contraction within expressions (-ffp-contract=on, which
    honours the per-function contract(off) of the shared FP64 code), fast
    FP32 divide/sqrt;
  * all device code: packed FP32 instructions disabled
    (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libshadowfax_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-std=c++17", "-O3", "-fPIC", "--offload-arch=" + ARCH,
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I" + CSRC, "-I" + os.path.join(ROOT, "include"),
          "-Wno-unused-command-line-argument"]

SOURCES = [
    ("riemann_fp64.hip", ["-ffp-contract=off"]),
    ("riemann_mixed.hip", ["-ffp-contract=on",
                           "-fno-hip-fp32-correctly-rounded-divide-sqrt"]),
    ("flux_pass.hip", ["-ffp-contract=off"]),
    ("cell_ops.hip", ["-ffp-contract=off"]),
    ("flux_pass_mixed.hip", ["-ffp-contract=on",
                             "-fno-hip-fp32-correctly-rounded-divide-sqrt"]),
    ("sfx_capi.cpp", []),
    ("sfx_mesh_capi.cpp", ["-ffp-contract=off"]),
    ("sfx_static.hip", []),
    ("mesh_voronoi.cpp", ["-ffp-contract=off"]),
    ("mesh_voronoi3d.cpp", ["-ffp-contract=off"]),
    ("ensemble_fp64.hip", ["-ffp-contract=off"]),
    ("ensemble_mixed.hip", ["-ffp-contract=on", "-fno-hip-fp32-correctly-rounded-divide-sqrt"]),
    ("sfx_ensemble.cpp", []),
    ("snapshot_hdf5.cpp", []),
]


def build(verbose=False, jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    procs = []
    for src, flags in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC] + COMMON + flags + ["-c", path, "-o", obj]
        if src.endswith(".cpp"):
            cmd = [HIPCC] + COMMON + flags + ["-x", "hip", "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT)))
        objs.append(obj)
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append(src)
            sys.stderr.write(out.decode(errors="replace"))
        elif verbose and out:
            sys.stderr.write(out.decode(errors="replace"))
    if failed:
        raise RuntimeError("hipcc failed for: " + ", ".join(failed))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + \
        objs + ["-Wl,-soname,libshadowfax_amd.so"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    build_probe(verbose)
    return LIB


PROBE = os.path.join(HERE, "libsfxprobe.so")


def build_probe(verbose=False):
    """libsfxprobe.so: the copy and FMA-loop kernels bench.py measures the
    GPU's ceilings with (not part of the face-flux library)."""
    tmp = PROBE + ".tmp"
    cmd = [HIPCC] + COMMON + ["-shared", "-o", tmp, os.path.join(CSRC, "probe.hip")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, PROBE)
    return PROBE


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
