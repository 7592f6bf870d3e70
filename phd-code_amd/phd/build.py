"""In-tree build of the package's native libraries.

libphd_flux.so : hipcc --offload-arch=gfx950 (HIP kernels + C ABI)
libphd_mesh.so : g++ -fopenmp (host Voronoi builder)

Both land next to this file, where phd._native loads them.  Run as
`python -m phd.build` from phd-code_amd/, or via __graft_entry__.build().
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
INCLUDE = os.path.join(REPO, "include")
CSRC = os.path.join(HERE, "csrc")

HIP_SOURCES = ["phd_flux.hip", "adjoint.hip", "domain.hip", "noalloc.hip", "geometry.hip"]
# -ffp-contract=off: no FMA contraction, so BITWISE kernels perform the same
# IEEE operations as the reference; packed-fp32 instructions are disabled.
HIPCC_FLAGS = ["-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-shared",
               "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang",
               "-packed-fp32-ops", "-I" + INCLUDE]
GXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-I" + INCLUDE]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.exists(c) or c == "hipcc"):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out, flush=True)
    if p.returncode != 0:
        raise RuntimeError("build failed: %s" % " ".join(cmd))


def build_flux():
    out = os.path.join(HERE, "libphd_flux.so")
    tmp = out + ".tmp"
    _run([hipcc()] + HIPCC_FLAGS + [os.path.join(CSRC, s) for s in HIP_SOURCES] + ["-o", tmp])
    os.replace(tmp, out)
    return out


def build_mesh():
    out = os.path.join(HERE, "libphd_mesh.so")
    tmp = out + ".tmp"
    _run(["g++"] + GXX_FLAGS + [os.path.join(CSRC, "voronoi.cpp"), "-o", tmp])
    os.replace(tmp, out)
    return out


def build_all():
    return [build_mesh(), build_flux()]


if __name__ == "__main__":
    build_all()
    sys.exit(0)
