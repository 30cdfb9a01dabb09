"""Build recipes for the native parts of the package.

All artefacts are written in-tree under ``pyblastafterglowmag_amd/_lib`` so
that they travel with the source tree.  The GPU library is cross-compiled for
gfx950 with hipcc; the HDF5 helper is plain C against the image's HDF5 1.10
(``/opt/conda``), exposed through a private folder of symlinks so that only the
HDF5 shared objects (and not conda's libstdc++) are found at run time.
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(PKG, "_lib")
CSRC = os.path.join(PKG, "csrc")
ROOT = os.path.dirname(PKG)
CONDA = "/opt/conda"
HDF5_LINKS = os.path.join(LIB, "hdf5")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# packed-FP32 instructions are disabled for every HIP translation unit
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]


def _run(cmd, **kw):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd, **kw)


def hdf5_links():
    """Private folder exposing only HDF5 and its compression/transport deps."""
    os.makedirs(HDF5_LINKS, exist_ok=True)
    pats = ("libhdf5", "libsz.so", "libaec.so", "libz.so", "libcurl.so", "libssl.so", "libcrypto.so")
    for name in os.listdir(os.path.join(CONDA, "lib")):
        if name.startswith(pats) and ".so" in name:
            dst = os.path.join(HDF5_LINKS, name)
            if not os.path.lexists(dst):
                os.symlink(os.path.join(CONDA, "lib", name), dst)
    return HDF5_LINKS


def build_h5io():
    os.makedirs(LIB, exist_ok=True)
    links = hdf5_links()
    out = os.path.join(LIB, "libpba_h5io.so")
    _run(["gcc", "-O2", "-shared", "-fPIC", os.path.join(CSRC, "h5io.c"),
          "-I" + os.path.join(CONDA, "include"), "-L" + links, "-Wl,-rpath," + links,
          "-lhdf5", "-o", out])
    return out


GPU_SOURCES = ["capi.hip", "dyn.hip", "ele.hip", "eats.hip", "eats_vec.hip", "eats_pw.hip", "skymap.hip", "mlp.hip"]


def build_gpu():
    """libpbagpu.so: every HIP kernel of the hot path plus the C ABI, for gfx950."""
    os.makedirs(LIB, exist_ok=True)
    objs = []
    for src in GPU_SOURCES:
        obj = os.path.join(LIB, src.replace(".hip", ".o"))
        _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    out = os.path.join(LIB, "libpbagpu.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", out])
    return out


def build_cabi_driver():
    """tests/cabi/pba_gpu_run.cpp: a C++ program that calls the C ABI with no Python in the process
    (the GPU test test_gpu_cabi.py runs it)."""
    src = os.path.join(ROOT, "tests", "cabi", "pba_gpu_run.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(LIB, "pba_gpu_run")
    _run([HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), src,
          "-L" + LIB, "-Wl,-rpath,$ORIGIN", "-lpbagpu", "-o", out])
    return out


ORACLE = os.path.join(ROOT, "oracle")


def build_oracle():
    """CPU oracle (test infrastructure): plain g++ build of oracle/*.cpp."""
    out_dir = os.path.join(ORACLE, "_lib")
    os.makedirs(out_dir, exist_ok=True)
    srcs = sorted(os.path.join(ORACLE, f) for f in os.listdir(ORACLE) if f.endswith(".cpp"))
    out = os.path.join(out_dir, "liboracle.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-I" + os.path.join(ROOT, "include")]
         + srcs + ["-o", out])
    return out


def build_reference():
    """Reference CPU program from /root/reference (only where that tree exists)."""
    script = os.path.join(ORACLE, "build_ref.sh")
    if os.path.isdir("/root/reference/src"):
        _run(["bash", script])


def build_all(reference=True):
    build_h5io()
    build_gpu()
    build_cabi_driver()
    build_oracle()
    if reference:
        build_reference()


if __name__ == "__main__":
    build_all(reference="--no-reference" not in sys.argv)
