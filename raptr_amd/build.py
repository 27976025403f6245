"""Build recipe of the gfx950 HIP library (``raptr_amd/lib/libraptr_hip.so``).

Each ``csrc/*.hip`` translation unit is compiled with hipcc for gfx950 and
the objects are linked into one shared library exposing the C ABI declared
in ``include/raptr_hip.h``.  Numerics-relevant flags:

* ``-ffp-contract=off``: no FMA contraction, so the kernels evaluate the
  reference's expressions in its operation order (bit parity with the host
  restatement);
* HIP's default correctly rounded FP32 divide and square root (never
  ``-ffast-math``);
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 VALU
  instructions.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libraptr_hip.so")
ARCH = os.environ.get("RAPTR_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC] + CFLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    # hipcc's host pass warns about the device-only target feature; keep the
    # rest of the diagnostics visible
    msg = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l
                    and "warning generated" not in l)
    return obj, msg.strip()


def build(verbose=False, jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_compile, srcs))
    for obj, msg in results:
        if msg and verbose:
            print(msg, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + \
          [o for o, _ in results]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if build_h5():
        build_cpp_io_test()
    build_cpp_tests()
    return LIB


HDF5_PREFIX = os.environ.get("RAPTR_HDF5_PREFIX", "/opt/conda")
H5_LIB = os.path.join(LIBDIR, "libraptr_h5.so")


def build_h5():
    """Host library for the HDF5 measurement/density files
    (csrc/h5io.cpp against the image's HDF5 C library); skipped when HDF5
    is absent, in which case raptr_amd.h5io raises on use."""
    inc = os.path.join(HDF5_PREFIX, "include")
    lib = os.path.join(HDF5_PREFIX, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        return None
    os.makedirs(LIBDIR, exist_ok=True)
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-static-libstdc++",
           "-static-libgcc", f"-I{os.path.join(ROOT, 'include')}", f"-I{inc}",
           os.path.join(CSRC, "h5io.cpp"), "-o", H5_LIB + ".tmp", f"-L{lib}", "-lhdf5",
           f"-Wl,-rpath,{lib}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for h5io.cpp:\n{r.stdout}\n{r.stderr}")
    os.replace(H5_LIB + ".tmp", H5_LIB)
    return H5_LIB


def build_cpp_io_test():
    """C++ client of include/raptr/io.hpp (host only, links libraptr_h5)."""
    src = os.path.join(ROOT, "tests", "cpp", "test_io.cpp")
    out = os.path.join(LIBDIR, "test_io")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", f"-I{os.path.join(ROOT, 'include')}", src,
           "-o", out + ".tmp", f"-L{LIBDIR}", "-lraptr_h5", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for {src}:\n{r.stdout}\n{r.stderr}")
    os.replace(out + ".tmp", out)


CPP_TESTS = [os.path.join(ROOT, "tests", "cpp", "test_reference_api.cpp")]


def build_cpp_tests():
    """C++ programs written against include/raptr/*.hpp, linked to the HIP
    library (rpath to raptr_amd/lib)."""
    for src in CPP_TESTS:
        out = os.path.join(LIBDIR, os.path.basename(src)[:-4])
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", f"-I{os.path.join(ROOT, 'include')}", src,
               "-o", out + ".tmp", f"-L{LIBDIR}", "-lraptr_hip", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"g++ failed for {src}:\n{r.stdout}\n{r.stderr}")
        os.replace(out + ".tmp", out)


if __name__ == "__main__":
    print(build(verbose=True))
