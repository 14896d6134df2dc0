"""Builds libfnet_hip.so (HIP, gfx950) in-tree with hipcc.

``python -m fortnet_amd.build`` compiles every ``csrc/*.hip`` translation unit in parallel and
links the shared library next to this file, where ``fortnet_amd._lib`` loads it; it also links
``libfnet_probe.so`` (``probe/``: a plain copy kernel that bench.py uses to scale the HBM byte
counters when it runs under a profiler) and, with amdflang, the Fortran host layer
(``fortran/``: the ISO_C_BINDING module, the host types and the ``fnet_fortran_check`` driver).
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfnet_hip.so")
PROBE_LIB = os.path.join(HERE, "libfnet_probe.so")  # copy kernel for counter calibration
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("FNET_FC", "/opt/rocm/bin/amdflang")
FORTRAN = os.path.join(HERE, "fortran")
FCHECK = os.path.join(FORTRAN, "fnet_fortran_check")  # Fortran host driver (executable)
# HDF5 I/O layer: plain C over the HDF5 C and high-level libraries of the image (HDF5 1.10)
CC = os.environ.get("FNET_CC", "gcc")
HDF5_PREFIX = os.environ.get("FNET_HDF5_PREFIX", "/opt/conda")
H5_LIB = os.path.join(HERE, "libfnet_h5.so")
ARCH = os.environ.get("FNET_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions are disabled for every device compilation.
FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-Wall",
         "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I" + os.path.join(HERE, "..", "include")]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, out))
    return obj, out


def _link(objs, lib):
    tmp = lib + ".tmp%d" % os.getpid()
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout)
    os.replace(tmp, lib)


def build_fortran():
    """Compiles the Fortran host layer and links the driver against libfnet_hip.so (found next
    to it through the run path, so the tree can move).  Returns the executable, or None without
    a Fortran compiler."""
    if not os.path.exists(FC):
        return None
    mod = os.path.join(OBJ, "fortran")
    os.makedirs(mod, exist_ok=True)
    objs = []
    for name in ("fnet_hip_c", "fnet_host", "fnet_fortran_check"):
        obj = os.path.join(mod, name + ".o")
        cmd = [FC, "-O2", "-module-dir", mod, "-I", mod, "-c",
               os.path.join(FORTRAN, name + ".f90"), "-o", obj]
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode != 0:
            raise RuntimeError("amdflang failed for %s:\n%s" % (name, r.stdout))
        objs.append(obj)
    tmp = FCHECK + ".tmp%d" % os.getpid()
    cmd = [FC] + objs + ["-L" + HERE, "-lfnet_hip", "-Wl,-rpath,$ORIGIN/..", "-o", tmp]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("Fortran link failed:\n" + r.stdout)
    os.replace(tmp, FCHECK)
    return FCHECK


def build_h5():
    """Builds libfnet_h5.so (csrc/fnet_h5.c) against the HDF5 libraries under HDF5_PREFIX, found
    at run time through the run path.  Returns the library, or None without HDF5."""
    inc = os.path.join(HDF5_PREFIX, "include")
    libdir = os.path.join(HDF5_PREFIX, "lib")
    if not (os.path.exists(os.path.join(inc, "hdf5_hl.h")) and
            os.path.exists(os.path.join(libdir, "libhdf5_hl.so"))):
        return None
    tmp = H5_LIB + ".tmp%d" % os.getpid()
    cmd = [CC, "-O2", "-std=c99", "-Wall", "-shared", "-fPIC",
           "-I" + os.path.join(HERE, "..", "include"), "-I" + inc,
           os.path.join(CSRC, "fnet_h5.c"), "-L" + libdir, "-lhdf5_hl", "-lhdf5",
           "-Wl,-rpath," + libdir, "-o", tmp]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("HDF5 layer build failed:\n" + r.stdout)
    os.replace(tmp, H5_LIB)
    return H5_LIB


GEN_LIB = os.path.join(HERE, "libfnet_gen.so")  # host helper of the workload generator


def build_gen():
    """Builds libfnet_gen.so (csrc/fnet_gen.c): the random sequential addition of the synthetic
    cells, with FMA contraction off so that it matches the numpy restatement bit for bit."""
    tmp = GEN_LIB + ".tmp%d" % os.getpid()
    cmd = [CC, "-O2", "-std=c99", "-Wall", "-shared", "-fPIC", "-ffp-contract=off",
           os.path.join(CSRC, "fnet_gen.c"), "-lm", "-o", tmp]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("generator helper build failed:\n" + r.stdout)
    os.replace(tmp, GEN_LIB)
    return GEN_LIB


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    probe = os.path.join(HERE, "probe", "fnet_copy_probe.hip")
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs) + 1)) as ex:
        results = list(ex.map(_compile, srcs + [probe]))
    if verbose:
        for _, out in results:
            if out.strip():
                print(out)
    _link([o for o, _ in results[:-1]], LIB)
    _link([results[-1][0]], PROBE_LIB)
    build_fortran()
    build_h5()
    build_gen()
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
