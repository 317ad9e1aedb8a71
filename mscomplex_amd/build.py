"""Build libmscx.so (HIP kernels for gfx950 + C++ host runtime) in-tree.

Usage: python -m mscomplex_amd.build   (or mscomplex_amd.build.build())
"""
import concurrent.futures as _cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmscx.so")
ARCH = os.environ.get("MSCX_OFFLOAD_ARCH", "gfx950")

# Packed-FP32 instructions are disabled for all device code.
HIP_FLAGS = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-command-line-argument",
]
HOST_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-result",
              "-Wno-unused-command-line-argument"]

DEVICE_SOURCES = ["kernels.hip", "noalloc.hip", "dmerge.hip", "devsort.hip"]
HOST_SOURCES = ["runtime.cpp", "tile.cpp", "complex.cpp", "discs.cpp", "pipeline.cpp", "reader.cpp", "server.cpp",
                "gen.cpp"]
# the reference's dmsc command line over the C ABI (links libmscx.so)
CLI_SOURCE = "dmsc_main.cpp"
CLI = os.path.join(HERE, "mscx_dmsc")
DRIVER_SOURCE = "dataset_driver.cpp"
DRIVER = os.path.join(HERE, "mscx_dataset_driver")


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    return "hipcc"


def _host_cxx():
    c = os.environ.get("MSCX_HOST_CXX")
    if c:
        return c
    return shutil.which("g++") or _hipcc()


def _compile(src, device):
    out = os.path.join(OBJ, os.path.basename(src) + ".o")
    if device:
        cmd = [_hipcc(), *HIP_FLAGS, "-c", src, "-o", out]
    else:
        # host code goes through the system C++ compiler: its code for the
        # simplification heap's sift-down is measurably faster than clang's
        cmd = [_host_cxx(), *HOST_FLAGS, "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-c", src, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return out


def build(verbose=False):
    """Compile every source and link mscomplex_amd/libmscx.so (always rebuilds)."""
    os.makedirs(OBJ, exist_ok=True)
    jobs = [(os.path.join(CSRC, s), True) for s in DEVICE_SOURCES]
    jobs += [(os.path.join(CSRC, s), False) for s in HOST_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with _cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        objs = list(ex.map(lambda j: _compile(*j), jobs))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-lpthread",
           "-Wl,-soname,libmscx.so"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    tmp = CLI + ".tmp"
    cmd = [_host_cxx(), "-O2", "-std=c++17", "-Wall", os.path.join(CSRC, CLI_SOURCE), "-o", tmp, LIB,
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, CLI)
    # the GridDataManager order through the C++ drop-in header (an
    # integration example; the GPU tests run it against the oracle)
    tmp = DRIVER + ".tmp"
    cmd = [_host_cxx(), "-O2", "-std=c++17", "-Wall", os.path.join(CSRC, DRIVER_SOURCE), "-o", tmp, LIB,
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, DRIVER)
    if verbose:
        print(f"built {LIB}, {CLI} and {DRIVER}")
    return LIB


SANITIZERS = {
    # host code instrumented; device code is built as usual (GPU sanitizers
    # are not part of this build)
    "address": ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"],
    "thread": ["-fsanitize=thread"],
}


def build_sanitized(kind, verbose=False):
    """A CPU-sanitizer build of the library in mscomplex_amd/_san_<kind>/
    (host sources through g++ with AddressSanitizer + UndefinedBehavior-
    Sanitizer, or ThreadSanitizer).  Load it with MSCX_LIB_PATH set to the
    returned path and the sanitizer runtime preloaded, e.g.

      MSCX_LIB_PATH=$(python -m mscomplex_amd.build --sanitize address) \
      LD_PRELOAD=$(gcc -print-file-name=libasan.so):$(gcc -print-file-name=libubsan.so) \
      ASAN_OPTIONS=detect_leaks=0 python -m pytest tests/test_complex_cpu.py
    """
    flags = SANITIZERS[kind]
    out_dir = os.path.join(HERE, f"_san_{kind}")
    os.makedirs(out_dir, exist_ok=True)

    def comp(src, device):
        out = os.path.join(out_dir, os.path.basename(src) + ".o")
        if device:
            cmd = [_hipcc(), *HIP_FLAGS, "-c", src, "-o", out]
        else:
            cmd = ["g++", "-O1", "-g", "-fno-omit-frame-pointer", "-fPIC", "-std=c++17", "-Wall",
                   "-Wno-unused-result", *flags, "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-c", src, "-o", out]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        return out

    jobs = [(os.path.join(CSRC, x), True) for x in DEVICE_SOURCES]
    jobs += [(os.path.join(CSRC, x), False) for x in HOST_SOURCES]
    with _cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        objs = list(ex.map(lambda j: comp(*j), jobs))
    lib = os.path.join(out_dir, "libmscx.so")
    cmd = ["g++", "-shared", *flags, "-o", lib + ".tmp", *objs, "-L/opt/rocm/lib", "-lamdhip64", "-lpthread",
           "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(lib + ".tmp", lib)
    if verbose:
        print(lib)
    return lib


if __name__ == "__main__":
    if len(sys.argv) == 3 and sys.argv[1] == "--sanitize":
        build_sanitized(sys.argv[2], verbose=True)
    else:
        build(verbose=True)
    sys.exit(0)
