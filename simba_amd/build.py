"""Builds the gfx950 library and command-line tools in-tree.

  simba_amd/libsimba_amd.so   kernels + host drivers + the C ABI
  simba_amd/bin/simba_exe     reference-compatible SIMBA CLI
  simba_amd/bin/vfb_exe       reference-compatible VFB CLI

Every HIP source is compiled with hipcc for gfx950 only, with packed FP32
instructions disabled.  Run: python -m simba_amd.build
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsimba_amd.so")
BIN = os.path.join(HERE, "bin")

ARCH = "gfx950"
COMMON = ["-O3", "-std=c++20", "-fPIC", "-Wall", "-Wno-unused-result",
          f"-I{ROOT}/include", f"-I{CSRC}"]
HIP_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops"]

LIB_SOURCES = [
    "simba_kernels.hip", "vfb_kernels.hip", "mlp_kernels.hip", "noalloc.hip",
    "simba_host.cpp", "vfb_host.cpp", "batch_host.cpp", "mlp_host.cpp", "api.cpp", "expr.cpp",
    "instance_io.cpp",
]
# CLI binaries: (source, extra defines).  simba_rtid_exe is simba_exe with
# the SIMBA-RTid order (the reference's nocache_rtid.cu build).
CLI_SOURCES = {"simba_exe": ("cli_simba.cpp", []),
               "simba_rtid_exe": ("cli_simba.cpp", ["-DMBA_CLI_RTID=1"]),
               "vfb_exe": ("cli_vfb.cpp", [])}


def hipcc():
    for p in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if p and os.path.exists(p):
            return p
    raise RuntimeError("hipcc not found")


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    cmd = [hipcc()] + COMMON + HIP_FLAGS
    if src.endswith(".hip"):
        cmd += ["-x", "hip"]
    cmd += ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stderr}")
    return obj


def build(jobs=None, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    srcs = [s for s in LIB_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stderr}")
    os.replace(tmp, LIB)
    for exe, (src, defs) in CLI_SOURCES.items():
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        out = os.path.join(BIN, exe)
        cmd = [hipcc()] + COMMON + defs + ["-o", out, path, f"-L{HERE}", "-lsimba_amd",
                                           "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"CLI build failed: {exe}\n{r.stderr}")
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
