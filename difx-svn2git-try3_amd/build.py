"""Build recipe for libfxcorr_gpu.so (HIP, gfx950) and the C++ host shim.

Every HIP translation unit is compiled for gfx950 only, with packed FP32
instructions disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
OUT = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-result",
    "-Wno-unknown-pragmas",
    f"-I{os.path.join(REPO, 'include')}",
]

SOURCES = ["fx_api.cpp", "fx_plan.cpp", "fx_station.hip", "fx_xmac.hip", "fx_xmac_mfma.hip", "fx_pcal.hip", "fx_mux.cpp",
           "fx_mux.hip", "fx_kurtosis.hip"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def _check_no_scratch(r, src):
    """Kernels must not use scratch (private) memory: a spilled struct or a
    dynamically indexed local array costs a memory round trip per access."""
    fn, bad = None, []
    for line in r.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            fn = m.group(1)
        m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", line)
        if m and int(m.group(1)) > 0:
            bad.append(f"{fn}: {m.group(1)} bytes/lane")
    if bad:
        raise RuntimeError(f"{src}: kernels using scratch memory:\n  " + "\n  ".join(bad))


def build_gpu_lib(verbose: bool = False) -> str:
    os.makedirs(OUT, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(OUT, src.replace(".", "_") + ".o")
        lang = ["-x", "hip"]
        remarks = ["-Rpass-analysis=kernel-resource-usage"] if src.endswith(".hip") else []
        cmds.append(([HIPCC, *HIP_FLAGS, *remarks, *lang, "-c", path, "-o", obj], src))
        objs.append(obj)

    def compile_one(job):
        cmd, src = job
        _check_no_scratch(_run(cmd), src)

    with cf.ThreadPoolExecutor(max_workers=min(8, len(cmds))) as ex:
        list(ex.map(compile_one, cmds))
    lib = os.path.join(OUT, "libfxcorr_gpu.so")
    tmp = lib + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, lib)
    if verbose:
        print("built", lib)
    return lib


def build_host_shim(verbose: bool = False) -> str:
    """C++ Mode/Core mirror classes over the C ABI (libfxcorr_host.so)."""
    os.makedirs(OUT, exist_ok=True)
    srcs = [os.path.join(HOST, f) for f in sorted(os.listdir(HOST)) if f.endswith(".cpp")] if os.path.isdir(HOST) else []
    if not srcs:
        return ""
    lib = os.path.join(OUT, "libfxcorr_host.so")
    tmp = lib + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall",
          f"-I{os.path.join(REPO, 'include')}", f"-I{HOST}", "-o", tmp, *srcs,
          f"-L{OUT}", "-lfxcorr_gpu", f"-Wl,-rpath,{OUT}"])
    os.replace(tmp, lib)
    if verbose:
        print("built", lib)
    return lib


def build_host_tests(verbose: bool = False) -> list:
    """C++ programs that drive the host shim (tests/cpp/*.cpp)."""
    tdir = os.path.join(REPO, "tests", "cpp")
    outs = []
    if not os.path.isdir(tdir):
        return outs
    for f in sorted(os.listdir(tdir)):
        if not f.endswith(".cpp"):
            continue
        exe = os.path.join(OUT, f[:-4])
        _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{os.path.join(REPO, 'include')}", f"-I{HOST}",
              "-o", exe, os.path.join(tdir, f), f"-L{OUT}", "-lfxcorr_host", "-lfxcorr_gpu",
              f"-Wl,-rpath,{OUT}"])
        outs.append(exe)
        if verbose:
            print("built", exe)
    return outs


def build_oracle(verbose: bool = False) -> str:
    """The CPU oracle (test infrastructure), via its own Makefile."""
    r = subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "-B"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    return os.path.join(REPO, "oracle", "_build", "libfxoracle.so")


def build_all(verbose: bool = False):
    build_gpu_lib(verbose)
    build_host_shim(verbose)
    build_host_tests(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
    sys.exit(0)
