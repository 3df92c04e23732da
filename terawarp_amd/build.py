"""Build libterawarp.so (HIP kernels for gfx950 + host C++) in-tree.

    python -m terawarp_amd.build          # or terawarp_amd.build.build()

Every .hip file is compiled by hipcc for gfx950 with FP contraction off (the
reference-order FP64 paths depend on it; the fast paths call fma
explicitly) and with packed FP32 instructions disabled.  Host-only C++ is
compiled by g++.  Objects go to terawarp_amd/_build/, the library and the
command-line tool to terawarp_amd/.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libterawarp.so")
CLI = os.path.join(HERE, "terawarp")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-I" + INCLUDE,
]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wextra", "-I" + INCLUDE,
             "-I" + os.path.join(ROCM, "include"), "-D__HIP_PLATFORM_AMD__"]


def _sources():
    hip, cxx = [], []
    for name in sorted(os.listdir(CSRC)):
        path = os.path.join(CSRC, name)
        if name.endswith(".hip"):
            hip.append(path)
        elif name.endswith(".cpp") and not name.startswith("cli_"):
            cxx.append(path)
    return hip, cxx


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r.stderr


def build(verbose=False, jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    hip, cxx = _sources()
    cmds = []
    objs = []
    for src in hip:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj])
        objs.append(obj)
    for src in cxx:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        cmds.append(["g++"] + CXX_FLAGS + ["-c", src, "-o", obj])
        objs.append(obj)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for err in ex.map(_run, cmds):
            if verbose and err.strip():
                sys.stderr.write(err)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lgomp"])
    os.replace(tmp, LIB)
    cli_src = os.path.join(CSRC, "cli_terawarp.cpp")
    if os.path.exists(cli_src):
        _run(["g++"] + CXX_FLAGS + [cli_src, "-o", CLI, "-L" + HERE, "-lterawarp",
                                    "-Wl,-rpath,$ORIGIN", "-L" + os.path.join(ROCM, "lib"), "-lamdhip64"])
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
